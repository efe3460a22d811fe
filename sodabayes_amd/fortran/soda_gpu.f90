! Fortran binding of libsodabayes (include/sodabayes.h) through iso_c_binding.
!
! A host program that keeps the reference's option parsing and output
! writers uses this module in place of the SODA branch's two calls
! (SODABayes.f90:139 setup_soda, :142 soda_bayes): sb_create + sb_build_aug,
! then sb_mcmc_init + sb_mcmc_run + sb_get.  Arrays cross as plain pointers;
! SNP ids and block positions are 0-based on the C side.
module soda_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: SB_OK = 0, SB_ERR_ARG = -1, SB_ERR_HIP = -2, &
        SB_ERR_NOMEM = -3, SB_ERR_STATE = -4, SB_ERR_NOT_PD = -5
    integer(c_int), parameter :: SB_RNG_REPLAY = 0, SB_RNG_COUNTER = 1
    integer(c_int), parameter :: SB_PREC_FP64 = 0, SB_PREC_MIXED = 1
    integer(c_int), parameter :: SB_Q_G = 1, SB_Q_CLASS = 2, SB_Q_VARA_S = 3, &
        SB_Q_YADJ = 4, SB_Q_YAADJ = 5, SB_Q_P = 6, SB_Q_SCALARS = 7, SB_Q_GSTORE = 8, &
        SB_Q_INDISTSTORE = 9, SB_Q_VARA_SSTORE = 10, SB_Q_STORES = 11, &
        SB_Q_SNPSTORE = 12, SB_Q_VARSTORE = 13, SB_Q_PSTORE = 14, SB_Q_XA = 15, &
        SB_Q_XXD = 16, SB_Q_GRAM = 17, SB_Q_SNPINDIST = 18, SB_Q_VARINDIST = 19, &
        SB_Q_KTIMES = 20, SB_Q_SKIPPED = 21, SB_Q_NAIND = 22, SB_Q_WHY = 23

    type, bind(C) :: sb_dims
        integer(c_int32_t) :: method, nind, nloci, nmix, ngroup
        integer(c_int32_t) :: nblocks, blocksize, blockt
        integer(c_int32_t) :: nchains, rng_mode, precision, bed_on_device
    end type sb_dims

    type, bind(C) :: sb_priors
        real(c_double) :: pizero, scale_va, scale_ve, df_va, df_ve, vara0
        real(c_double) :: mix(20), alpha(20)
        integer(c_int32_t) :: numit, burnin, thin, resync
        integer(c_int64_t) :: seed
    end type sb_priors

    interface
        integer(c_int) function sb_rng_seed(seed, stream, state) bind(C, name='sb_rng_seed')
            import :: c_int, c_int32_t, c_int64_t
            integer(c_int64_t), value :: seed
            integer(c_int32_t), value :: stream
            integer(c_int64_t) :: state(2)
        end function
        integer(c_int) function sb_permutate(n, s, state, permvec) bind(C, name='sb_permutate')
            import :: c_int, c_int32_t, c_int64_t
            integer(c_int32_t), value :: n, s
            integer(c_int64_t) :: state(2)
            integer(c_int32_t) :: permvec(*)
        end function
        integer(c_int) function sb_create(dims, bed, freq, permvec, snpgrp, h) bind(C, name='sb_create')
            import :: c_int, c_int32_t, c_double, c_ptr, sb_dims
            type(sb_dims), intent(in) :: dims
            type(c_ptr), value :: bed
            real(c_double), intent(in) :: freq(*)
            integer(c_int32_t), intent(in) :: permvec(*)
            type(c_ptr), value :: snpgrp
            type(c_ptr) :: h
        end function
        integer(c_int) function sb_destroy(h) bind(C, name='sb_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function sb_rng_set(h, state) bind(C, name='sb_rng_set')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: state(2)
        end function
        integer(c_int) function sb_rng_get(h, state) bind(C, name='sb_rng_get')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int64_t) :: state(2)
        end function
        integer(c_int) function sb_build_aug(h, xxd, fail_block, fail_col) bind(C, name='sb_build_aug')
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double) :: xxd(*)
            integer(c_int32_t) :: fail_block, fail_col
        end function
        integer(c_int) function sb_mcmc_init(h, pr, why) bind(C, name='sb_mcmc_init')
            import :: c_int, c_double, c_ptr, sb_priors
            type(c_ptr), value :: h
            type(sb_priors), intent(in) :: pr
            real(c_double), intent(in) :: why(*)
        end function
        integer(c_int) function sb_mcmc_init_chains(h, pr, why, ldy, heldout, ldh) &
                bind(C, name='sb_mcmc_init_chains')
            import :: c_int, c_int64_t, c_double, c_ptr, sb_priors
            type(c_ptr), value :: h
            type(sb_priors), intent(in) :: pr
            real(c_double), intent(in) :: why(*)
            integer(c_int64_t), value :: ldy, ldh
            type(c_ptr), value :: heldout
        end function
        integer(c_int) function sb_mcmc_run(h, first, count, hyp) bind(C, name='sb_mcmc_run')
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: first, count
            real(c_double) :: hyp(*)
        end function
        integer(c_int32_t) function sb_hyp_width(h) bind(C, name='sb_hyp_width')
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function sb_get(h, what, chain, out, count) bind(C, name='sb_get')
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: what, chain
            real(c_double) :: out(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function sb_get_info(h, out, n) bind(C, name='sb_get_info')
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: h
            real(c_double) :: out(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function sb_predict(nind, nloci, bed, bed_on_device, freq, effects, mu, pred) &
                bind(C, name='sb_predict')
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value :: nind, nloci, bed_on_device
            type(c_ptr), value :: bed
            real(c_double), intent(in) :: freq(*), effects(*)
            real(c_double), value :: mu
            real(c_double) :: pred(*)
        end function
        integer(c_int) function sb_state_size(h, bytes) bind(C, name='sb_state_size')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int64_t) :: bytes
        end function
        integer(c_int) function sb_state_save(h, buf, size) bind(C, name='sb_state_save')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h, buf
            integer(c_int64_t), value :: size
        end function
        integer(c_int) function sb_state_load(h, buf, size) bind(C, name='sb_state_load')
            import :: c_int, c_int64_t, c_ptr
            type(c_ptr), value :: h, buf
            integer(c_int64_t), value :: size
        end function
        type(c_ptr) function sb_last_error() bind(C, name='sb_last_error')
            import :: c_ptr
        end function
    end interface

contains

    ! Message of the last library failure as a Fortran string.
    function sb_error_message() result(msg)
        character(len=256) :: msg
        character(kind=c_char), pointer :: cs(:)
        integer :: i
        msg = ''
        call c_f_pointer(sb_last_error(), cs, [256])
        do i = 1, 256
            if (cs(i) == c_null_char) exit
            msg(i:i) = cs(i)
        end do
    end function sb_error_message

end module soda_gpu
