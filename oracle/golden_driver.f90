! Golden-vector driver for the SODA oracle (test infrastructure only).
!
! Links the reference's own Fortran modules (compiled from /root/reference/src
! by oracle/Makefile into oracle/_ref/) and dumps full-precision values that
! the C restatement in oracle/soda_oracle.c and the GPU path are checked
! against.  Nothing here is part of the product.
!
! Selected by environment variables (the reference's own command-line parser
! still reads the normal options, e.g. -bfile -method -seed):
!   GOLDEN_MODE = rng   : stream-0 variates from parallelrng.f90 for -seed
!               = soda  : run_program_sodaBayes' MCMC sequence
!                         (SODABayes.f90:123-142), then soda_bayes, dumping the
!                         setup products and the final chain state
!               = pd | psd | power : one routine of soda.f90 on the matrix in
!                         GOLDEN_IN (int64 n, then n*n float64, column-major)
!   GOLDEN_OUT  = path of the tagged binary dump
!
! Dump record: character(16) name, int64 kind (1=float64, 2=int64),
! int64 count, then count values.
program golden_driver
    use globals
    use ParallelRNG
    use utilities
    use useroptions
    use soda
    use models
    use omp_lib
    implicit none
    character(len=32) :: mode
    character(len=512) :: outpath, inpath
    integer :: ios

    call get_environment_variable('GOLDEN_MODE', mode)
    call get_environment_variable('GOLDEN_OUT', outpath)
    call get_environment_variable('GOLDEN_IN', inpath)
    open(unit=77, file=trim(outpath), access='stream', form='unformatted', &
         status='replace', iostat=ios)
    if (ios /= 0) stop 'golden_driver: cannot open GOLDEN_OUT'

    select case (trim(mode))
    case ('rng')
        call run_rng()
    case ('soda')
        call run_soda()
    case ('pd', 'psd', 'power')
        call run_matrix(trim(mode))
    case default
        stop 'golden_driver: unknown GOLDEN_MODE'
    end select
    close(77)

contains

    subroutine put_d(name, a, n)
        character(len=*), intent(in) :: name
        integer, intent(in) :: n
        real(real64), intent(in) :: a(n)
        character(len=16) :: nm
        nm = name
        write(77) nm, 1_int64, int(n, int64), a
    end subroutine put_d

    subroutine put_s(name, x)
        character(len=*), intent(in) :: name
        real(real64), intent(in) :: x
        real(real64) :: a(1)
        a(1) = x
        call put_d(name, a, 1)
    end subroutine put_s

    subroutine put_i(name, a, n)
        character(len=*), intent(in) :: name
        integer, intent(in) :: n
        integer, intent(in) :: a(n)
        character(len=16) :: nm
        integer(int64) :: b(n)
        nm = name
        b = int(a, int64)
        write(77) nm, 2_int64, int(n, int64), b
    end subroutine put_i

    subroutine put_l(name, a, n)
        character(len=*), intent(in) :: name
        integer, intent(in) :: n
        logical, intent(in) :: a(n)
        integer :: b(n)
        b = 0
        where (a) b = 1
        call put_i(name, b, n)
    end subroutine put_l

    subroutine put_n(name, x)
        character(len=*), intent(in) :: name
        integer, intent(in) :: x
        integer :: a(1)
        a(1) = x
        call put_i(name, a, 1)
    end subroutine put_n

    ! Catalogue of stream-0 draws (parallelrng.f90:163-455) and the first
    ! uniforms of streams 1..3 (jump-separated, parallelrng.f90:73-77).
    subroutine run_rng()
        character(len=64) :: sval
        integer(int64) :: seed(2), s1
        real(real64) :: u(64), x(16), dir(4), a4(4)
        real(real64) :: shapes(5)
        integer :: i, k
        call get_environment_variable('GOLDEN_SEED', sval)
        read(sval, *) s1
        seed(1) = abs(s1)
        seed(2) = seed(1) + 1
        call rand_set_seed(4, seed)
        call random(64, u, 0)
        call put_d('uniform', u, 64)
        do i = 1, 16
            x(i) = rand_normal(0.3d0, 1.7d0, 0)
        end do
        call put_d('normal', x, 16)
        ! sd <= 0 returns the mean and consumes nothing
        x(1) = rand_normal(2.5d0, 0.0d0, 0)
        call random(x(2), 0)
        call put_d('normal_sd0', x, 2)
        shapes = (/ 0.505d0, 1.0d0, 2.5d0, 7.3d0, 10005.0d0 /)
        do k = 1, 5
            do i = 1, 16
                x(i) = rand_gamma(shapes(k), 1.7d0, 0)
            end do
            call put_d('gamma'//achar(48 + k), x, 16)
        end do
        do i = 1, 16
            x(i) = rand_scaled_inverse_chi_square(4.01d0, 0.3d0, 0)
        end do
        call put_d('sinvchi2', x, 16)
        do i = 1, 16
            x(i) = rand_scaled_inverse_chi_square(2000.01d0, 77.0d0, 0)
        end do
        call put_d('sinvchi2_big', x, 16)
        a4 = (/ 11.0d0, 3.0d0, 1.0d0, 1.0d0 /)
        do i = 1, 4
            dir = rand_dirichlet(4, a4, 0)
            call put_d('dirichlet'//achar(48 + i), dir, 4)
        end do
        do k = 1, 3
            call random(8, u(1:8), k)
            call put_d('stream'//achar(48 + k), u(1:8), 8)
        end do
        call random(8, u(1:8), 0)
        call put_d('uniform_tail', u(1:8), 8)
    end subroutine run_rng

    ! run_program_sodaBayes (SODABayes.f90:118-143) with dumps around the
    ! two calls the GPU path replaces: setup_soda and soda_bayes.
    subroutine run_soda()
        integer :: k, npar, ios2, dblock
        character(len=1024) :: parameters(100)
        character(len=512) :: optpath
        character(len=64) :: sblock, sshift
        real(real64) :: dshift
        ! Options come one token per line from GOLDEN_OPTS and go through the
        ! reference's .par option parser, as a .par job does
        ! (SODABayes.f90:68-70): on the command line a 3-character key such
        ! as -pi is mistaken for a .par file name (useroptions.f90:171).
        call get_environment_variable('GOLDEN_OPTS', optpath)
        npar = 0
        parameters = ''
        open(unit=79, file=trim(optpath), status='old', form='formatted')
        do
            read(79, '(a)', iostat=ios2) parameters(npar + 1)
            if (ios2 /= 0) exit
            npar = npar + 1
        end do
        close(79)
        call get_cmdline
        call get_default_options
        call parse_parfile_options(npar, parameters)
        call parse_cmdline
        call omp_set_dynamic(.false.)
        maxthreads = omp_get_max_threads()
        call parse
        open(unit=21, file=logfil, status='unknown', form='formatted')
        call get_size
        call allocate_data
        call load_geno
        call xscale
        if (.not. mcmc) stop 'golden_driver: soda mode needs an MCMC job'
        call load_pheno
        call set_random_seed
        if (SODAOFF) stop 'golden_driver: soda mode needs SODA on'
        call setup_soda
        if (SODAOFF) stop 'golden_driver: SODA switched off by setup_soda'
        call get_environment_variable('GOLDEN_DSHIFT_BLOCK', sblock, status=ios2)
        if (ios2 == 0 .and. len_trim(sblock) > 0) then
            read(sblock, *) dblock
            call get_environment_variable('GOLDEN_DSHIFT', sshift)
            read(sshift, *) dshift
            call shifted_aug(dblock + 1, dshift)
        end if
        call put_n('nind', nind)
        call put_n('nloci', nloci)
        call put_n('maxthreads', maxthreads)
        call put_n('nthreads', nthreads)
        call put_n('nblocks', nblocks)
        call put_n('blocksize', blocksize)
        call put_n('blockt', blockt)
        call put_n('naind', naind)
        call put_n('nskipped', nskipped)
        call put_d('why', why, nind)
        call put_d('freq', freqstore, nloci)
        call put_i('permvec', permvec, nloci)
        call put_d('xxd', xxd, nblocks)
        call put_d('XA', reshape(XA, (/ size(XA) /)), size(XA))
        call put_l('skip_j', skip_j, nloci)
        if (nskipped > 0) call put_i('i_skipped', i_skipped, nskipped)
        call write_log
        call soda_bayes
        call put_s('mu_store', mu_store)
        call put_s('vare', vare)
        call put_s('vara', vara)
        call put_s('vare_store', vare_store)
        call put_d('g', g, nloci)
        call put_d('gstore', gstore, nloci)
        call put_d('yadj', yadj, nind)
        call put_d('yAadj', yAadj, size(yAadj))
        call put_d('varstore', reshape(varstore, (/ size(varstore) /)), size(varstore))
        call put_d('varindist', reshape(varindist, (/ size(varindist) /)), size(varindist))
        call put_d('pred', pred, nind)
        if (BayesMethod /= BayesA) then
            call put_i('snptracker', reshape(snptracker, (/ 2*nloci /)), 2*nloci)
            call put_i('snpindist', reshape(snpindist, (/ size(snpindist) /)), size(snpindist))
            call put_d('snpstore', reshape(snpstore, (/ size(snpstore) /)), size(snpstore))
            call put_d('indiststore', reshape(indiststore, (/ size(indiststore) /)), size(indiststore))
            call put_s('included_store', included_store)
            call put_d('p', reshape(p, (/ size(p) /)), size(p))
        end if
        if (BayesMethod /= BayesA .and. BayesMethod /= BayesB .and. BayesMethod /= BayesC) then
            call put_d('pstore', reshape(pstore, (/ size(pstore) /)), size(pstore))
        end if
        if (BayesMethod == BayesA .or. BayesMethod == BayesB) then
            call put_d('vara_s', vara_s, nloci)
            call put_d('vara_sstore', vara_sstore, nloci)
        else
            call put_s('vara_store', vara_store)
        end if
        k = 0
        close(21)
    end subroutine run_soda

    ! Test hook (no counterpart in the reference): get_aug_matrix's loop
    ! (soda.f90:445-518) again, with block `bshift`'s d lowered by `shift`, so
    ! that dI - G is indefinite there and the reference's own choleskyPD /
    ! choleskyPSD / skip bookkeeping run.  d is taken from xxd, so the
    ! power method (and its stream-0 draws) is not repeated.
    subroutine shifted_aug(bshift, shift)
        integer, intent(in) :: bshift
        real(real64), intent(in) :: shift
        integer :: i, j, k, jj, kk, ls, le, n, rank, status
        integer, allocatable :: piv(:)
        real(real64), allocatable :: xxtemp(:,:)
        real(real64) :: d
        allocate(xxtemp(blocksize, blocksize), piv(blocksize))
        ! setup_soda frees the packed genotypes on return (soda.f90:436-441)
        nbatch = (nind - 1) / (int64 * 4) + 1
        allocate(genbits(nbatch, nloci))
        call load_geno_bits
        skip_j = .false.
        i_skipped = 0
        nskipped = 0
        naind = naind + naless
        naless = nloci
        XA = 0.0d0
        do i = 1, nblocks
            if (i <= blockt) then
                ls = (i - 1) * blocksize + 1
                le = min(ls + blocksize - 1, nloci)
            else
                ls = blocksize * blockt + (i - blockt - 1) * (blocksize - 1) + 1
                le = min(ls + blocksize - 2, nloci)
            end if
            n = le - ls + 1
            xxtemp = 0.0d0
            do j = 1, n
                jj = permvec(ls + j - 1)
                do k = 1, j
                    kk = permvec(ls + k - 1)
                    xxtemp(k, j) = xxproduct(jj, kk)
                    xxtemp(j, k) = xxtemp(k, j)
                end do
            end do
            d = xxd(i)
            if (i == bshift) d = d - shift
            xxd(i) = d
            xxtemp(:,:) = -xxtemp(:,:)
            do j = 1, n
                xxtemp(j, j) = d + xxtemp(j, j)
            end do
            call choleskyPD(xxtemp(1:n,1:n), XA(1:n,ls:le), n, status)
            if (status /= 0) then
                call choleskyPSD(xxtemp(1:n,1:n), piv, n, rank, status)
                if (status /= 0) stop 'golden_driver: D-XX is not positive semidefinite'
                do j = 1, n
                    jj = piv(j)
                    if (j <= rank) then
                        XA(1:j, ls+jj-1) = xxtemp(1:j, j)
                    else
                        skip_j(ls+jj-1) = .true.
                        nskipped = nskipped + 1
                        i_skipped(nskipped) = permvec(ls+jj-1)
                    end if
                end do
            else
                naless = 0
            end if
            naless = min(naless, blocksize - n + nskipped)
        end do
        naind = naind - naless
        deallocate(xxtemp, piv, genbits)
    end subroutine shifted_aug

    subroutine run_matrix(which)
        character(len=*), intent(in) :: which
        integer(int64) :: n8, s1
        integer :: n, status, rank
        real(real64), allocatable :: A(:,:), U(:,:)
        integer, allocatable :: piv(:)
        real(real64) :: lam
        integer(int64) :: seed(2)
        character(len=64) :: sval
        character(len=512) :: inpath2
        call get_environment_variable('GOLDEN_IN', inpath2)
        open(unit=78, file=trim(inpath2), access='stream', form='unformatted', status='old')
        read(78) n8
        n = int(n8)
        allocate(A(n,n), U(n,n), piv(n))
        read(78) A
        close(78)
        select case (which)
        case ('pd')
            call choleskyPD(A, U, n, status)
            call put_n('status', status)
            call put_d('U', reshape(U, (/ n*n /)), n*n)
        case ('psd')
            call choleskyPSD(A, piv, n, rank, status)
            call put_n('status', status)
            call put_n('rank', rank)
            call put_i('piv', piv, n)
            call put_d('A', reshape(A, (/ n*n /)), n*n)
        case ('power')
            call get_environment_variable('GOLDEN_SEED', sval)
            read(sval, *) s1
            seed(1) = abs(s1)
            seed(2) = seed(1) + 1
            call rand_set_seed(1, seed)
            call power_method(n, A, lam, 0)
            call put_s('lambda', lam)
        end select
    end subroutine run_matrix

end program golden_driver
