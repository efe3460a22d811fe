! SODA branch of the reference program on the GPU.
!
! Replaces the two argument-less calls of the reference's SODA branch
! (SODABayes.f90:139 setup_soda, :142 soda_bayes) and its compute_dgv
! (utilities.f90:616-622) with calls into libsodabayes through the soda_gpu
! binding.  It works on the reference's own module globals (globals.f90), so
! the reference's option parser, readers, write_log, output_model and
! write_dgv run unchanged around it.
!
! Path selection (the reference's command line is kept as it is, so these
! come from the environment):
!   SODABAYES_RNG       = replay (default) | counter
!   SODABAYES_PRECISION = fp64 (default)   | mixed
!   SODABAYES_CHAINS    = number of chains (default 1; more than one needs
!                         counter + mixed).  Chain 0 writes the usual files;
!                         chain k >= 1 writes <out>_chain<k>.eff/.mod/.gv.
! With replay + fp64 the run reproduces the reference's -nthreads 1 chain
! (every draw from its stream 0, whose state is taken from and handed back
! to the reference's generator module).
module soda_gpu_host
    use iso_c_binding
    use globals
    use ParallelRNG, only: ss
    use utilities, only: output_model, write_dgv, load_snp_grp
    use soda_gpu
    implicit none
    private
    public :: soda_setup_gpu, soda_bayes_gpu, soda_predict_gpu, soda_gpu_release
    public :: soda_gpu_packed_input, soda_allocate_packed, soda_load_packed

    type(c_ptr), save :: h = c_null_ptr
    integer(c_int8_t), allocatable, target, save :: bed(:)
    integer(c_int32_t), save :: rng_mode = SB_RNG_REPLAY, precision = SB_PREC_FP64, nchains = 1

contains

    ! Stops the job on a library failure: message on stdout and in the log
    ! (unit 21), as the reference reports its fatal errors, then a nonzero
    ! exit status.
    subroutine fail(what, rc)
        character(len=*), intent(in) :: what
        integer(c_int), intent(in) :: rc
        character(len=300) :: msg
        write(msg, '(a,a,i0,a,a)') what, ' failed (', rc, '): ', trim(sb_error_message())
        write(*, '(a)') trim(msg)
        write(21, '(a)') trim(msg)
        call flush(21)
        error stop 'SODA on the GPU failed'
    end subroutine fail

    subroutine read_env()
        character(len=64) :: v
        integer :: ln, st, c
        rng_mode = SB_RNG_REPLAY
        precision = SB_PREC_FP64
        nchains = 1
        call get_environment_variable('SODABAYES_RNG', v, ln, st)
        if (st == 0 .and. (trim(v) == 'counter' .or. trim(v) == 'COUNTER')) rng_mode = SB_RNG_COUNTER
        call get_environment_variable('SODABAYES_PRECISION', v, ln, st)
        if (st == 0 .and. (trim(v) == 'mixed' .or. trim(v) == 'MIXED')) precision = SB_PREC_MIXED
        call get_environment_variable('SODABAYES_CHAINS', v, ln, st)
        if (st == 0 .and. ln > 0) then
            read(v, *, iostat=st) c
            if (st == 0 .and. c >= 1) nchains = c
        end if
    end subroutine read_env

    ! SNP-major .bed payload (nloci * ceiling(nind/4) bytes after the three
    ! magic bytes, checked as load_geno checks them, utilities.f90:423-438).
    subroutine read_bed()
        integer :: u, ios
        integer(c_int8_t) :: magic(3)
        integer(int64) :: nbytes
        nbytes = int(nloci, int64) * ((nind + 3) / 4)
        if (allocated(bed)) deallocate(bed)
        allocate(bed(nbytes))
        open(newunit=u, file=trim(genfil), access='stream', form='unformatted', status='old', &
             action='read', iostat=ios)
        if (ios /= 0) stop 'cannot open the .bed file'
        read(u) magic
        if (magic(1) /= int(z'6c', c_int8_t) .or. magic(2) /= int(z'1b', c_int8_t)) then
            write(*, *) 'Binary genotype file may not be a plink file'
            write(21, '(a)') 'Binary genotype file may not be a plink file'
            call flush(21)
            stop
        end if
        if (magic(3) /= 1_c_int8_t) then
            write(*, '(a)') 'should not be individual mode - check >>>'
            write(21, '(a)') 'SNP file should not be in individual mode'
            call flush(21)
            stop
        end if
        read(u) bed
        close(u)
    end subroutine read_bed

    ! The single-site sampler (-SODAOff, or a partition into blocks of one
    ! SNP, soda.f90:407-421) stays the reference's host code when
    ! SODABAYES_REGULAR_HOST is set; otherwise it runs on the GPU.
    logical function regular_on_host()
        integer :: ln, st
        character(len=8) :: v
        call get_environment_variable('SODABAYES_REGULAR_HOST', v, ln, st)
        regular_on_host = st == 0
    end function regular_on_host

    logical function single_site()
        integer :: bs
        if (SODAOFF) then
            single_site = .true.
            return
        end if
        if (setblocksize) then
            bs = blocksize
        else if (setblocks) then
            bs = ceiling(dble(nloci) / nblocks)
        else
            bs = nthreads
        end if
        single_site = bs <= 1
    end function single_site

    ! Whether the job needs no dense genotype matrix on the host: prediction,
    ! or a chain on the GPU.  SODABAYES_HOST_DENSE_X keeps the reference's
    ! dense loaders.
    logical function soda_gpu_packed_input()
        integer :: ln, st
        character(len=8) :: v
        call get_environment_variable('SODABAYES_HOST_DENSE_X', v, ln, st)
        soda_gpu_packed_input = .false.
        if (st == 0) return
        soda_gpu_packed_input = .not. mcmc .or. .not. (single_site() .and. regular_on_host())
    end function soda_gpu_packed_input

    ! allocate_data (utilities.f90:558-586) without the dense genotype matrix
    ! (8 * nind * nloci bytes), which the GPU path never reads: X gets a 1 x 1
    ! placeholder so that clean_up's deallocation stays valid.
    subroutine soda_allocate_packed()
        integer :: ios, ios2
        ios = 0
        ios2 = 0
        if (mcmc) then
            allocate(animalID(nind), snpID(nloci), why(nind), pred(nind), X(1, 1), g(nloci), yadj(nind), &
                     xpx(nloci), permvec(nloci), varindist(ngroup, nmix), varstore(ngroup, nmix), gstore(nloci), &
                     freqstore(nloci), stat=ios)
            if (BayesMethod == BayesA) then
                allocate(vara_s(nloci), vara_sstore(nloci), stat=ios2)
            else if (BayesMethod == BayesB) then
                allocate(vara_s(nloci), vara_sstore(nloci), p(ngroup, nmix), log_p(ngroup, nmix), &
                         snpindist(ngroup, nmix), snptracker(nloci, 2), snpstore(ngroup, nmix), &
                         indiststore(nloci, nmix), stat=ios2)
            else if (BayesMethod == BayesC) then
                allocate(p(ngroup, nmix), log_p(ngroup, nmix), snpindist(ngroup, nmix), snptracker(nloci, 2), &
                         snpstore(ngroup, nmix), indiststore(nloci, nmix), vare_gp(nmix), gp(nmix), stat=ios2)
            else
                allocate(p(ngroup, nmix), log_p(ngroup, nmix), snpindist(ngroup, nmix), snptracker(nloci, 2), &
                         snpstore(ngroup, nmix), indiststore(nloci, nmix), vare_gp(nmix), gp(nmix), &
                         pstore(ngroup, nmix), dirx(ngroup, nmix), stat=ios2)
            end if
        else
            allocate(animalID(nind), snpID(nloci), X(1, 1), freqstore(nloci), gstore(nloci), pred(nind), stat=ios)
        end if
        if (ios /= 0 .or. ios2 /= 0) then
            print *, 'ERROR :: Unable to allocate required storage'
            stop 'Unable to allocate required storage for data'
        end if
    end subroutine soda_allocate_packed

    ! load_geno + xscale (utilities.f90:402-460, 519-556) without the dense
    ! matrix: SNP ids from the .bim, the packed .bed, and the allele
    ! frequencies from its two-bit counts (missing read as 1, as load_geno
    ! reads it).  The counts are the integers xscale sums as doubles, so the
    ! frequencies are the reference's to the last bit; the .frq file is
    ! written in its format.  Prediction reads the frequencies from -freq.
    subroutine soda_load_packed()
        integer :: u, j, k, ios, nb, last
        character(len=80) :: dum, id
        integer(int64) :: cnt, o
        integer :: b, w
        open(newunit=u, file=trim(bimfil), status='old', form='formatted')
        do j = 1, nloci
            read(u, *, iostat=ios) dum, id
            if (ios /= 0) write(*, *) 'error: reading snpID from: ' // trim(bimfil)
            snpID(j) = trim(id)
        end do
        close(u)
        call read_bed()
        if (mcmc) then
            nb = (nind + 3) / 4
            last = nind - 4 * (nb - 1)   ! genotypes in the last byte of a SNP
            do j = 1, nloci
                o = int(j - 1, int64) * nb
                cnt = 0
                do k = 1, nb
                    b = iand(int(bed(o + k)), 255)
                    w = b - iand(ishft(b, -1), 85)          ! two-bit codes -> counts 0 / 1 / 1 / 2
                    if (k == nb .and. last < 4) w = iand(w, ishft(1, 2 * last) - 1)
                    cnt = cnt + iand(w, 3) + iand(ishft(w, -2), 3) + iand(ishft(w, -4), 3) + ishft(w, -6)
                end do
                freqstore(j) = dble(cnt) / (2.0d0 * nind)
            end do
            open(newunit=u, file=trim(freqfil), status='unknown')
            do j = 1, nloci
                write(u, '(A10,1X,F10.6)') snpID(j), freqstore(j)
            end do
            close(u, status='keep')
            if (BayesMethod == BayesRc) call load_snp_grp
        else
            open(newunit=u, file=trim(freqfil), status='old')
            do j = 1, nloci
                read(u, *) dum, freqstore(j)
            end do
            close(u, status='keep')
        end if
    end subroutine soda_load_packed

    ! setup_soda (soda.f90:403-443): block partition, permutation, packed
    ! genotypes on the device and the augmenting matrices.  Sets SODAOFF
    ! when the partition gives blocks of one SNP (soda.f90:418-421).  With
    ! SODAOFF the chain is the single-site sampler (models.f90:607-962), on
    ! the GPU (nblocks = blocksize = 0 at the boundary) unless
    ! SODABAYES_REGULAR_HOST keeps it on the host.  Returns whether the chain
    ! runs on the GPU (soda_bayes_gpu) or on the host (regular_bayes).
    logical function soda_setup_gpu()
        type(sb_dims) :: d
        integer(c_int64_t) :: st(2)
        integer(c_int32_t), allocatable, target :: perm0(:), grp0(:)
        integer(c_int32_t) :: fb, fc
        integer(c_int) :: rc
        real(c_double) :: na(2)
        real(c_double), allocatable :: ids(:)
        type(c_ptr) :: gptr
        integer :: k

        soda_setup_gpu = .false.
        if (single_site()) then
            SODAOFF = .true.
            if (regular_on_host()) return
            call setup_single_site()
            soda_setup_gpu = .true.
            return
        end if
        ! the partition of setup_soda (soda.f90:407-417)
        if (setblocksize) then
            nblocks = ceiling(dble(nloci) / blocksize)
            blockt = nblocks
        else if (setblocks) then
            blocksize = ceiling(dble(nloci) / nblocks)
            blockt = nloci - (blocksize - 1) * nblocks
        else
            blocksize = nthreads
            nblocks = ceiling(dble(nloci) / blocksize)
            blockt = nblocks
        end if
        naind = blocksize
        call read_env()
        if (nchains > 1 .and. (rng_mode /= SB_RNG_COUNTER .or. precision /= SB_PREC_MIXED)) &
            stop 'SODABAYES_CHAINS > 1 needs SODABAYES_RNG=counter and SODABAYES_PRECISION=mixed'
        write(*, '(a,a,a,a,a,i0)') 'SODA on the GPU: rng ', merge('replay ', 'counter', rng_mode == SB_RNG_REPLAY), &
            ', precision ', merge('fp64 ', 'mixed', precision == SB_PREC_FP64), ', chains ', nchains

        ! stream 0 of the reference's generator, as set_random_seed left it
        st(1) = ss(0)%s1
        st(2) = ss(0)%s2
        allocate(perm0(nloci))
        if (rblocks) then
            rc = sb_permutate(int(nloci, c_int32_t), int(blocksize, c_int32_t), st, perm0)
            if (rc /= SB_OK) call fail('sb_permutate', rc)
        else
            perm0 = [(k - 1, k = 1, nloci)]
        end if
        permvec = perm0 + 1

        call read_bed()
        gptr = c_null_ptr
        if (BayesMethod == BayesRc) then
            allocate(grp0(nloci))
            grp0 = snptracker(:, 1) - 1
            gptr = c_loc(grp0)
        end if
        d = sb_dims(int(BayesMethod, c_int32_t), int(nind, c_int32_t), int(nloci, c_int32_t), &
                    int(nmix, c_int32_t), int(ngroup, c_int32_t), int(nblocks, c_int32_t), &
                    int(blocksize, c_int32_t), int(blockt, c_int32_t), nchains, rng_mode, precision, 0_c_int32_t)
        rc = sb_create(d, c_loc(bed), freqstore, perm0, gptr, h)
        if (rc /= SB_OK) call fail('sb_create', rc)
        rc = sb_rng_set(h, st)
        if (rc /= SB_OK) call fail('sb_rng_set', rc)

        allocate(xxd(nblocks), XA(1, 1), yA(1), yAadj(naind), skip_j(nloci), i_skipped(nloci))
        skip_j = .false.
        i_skipped = 0
        rc = sb_build_aug(h, xxd, fb, fc)
        if (rc == SB_ERR_NOT_PD) then
            ! the reference's message when the pivoted factorisation fails too (soda.f90:496-500)
            write(21, '(a)') 'ERROR :: D-XX is not positive semi-definite'
            call fail('sb_build_aug', rc)
        else if (rc /= SB_OK) then
            call fail('sb_build_aug', rc)
        end if
        rc = sb_get(h, SB_Q_NAIND, 0, na, 2_c_int64_t)
        if (rc /= SB_OK) call fail('sb_get(naind)', rc)
        naind = int(na(1))
        nskipped = int(na(2))
        naless = 0
        if (nskipped > 0) then
            allocate(ids(nskipped))
            rc = sb_get(h, SB_Q_SKIPPED, 0, ids, int(nskipped, c_int64_t))
            if (rc /= SB_OK) call fail('sb_get(skipped)', rc)
            i_skipped(1:nskipped) = int(ids)
            skip_j(i_skipped(1:nskipped)) = .true.
        end if
        ! the power-method start vectors came from stream 0
        rc = sb_rng_get(h, st)
        ss(0)%s1 = st(1)
        ss(0)%s2 = st(2)
        soda_setup_gpu = .true.
    end function soda_setup_gpu

    ! The single-site sampler's handle: SNPs in id order (permvec = identity,
    ! models.f90:672-675), no blocks and no augmenting matrices.
    subroutine setup_single_site()
        type(sb_dims) :: d
        integer(c_int64_t) :: st(2)
        integer(c_int32_t), allocatable, target :: perm0(:), grp0(:)
        integer(c_int32_t) :: fb, fc
        integer(c_int) :: rc
        real(c_double) :: xxd0(1)
        type(c_ptr) :: gptr
        integer :: k
        call read_env()
        if (precision /= SB_PREC_FP64) then
            write(*, '(a)') 'SODA off: the single-site sampler runs in FP64'
            precision = SB_PREC_FP64
        end if
        if (nchains > 1 .and. rng_mode /= SB_RNG_COUNTER) &
            stop 'SODABAYES_CHAINS > 1 needs SODABAYES_RNG=counter'
        write(*, '(a,a,a,i0)') 'single-site sampler on the GPU: rng ', &
            merge('replay ', 'counter', rng_mode == SB_RNG_REPLAY), ', chains ', nchains
        allocate(perm0(nloci))
        perm0 = [(k - 1, k = 1, nloci)]
        permvec = perm0 + 1
        call read_bed()
        gptr = c_null_ptr
        if (BayesMethod == BayesRc) then
            allocate(grp0(nloci))
            grp0 = snptracker(:, 1) - 1
            gptr = c_loc(grp0)
        end if
        d = sb_dims(int(BayesMethod, c_int32_t), int(nind, c_int32_t), int(nloci, c_int32_t), &
                    int(nmix, c_int32_t), int(ngroup, c_int32_t), 0_c_int32_t, 0_c_int32_t, 0_c_int32_t, &
                    nchains, rng_mode, precision, 0_c_int32_t)
        rc = sb_create(d, c_loc(bed), freqstore, perm0, gptr, h)
        if (rc /= SB_OK) call fail('sb_create', rc)
        st(1) = ss(0)%s1
        st(2) = ss(0)%s2
        rc = sb_rng_set(h, st)
        if (rc /= SB_OK) call fail('sb_rng_set', rc)
        rc = sb_build_aug(h, xxd0, fb, fc)
        if (rc /= SB_OK) call fail('sb_build_aug', rc)
    end subroutine setup_single_site

    ! soda_bayes (models.f90:9-605) on the device: the .hyp header and lines
    ! (models.f90:24-53, 549-569), the posterior means (models.f90:575-594),
    ! output_model, then the genomic values of the training individuals.
    subroutine soda_bayes_gpu()
        type(sb_priors) :: pr
        real(c_double), allocatable :: hyp(:, :)
        integer(c_int) :: rc
        integer :: i, k, w, rep0, cnt, nrec, r, chunk
        character(len=12) :: c12
        character(len=15) :: c15

        open(unit=25, file=hypfil, status='unknown', form='formatted')
        if (BayesMethod == BayesA) then
            write(*, '(A12,3(1x,A15))') 'Iteration', 'mu', 'Ve', 'Va'
            write(25, '(A12,3(1x,A15))') 'Iteration', 'mu', 'Ve', 'Va'
        else if (BayesMethod == BayesB) then
            write(*, '((A12,1x,A12),4(1x,A15))') 'Iteration', 'ModelSize', 'mu', 'Ve', 'Va1', 'Va2'
            write(25, '((A12,1x,A12),4(1x,A15))') 'Iteration', 'ModelSize', 'mu', 'Ve', 'Va1', 'Va2'
        else
            write(*, '((A12,1x,A12),3(1x,A15))', advance='no') 'Iteration', 'ModelSize', 'mu', 'Vara', 'Ve'
            write(25, '((A12,1x,A12),3(1x,A15))', advance='no') 'Iteration', 'ModelSize', 'mu', 'Vara', 'Ve'
            if (ngroup == 1) then
                do i = 1, nmix
                    write(c12, '(I12)') i
                    c12 = adjustr('N' // trim(adjustl(c12)))
                    write(*, '(1x,A12)', advance='no') c12
                    write(25, '(1x,A12)', advance='no') c12
                end do
                do i = 1, nmix
                    write(c15, '(I15)') i
                    c15 = adjustr('Va' // trim(adjustl(c15)))
                    write(*, '(1x,A15)', advance='no') c15
                    write(25, '(1x,A15)', advance='no') c15
                end do
            end if
            write(*, *)
            write(25, *)
        end if

        ! priors as the reference's parser left them (float options already
        ! rounded through real32 by cast_float)
        pr%pizero = pizero
        pr%scale_va = scale_va
        pr%scale_ve = scale_ve
        pr%df_va = df_va
        pr%df_ve = df_ve
        pr%vara0 = vara
        pr%mix = 0.0d0
        pr%alpha = 0.0d0
        if (allocated(mix)) pr%mix(1:size(mix)) = mix
        if (allocated(alpha)) pr%alpha(1:size(alpha)) = alpha
        pr%numit = numit
        pr%burnin = burnin
        pr%thin = thin
        pr%resync = 0
        pr%seed = seed1
        rc = sb_mcmc_init(h, pr, why)
        if (rc /= SB_OK) call fail('sb_mcmc_init', rc)
        if (df_va <= 0.0d0) scale_va = 0.0d0   ! models.f90:104

        ! iterations in chunks, writing each chunk's .hyp lines
        w = sb_hyp_width(h)
        chunk = max(thin, (1000 / max(thin, 1)) * thin)
        rep0 = 0
        do while (rep0 < numit)
            cnt = min(chunk, numit - rep0)
            nrec = 0
            do r = rep0 + 1, rep0 + cnt
                if (mod(r, thin) == 0) nrec = nrec + 1
            end do
            allocate(hyp(w, max(nrec, 1)))
            rc = sb_mcmc_run(h, int(rep0 + 1, c_int32_t), int(cnt, c_int32_t), hyp)
            if (rc /= SB_OK) call fail('sb_mcmc_run', rc)
            do k = 1, nrec
                call write_hyp(hyp(:, k))
            end do
            deallocate(hyp)
            rep0 = rep0 + cnt
        end do
        close(25)

        ! chain 0 last, so that the module ends with its values, as after
        ! soda_bayes (mu = mu_store, models.f90:597)
        do k = nchains - 1, 0, -1
            if (nchains > 1) call chain_files(k)
            call fetch_chain(k)
            call output_model
            mu = mu_store
            call soda_predict_gpu()
            call write_dgv
        end do
        call fetch_state()
    end subroutine soda_bayes_gpu

    ! One .hyp line (models.f90:549-568) from a record
    ! [rep, included, mu, vara, vare, snpindist(ngroup*nmix), varindist(...)].
    subroutine write_hyp(rec)
        real(c_double), intent(in) :: rec(:)
        integer :: rep, included, nb, i
        integer, allocatable :: sid(:)
        real(real64), allocatable :: vid(:)
        nb = ngroup * nmix
        rep = nint(rec(1))
        included = nint(rec(2))
        allocate(sid(nb), vid(nb))
        sid = [(nint(rec(5 + i)), i = 1, nb)]
        vid = rec(6 + nb:5 + 2 * nb)
        if (BayesMethod == BayesA) then
            write(*, '(I12,3(1x,E15.6))') rep, rec(3), rec(5), vid
            write(25, '(I12,3(1x,E15.6))') rep, rec(3), rec(5), vid
        else if (BayesMethod == BayesB) then
            write(*, '((I12,1x,I12),4(1x,E15.6))') rep, included, rec(3), rec(5), vid
            write(25, '((I12,1x,I12),4(1x,E15.6))') rep, included, rec(3), rec(5), vid
        else
            write(*, '((I12,1x,I12),3(1x,E15.6))', advance='no') rep, included, rec(3), rec(4), rec(5)
            write(25, '((I12,1x,I12),3(1x,E15.6))', advance='no') rep, included, rec(3), rec(4), rec(5)
            if (ngroup == 1) then
                write(*, '(20(1x,I12))', advance='no') sid
                write(*, '(20(1x,E15.6))', advance='no') vid
                write(25, '(20(1x,I12))', advance='no') sid
                write(25, '(20(1x,E15.6))', advance='no') vid
            end if
            write(*, *)
            write(25, *)
        end if
        call flush(25)
    end subroutine write_hyp

    ! Output names of chain k (k = 0: the reference's names, parse_out).
    subroutine chain_files(k)
        integer, intent(in) :: k
        character(len=16) :: sfx
        sfx = ''
        if (k > 0) write(sfx, '(a,i0)') '_chain', k
        efffil = trim(outprefix) // trim(sfx) // '.eff'
        modfil = trim(outprefix) // trim(sfx) // '.mod'
        mbvfil = trim(outprefix) // trim(sfx) // '.gv'
    end subroutine chain_files

    subroutine get_q(what, k, out)
        integer(c_int32_t), intent(in) :: what
        integer, intent(in) :: k
        real(c_double), intent(out) :: out(:)
        integer(c_int) :: rc
        rc = sb_get(h, what, int(k, c_int32_t), out, int(size(out), c_int64_t))
        if (rc /= SB_OK) call fail('sb_get', rc)
    end subroutine get_q

    ! Posterior means of chain k into the module's store arrays, divided by
    ! the number of kept samples as models.f90:575-594 does.
    subroutine fetch_chain(k)
        integer, intent(in) :: k
        real(c_double) :: st(5)
        real(c_double), allocatable :: v(:)
        real(real64) :: counter
        call get_q(SB_Q_STORES, k, st)
        counter = st(1)
        call get_q(SB_Q_GSTORE, k, gstore)
        gstore = gstore / counter
        mu_store = st(2) / counter
        vare_store = st(3) / counter
        allocate(v(ngroup * nmix))
        call get_q(SB_Q_VARSTORE, k, v)
        varstore = reshape(v, [ngroup, nmix]) / counter
        if (BayesMethod /= BayesA) then
            included_store = st(5) / counter
            call get_q(SB_Q_SNPSTORE, k, v)
            snpstore = reshape(v, [ngroup, nmix]) / counter
            block
                real(c_double), allocatable :: ind(:)
                allocate(ind(int(nloci, int64) * nmix))
                call get_q(SB_Q_INDISTSTORE, k, ind)
                indiststore = reshape(ind, [nloci, nmix]) / counter
            end block
            if (BayesMethod /= BayesB .and. BayesMethod /= BayesC) then
                call get_q(SB_Q_PSTORE, k, v)
                pstore = reshape(v, [ngroup, nmix]) / counter
            end if
        end if
        if (BayesMethod == BayesA .or. BayesMethod == BayesB) then
            call get_q(SB_Q_VARA_SSTORE, k, vara_sstore)
            vara_sstore = vara_sstore / counter
        else
            vara_store = st(4) / counter
        end if
    end subroutine fetch_chain

    ! Final state of chain 0 into the module (what soda_bayes leaves there),
    ! and the stream-0 state back into the reference's generator.
    subroutine fetch_state()
        real(c_double) :: sc(4)
        real(c_double), allocatable :: v(:), cl(:)
        integer(c_int64_t) :: st(2)
        integer(c_int) :: rc
        call get_q(SB_Q_SCALARS, 0, sc)
        vare = sc(2)
        vara = sc(3)
        call get_q(SB_Q_G, 0, g)
        call get_q(SB_Q_YADJ, 0, yadj)
        if (allocated(yAadj)) call get_q(SB_Q_YAADJ, 0, yAadj)
        allocate(v(ngroup * nmix))
        call get_q(SB_Q_VARINDIST, 0, v)
        varindist = reshape(v, [ngroup, nmix])
        if (BayesMethod /= BayesA) then
            allocate(cl(nloci))
            call get_q(SB_Q_CLASS, 0, cl)
            snptracker(:, 2) = nint(cl)
            call get_q(SB_Q_SNPINDIST, 0, v)
            snpindist = reshape(nint(v), [ngroup, nmix])
            call get_q(SB_Q_P, 0, v)
            p = reshape(v, [ngroup, nmix])
        end if
        if (BayesMethod == BayesA .or. BayesMethod == BayesB) call get_q(SB_Q_VARA_S, 0, vara_s)
        if (rng_mode == SB_RNG_REPLAY) then
            rc = sb_rng_get(h, st)
            ss(0)%s1 = st(1)
            ss(0)%s2 = st(2)
        end if
    end subroutine fetch_state

    ! compute_dgv (utilities.f90:616-622) on the device: pred = mu + Z gstore
    ! with the standardisation of xscale, from the packed .bed payload.
    subroutine soda_predict_gpu()
        integer(c_int) :: rc
        if (.not. allocated(bed)) call read_bed()
        rc = sb_predict(int(nind, c_int32_t), int(nloci, c_int32_t), c_loc(bed), 0_c_int32_t, freqstore, &
                        gstore, mu, pred)
        if (rc /= SB_OK) call fail('sb_predict', rc)
    end subroutine soda_predict_gpu

    ! The SODA part of clean_up (utilities.f90:776-794) for the device.
    subroutine soda_gpu_release()
        integer(c_int) :: rc
        if (c_associated(h)) then
            rc = sb_destroy(h)
            h = c_null_ptr
        end if
        if (allocated(bed)) deallocate(bed)
    end subroutine soda_gpu_release

end module soda_gpu_host
