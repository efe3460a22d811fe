! sodabayes_gpu: the reference program with its SODA branch on the GPU.
!
! Same command line, .par job files and output files as the reference
! (SODABayes.f90).  The reference's own modules do the option parsing, the
! PLINK / phenotype readers, write_log, output_model and write_dgv; the
! soda_gpu_host module replaces setup_soda, soda_bayes and compute_dgv
! (SODABayes.f90:139-148) with the gfx950 library, and -SODAOff's
! single-site sampler (models.f90:607-962) too, unless SODABAYES_REGULAR_HOST
! keeps the reference's host version.
!
! Jobs on the GPU never build the reference's dense FP64 genotype matrix
! (8 * nind * nloci bytes on the host): the .bed is read packed, and the
! allele frequencies come from its counts (soda_load_packed).  A chain on
! the host keeps the reference's loaders.
!
! Deliberate deviations: the .bim file is closed after load_geno, so that
! a .par file may hold several jobs (the reference stops at the second one).
! With neither -blocksize nor -nblocks the reference
! partitions into blocks of nthreads SNPs, and with -nthreads 1 (its
! default) setup_soda switches SODA off but soda_bayes still runs and
! crashes (soda.f90:418-421, SODABayes.f90:134-142); here that job runs the
! single-site sampler instead.
program sodabayes_gpu
    use globals
    use utilities
    use useroptions
    use models, only: regular_bayes
    use soda_gpu_host
    use omp_lib
    implicit none

    call get_cmdline
    call parse_parfile
    if (useparfile) then
        call run_parfile()
    else
        call get_default_options
        call parse_cmdline
        call run_reported()
    end if

contains

    subroutine run_reported()
        integer :: status
        write(*, '(a)') 'job ' // trim(job_name) // ' started...'
        call run_job(status)
        if (status /= 0) then
            write(*, '(a)') 'job ' // trim(job_name) // ' did not complete.'
            stop
        end if
        write(*, '(a)') 'job ' // trim(job_name) // ' completed.'
    end subroutine run_reported

    ! Jobs of a .par file (SODABayes.f90:19-97): "begin <name>" ... "end"
    ! blocks of "option [value]" lines, '#' or '!' starting a comment; each
    ! job takes the defaults, then its block's options, then the command
    ! line.
    subroutine run_parfile()
        integer, parameter :: maxpar = 100
        character(len=1024), allocatable :: parameters(:)
        character(len=1024) :: line, word, up
        character(len=200) :: values(100)
        logical :: there, injob
        integer :: u, ios, npar, last, nitems, first

        inquire(file=parfil, exist=there)
        if (.not. there) then
            print *, 'par file ', trim(parfil), ' not found'
            stop
        end if
        open(newunit=u, file=trim(parfil), status='old', form='formatted', iostat=ios)
        if (ios /= 0) then
            print *, 'cannot open parfile:', parfil
            stop
        end if
        allocate(parameters(maxpar))
        injob = .false.
        npar = 0
        job_name = ''
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call left_trim(line)
            if (len_trim(line) == 0) cycle
            last = scan(line, '#!') - 1
            if (last == 0) cycle
            if (last < 0) last = len_trim(line)
            line = trim(line(1:last))
            read(line, *) word
            if (ssmatch(word, 'begin')) then
                if (injob) then
                    write(*, *) 'error: Need an END statement for job: ', job_name
                    stop
                end if
                up = line
                call to_upper(up)
                first = verify(up, 'BEGIN')
                job_name = line(first:last)
                call left_trim(job_name)
                parameters = ''
                npar = 0
                injob = .true.
            else if (ssmatch(word, 'end')) then
                if (.not. injob) then
                    write(*, *) 'error: need a BEGIN statement to start a job.'
                    stop
                end if
                call get_default_options
                call parse_parfile_options(npar, parameters)
                call parse_cmdline
                call run_reported()
                injob = .false.
            else
                nitems = nfields(line, ' ')
                if (nitems > 2) then
                    write(*, '(a)') trim(line) // ' has more than two items'
                    stop
                end if
                call get_fields(line, ' ', nitems, values)
                parameters(npar + 1:npar + nitems) = values(1:nitems)
                npar = npar + nitems
            end if
        end do
        close(u)
    end subroutine run_parfile

    ! run_program_sodaBayes (SODABayes.f90:114-165) with the SODA branch and
    ! the genomic values on the GPU.
    subroutine run_job(status)
        integer, intent(inout) :: status
        real(real64) :: t0, tsetup, t1
        character(len=100) :: s1, s2, s3
        logical :: bim_open, on_gpu

        status = 1
        call date_and_time(date=cdate, time=ctime)
        t0 = omp_get_wtime()
        call omp_set_dynamic(.false.)
        maxthreads = omp_get_max_threads()
        call parse
        open(unit=21, file=logfil, status='unknown', form='formatted')
        write(21, '(a)') 'Running program SODABayes: ' // trim(job_name)
        write(21, 101) 'Run started at', cdate(1:4), cdate(5:6), cdate(7:8), ctime(1:2), ctime(3:4), ctime(5:6)
        call get_size
        if (soda_gpu_packed_input()) then
            ! the GPU reads the packed .bed: no dense genotype matrix on the host
            call soda_allocate_packed()
            call soda_load_packed()
        else
            call allocate_data
            call load_geno
            ! load_geno leaves the .bim connected to unit 40 (utilities.f90:414),
            ! which stops the reference at the next job's get_size
            inquire(unit=40, opened=bim_open)
            if (bim_open) close(40)
            call xscale
        end if
        if (mcmc) then
            call load_pheno
            call set_random_seed
            on_gpu = soda_setup_gpu()   ! sets SODAOFF for blocks of one SNP
            tsetup = omp_get_wtime()
            call write_log
            if (on_gpu) then
                call soda_bayes_gpu()
            else
                call regular_bayes
            end if
        else
            call setup_prediction
            call soda_predict_gpu()
            call write_dgv
            call write_log
        end if
        call soda_gpu_release()
        call clean_up
        t1 = omp_get_wtime()
        call date_and_time(date=cdate, time=ctime)
        write(21, 101) 'Run ended at', cdate(1:4), cdate(5:6), cdate(7:8), ctime(1:2), ctime(3:4), ctime(5:6)
        if (mcmc) then
            write(s1, '(f16.2)') tsetup - t0
            write(s2, '(f16.2)') t1 - tsetup
            write(s3, '(f16.2)') t1 - t0
            write(21, 102) 'Walltime used (seconds)'
            write(21, '(a)') 'Setup: ' // adjustl(trim(s1)) // ' MCMC: ' // adjustl(trim(s2)) // &
                ' Total: ' // adjustl(trim(s3))
            close(21)
        end if
        status = 0
101     format(a20, 1x, a4, '-', a2, '-', a2, ' ', a2, ':', a2, ':', a2)
102     format(a31)
    end subroutine run_job

end program sodabayes_gpu
