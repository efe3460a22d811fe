! Calls the host-side entry points of libsodabayes through the soda_gpu
! binding (no GPU work): seeds stream 0 the way set_random_seed does
! (utilities.f90:504-517), permutes 1000 SNPs in blocks of 64 the way
! permutate does (utilities.f90:588-607), and prints the state words and
! the first permuted ids for the test suite to compare.
program binding_check
    use iso_c_binding
    use soda_gpu
    implicit none
    integer(c_int64_t) :: st(2)
    integer(c_int32_t) :: perm(1000)
    integer(c_int) :: rc
    rc = sb_rng_seed(123456789_c_int64_t, 0_c_int32_t, st)
    if (rc /= SB_OK) stop 1
    print '(2(i0,1x))', st(1), st(2)
    rc = sb_permutate(1000_c_int32_t, 64_c_int32_t, st, perm)
    if (rc /= SB_OK) stop 2
    print '(10(i0,1x))', perm(1:10)
    print '(2(i0,1x))', st(1), st(2)
    rc = sb_rng_seed(1_c_int64_t, -1_c_int32_t, st)
    print '(i0,1x,a)', rc, trim(sb_error_message())
end program binding_check
