! Links the Fortran bindings against libtrhepd_hip.so and opens a context on device 0.
! Prints the status of each call; exits with 0 when every call returned TRHEPD_OK.
program abi_check
        use trhepd_gpu
        implicit none
        type(trhepd_opts) :: o
        type(c_ptr) :: ctx
        integer(c_int) :: rc
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        o%device = 0
        o%precision = TRHEPD_PREC_FP64
        o%hbm_budget = 0
        ctx = c_null_ptr
        rc = trhepd_ctx_create(o, ctx)
        print '(a,i0)', 'trhepd_ctx_create status ', rc
        if (rc /= TRHEPD_OK) then
                if (c_associated(ctx)) then
                        call c_f_pointer(trhepd_last_error(ctx), msg, [1024])
                        n = 0
                        do while (n < 1024)
                                if (msg(n + 1) == c_null_char) exit
                                n = n + 1
                        end do
                        print '(1024a)', msg(1:n)
                end if
                stop 3
        end if
        rc = trhepd_set_renorm_period(ctx, -1)
        print '(a,i0)', 'trhepd_set_renorm_period status ', rc
        rc = trhepd_ctx_destroy(ctx)
        print '(a,i0)', 'trhepd_ctx_destroy status ', rc
end program
