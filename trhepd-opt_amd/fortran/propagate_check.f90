! Drives the hot entry points from Fortran, as a patched srfref would (surflf.f90:28-41 for
! the table, surflf.f90:118-139 for the angle loop): reads one deck's potential descriptor and
! one batch of angles from a binary file, builds the table with trhepd_table_build, runs the
! slice loop with trhepd_propagate_batch through the iso_c_binding interface (complex(8)
! arrays by reference, scalars by value), and writes f and f2 back.
!
!   propagate_check <input> <output>
!
! Input (stream, native byte order): 12 int32 nv, nvb, natms, natmb, nelm, nabr, nabi, ns,
! int_method, nb, nang, nb0; 3 real(8) cc, smesh, dz; then ielm(natm) int32, z(natm),
! dchg(nelm), ghk(nv), ocr(natm), b(6, nelm), asf(nv, 6, nelm) real(8); st0(nv, natms),
! st1(nvb, natmb), st2(nvb, natmb) complex(8); iv(nb, nb) int32; gamma(nb, nang),
! f(nb, nb, nang) complex(8); wnsga2(nang) real(8).  Output: f(nb, nb, nang), f2(nb, nang).
program propagate_check
        use trhepd_gpu
        implicit none
        integer(c_int) :: hdr(12), nv, nvb, natms, natmb, nelm, ns, method, nb, nang, nb0, natm, tid, rc
        real(c_double) :: sc(3)
        integer(c_int), allocatable, target :: ielm(:)
        real(c_double), allocatable, target :: z(:), dchg(:), ghk(:), ocr(:), b(:), asf(:)
        complex(c_double_complex), allocatable, target :: st0(:), st1(:), st2(:)
        integer(c_int), allocatable :: iv(:, :)
        complex(c_double_complex), allocatable :: gam(:, :), f(:, :, :)
        real(c_double), allocatable :: wn(:), f2(:, :)
        type(trhepd_potential) :: pot
        type(trhepd_opts) :: o
        type(c_ptr) :: ctx
        character(len=4096) :: fin, fout

        call get_command_argument(1, fin)
        call get_command_argument(2, fout)
        open (10, file=trim(fin), access='stream', form='unformatted', status='old')
        read (10) hdr
        nv = hdr(1); nvb = hdr(2); natms = hdr(3); natmb = hdr(4); nelm = hdr(5)
        ns = hdr(8); method = hdr(9); nb = hdr(10); nang = hdr(11); nb0 = hdr(12)
        read (10) sc
        natm = natms + natmb
        allocate (ielm(natm), z(natm), dchg(nelm), ghk(nv), ocr(natm), b(6*nelm), asf(nv*6*nelm))
        allocate (st0(max(1, nv*natms)), st1(max(1, nvb*natmb)), st2(max(1, nvb*natmb)))
        allocate (iv(nb, nb), gam(nb, nang), f(nb, nb, nang), wn(nang), f2(nb, nang))
        read (10) ielm, z, dchg, ghk, ocr, b, asf
        if (natms > 0) read (10) st0(1:nv*natms)
        if (natmb > 0) read (10) st1(1:nvb*natmb), st2(1:nvb*natmb)
        read (10) iv, gam, f, wn
        close (10)

        pot%nv = nv; pot%nvb = nvb; pot%natms = natms; pot%natmb = natmb; pot%nelm = nelm
        pot%nabr = hdr(6); pot%nabi = hdr(7); pot%cc = sc(1); pot%smesh = sc(2)
        pot%ielm = c_loc(ielm); pot%z = c_loc(z); pot%dchg = c_loc(dchg); pot%ghk = c_loc(ghk)
        pot%ocr = c_loc(ocr); pot%b = c_loc(b); pot%asf = c_loc(asf)
        pot%st0 = c_null_ptr; pot%st1 = c_null_ptr; pot%st2 = c_null_ptr
        if (natms > 0) pot%st0 = c_loc(st0)
        if (natmb > 0) then
                pot%st1 = c_loc(st1)
                pot%st2 = c_loc(st2)
        end if

        o%device = 0
        o%precision = TRHEPD_PREC_FP64
        o%hbm_budget = 0
        rc = trhepd_ctx_create(o, ctx)
        if (rc /= TRHEPD_OK) then
                print '(a,i0)', 'trhepd_ctx_create status ', rc
                stop 3
        end if
        call trhepd_check(trhepd_table_build(ctx, pot, sc(3), ns, method, tid), ctx, 'trhepd_table_build')
        call trhepd_check(trhepd_propagate_batch(ctx, tid, nb, nang, iv, gam, f, f2, nb0, wn), ctx, &
                          'trhepd_propagate_batch')
        call trhepd_check(trhepd_table_release(ctx, tid), ctx, 'trhepd_table_release')
        rc = trhepd_ctx_destroy(ctx)

        open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
        write (11) f, f2
        close (11)
        print '(a,i0,a,i0,a)', 'propagate_check: ', nang, ' angles of ', nb, ' beams'
end program
