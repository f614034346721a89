! trhepd_gpu.f90 -- iso_c_binding interface to libtrhepd_hip.so (include/trhepd.h).
!
! The Fortran driver of sim-trhepd-rheed keeps its decks, its bulk stage and its output;
! srfref (sim-trhepd-rheed/src/surflf.f90:7-166) hands the per-angle integration to the
! device through these bindings.  Every argument is passed exactly as the driver holds it:
! column-major complex(8) arrays, 1-based index maps.
module trhepd_gpu
        use iso_c_binding
        implicit none

        ! status codes (TRHEPD_OK ... TRHEPD_E_NOTFOUND)
        integer(c_int), parameter :: TRHEPD_OK = 0, TRHEPD_E_ARG = 1, TRHEPD_E_HIP = 2, &
                TRHEPD_E_NOMEM = 3, TRHEPD_E_BUDGET = 4, TRHEPD_E_SINGULAR = 5, &
                TRHEPD_E_NONFINITE = 6, TRHEPD_E_NOTFOUND = 7
        ! arithmetic of the stage products
        integer(c_int), parameter :: TRHEPD_PREC_FP64 = 0, TRHEPD_PREC_MIXED = 1

        type, bind(C) :: trhepd_opts
                integer(c_int) :: device
                integer(c_int) :: precision
                integer(c_size_t) :: hbm_budget
        end type

        ! the mod_scpot3 globals (surfio.f90:18-27) that comp_scpot reads
        type, bind(C) :: trhepd_potential
                integer(c_int) :: nv, nvb, natms, natmb, nelm, nabr, nabi
                real(c_double) :: cc, smesh
                type(c_ptr) :: ielm, z, dchg, ghk, ocr, b, asf, st0, st1, st2
        end type

        interface
                integer(c_int) function trhepd_ctx_create(opts, ctx) bind(C, name="trhepd_ctx_create")
                        import :: c_int, c_ptr, trhepd_opts
                        type(trhepd_opts), intent(in) :: opts
                        type(c_ptr), intent(out) :: ctx
                end function
                integer(c_int) function trhepd_ctx_destroy(ctx) bind(C, name="trhepd_ctx_destroy")
                        import :: c_int, c_ptr
                        type(c_ptr), value :: ctx
                end function
                type(c_ptr) function trhepd_last_error(ctx) bind(C, name="trhepd_last_error")
                        import :: c_ptr
                        type(c_ptr), value :: ctx
                end function
                integer(c_int) function trhepd_table_build(ctx, pot, dz, ns, method, table_id) &
                                bind(C, name="trhepd_table_build")
                        import :: c_int, c_ptr, c_double, trhepd_potential
                        type(c_ptr), value :: ctx
                        type(trhepd_potential), intent(in) :: pot
                        real(c_double), value :: dz
                        integer(c_int), value :: ns, method
                        integer(c_int), intent(out) :: table_id
                end function
                ! RKN table over a list of slice thicknesses (no counterpart in the reference)
                integer(c_int) function trhepd_table_build_steps(ctx, pot, method, ns, steps, z0, table_id) &
                                bind(C, name="trhepd_table_build_steps")
                        import :: c_int, c_ptr, c_double, trhepd_potential
                        type(c_ptr), value :: ctx
                        type(trhepd_potential), intent(in) :: pot
                        integer(c_int), value :: method, ns
                        real(c_double), intent(in) :: steps(*)
                        real(c_double), value :: z0
                        integer(c_int), intent(out) :: table_id
                end function
                ! h211b_control.f90:1-16, callable from C and Fortran alike
                real(c_double) function trhepd_h211b_limit(u) bind(C, name="trhepd_h211b_limit")
                        import :: c_double
                        real(c_double), value :: u
                end function
                real(c_double) function trhepd_h211b_filter(c1, c0, rho, k) bind(C, name="trhepd_h211b_filter")
                        import :: c_int, c_double
                        real(c_double), value :: c1, c0, rho
                        integer(c_int), value :: k
                end function
                ! slice thicknesses for [0, zmax] chosen by the H211b controller on pilot angles
                integer(c_int) function trhepd_adaptive_steps(ctx, pot, method, zmax, tol, h0, hmin, hmax, nb, nang, &
                                iv, gamma, f, nb0, max_steps, steps, nsteps, nforced) &
                                bind(C, name="trhepd_adaptive_steps")
                        import :: c_int, c_ptr, c_double, c_double_complex, trhepd_potential
                        type(c_ptr), value :: ctx
                        type(trhepd_potential), intent(in) :: pot
                        integer(c_int), value :: method, nb, nang, nb0, max_steps
                        real(c_double), value :: zmax, tol, h0, hmin, hmax
                        integer(c_int), intent(in) :: iv(*)
                        complex(c_double_complex), intent(in) :: gamma(*), f(*)
                        real(c_double), intent(out) :: steps(*)
                        integer(c_int), intent(out) :: nsteps, nforced
                end function
                integer(c_int) function trhepd_table_release(ctx, table_id) bind(C, name="trhepd_table_release")
                        import :: c_int, c_ptr
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: table_id
                end function
                integer(c_int) function trhepd_propagate_batch(ctx, table_id, nb, nang, iv, gamma, f, f2, &
                                nb0, wnsga2) bind(C, name="trhepd_propagate_batch")
                        import :: c_int, c_ptr, c_double, c_double_complex
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: table_id, nb, nang, nb0
                        integer(c_int), intent(in) :: iv(nb, nb)
                        complex(c_double_complex), intent(in) :: gamma(nb, nang)
                        complex(c_double_complex), intent(inout) :: f(nb, nb, nang)
                        real(c_double), intent(out) :: f2(nb, nang)
                        real(c_double), intent(in) :: wnsga2(nang)
                end function
                integer(c_int) function trhepd_set_renorm_period(ctx, period) bind(C, name="trhepd_set_renorm_period")
                        import :: c_int, c_ptr
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: period
                end function
                integer(c_int) function trhepd_set_phase_offsets(ctx, mode) bind(C, name="trhepd_set_phase_offsets")
                        import :: c_int, c_ptr
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: mode
                end function
                integer(c_int) function trhepd_set_mixed_digits(ctx, digits) bind(C, name="trhepd_set_mixed_digits")
                        import :: c_int, c_ptr
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: digits
                end function
                ! blkref's angle and group loops (bulk.f90:51-124); phase is complex(nb) or
                ! c_null_ptr, layers integer(nang, ngr) or c_null_ptr
                integer(c_int) function trhepd_bulk_reflect(ctx, table_id, ngr, nbg, nbgm, iv, nang, gamma, &
                                phase, ml, epsb, blocks, layers) bind(C, name="trhepd_bulk_reflect")
                        import :: c_int, c_ptr, c_double, c_double_complex
                        type(c_ptr), value :: ctx
                        integer(c_int), value :: table_id, ngr, nbgm, nang, ml
                        integer(c_int), intent(in) :: nbg(ngr), iv(nbgm, *)
                        complex(c_double_complex), intent(in) :: gamma(*)
                        type(c_ptr), value :: phase
                        real(c_double), value :: epsb
                        complex(c_double_complex), intent(out) :: blocks(*)
                        type(c_ptr), value :: layers
                end function
        end interface

contains

        ! prints the library's message and stops on a non-zero status (the reference has no
        ! error path at this interface: surflf.f90:40 aborts on an unknown method)
        subroutine trhepd_check(rc, ctx, what)
                integer(c_int), intent(in) :: rc
                type(c_ptr), intent(in) :: ctx
                character(*), intent(in) :: what
                character(kind=c_char), pointer :: msg(:)
                integer :: n
                if (rc == TRHEPD_OK) return
                call c_f_pointer(trhepd_last_error(ctx), msg, [1024])
                n = 0
                do while (n < 1024)
                        if (msg(n + 1) == c_null_char) exit
                        n = n + 1
                end do
                write (0, '(a,a,i0,a,1024a)') what, ': trhepd status ', rc, ': ', msg(1:n)
                error stop 1
        end subroutine

end module
