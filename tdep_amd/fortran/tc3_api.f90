! iso_c_binding interface to libtc3.so (include/tc3.h).
!
! A Fortran caller of the thermal_conductivity path (src/thermal_conductivity/main.f90)
! replaces dr%generate, sr%generate, get_kappa, iterative_solution and get_kappa_offdiag
! with the calls below.  Arrays are passed as flat, column-major buffers in Hartree atomic
! units; integer status 0 = success, 1-7 = lo_exitcode_* (konstanter.f90:160-172).
module tc3_api
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: TC3_PREC_FP64 = 0, TC3_PREC_FP32 = 1, TC3_PREC_BF16X3 = 2

    interface
        integer(c_int) function tc3_create(ctx, device) bind(C, name="tc3_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device
        end function

        subroutine tc3_destroy(ctx) bind(C, name="tc3_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine

        type(c_ptr) function tc3_last_error(ctx) bind(C, name="tc3_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function tc3_set_structure(ctx, na, latvec, rfrac, species, mass) &
                bind(C, name="tc3_set_structure")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: na
            real(c_double), intent(in) :: latvec(3, 3), rfrac(3, *), mass(*)
            integer(c_int), intent(in) :: species(*)
        end function

        integer(c_int) function tc3_natural_isotopes(symbol, maxn, conc, mass, n) &
                bind(C, name="tc3_natural_isotopes")
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: symbol(*)
            integer(c_int), value :: maxn
            real(c_double), intent(out) :: conc(*), mass(*)
            integer(c_int), intent(out) :: n
        end function

        integer(c_int) function tc3_set_qmesh_sym(ctx, dims, timereversal, qpsymmetry) &
                bind(C, name="tc3_set_qmesh_sym")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: dims(3)
            integer(c_int), value :: timereversal, qpsymmetry
        end function

        integer(c_int) function tc3_set_thirdorder(ctx, enable) bind(C, name="tc3_set_thirdorder")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: enable
        end function

        integer(c_int) function tc3_set_fc2(ctx, npair, a1, a2, lv2, m) bind(C, name="tc3_set_fc2")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: npair
            integer(c_int), intent(in) :: a1(*), a2(*)
            real(c_double), intent(in) :: lv2(3, *), m(3, 3, *)
        end function

        integer(c_int) function tc3_set_fc3(ctx, ntrip, a1, a2, a3, lv2, lv3, m) bind(C, name="tc3_set_fc3")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ntrip
            integer(c_int), intent(in) :: a1(*), a2(*), a3(*)
            real(c_double), intent(in) :: lv2(3, *), lv3(3, *), m(3, 3, 3, *)
        end function

        integer(c_int) function tc3_set_montecarlo4(ctx, qg4ph) bind(C, name="tc3_set_montecarlo4")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: qg4ph(3)
        end function

        integer(c_int) function tc3_set_fc4(ctx, nquart, a1, a2, a3, a4, lv2, lv3, lv4, m) bind(C, name="tc3_set_fc4")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nquart
            integer(c_int), intent(in) :: a1(*), a2(*), a3(*), a4(*)
            real(c_double), intent(in) :: lv2(3, *), lv3(3, *), lv4(3, *), m(3, 3, 3, 3, *)
        end function

        integer(c_int) function tc3_set_isotopes(ctx, disorder, enable) bind(C, name="tc3_set_isotopes")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: disorder(*)
            integer(c_int), value :: enable
        end function

        integer(c_int) function tc3_set_montecarlo(ctx, qg3ph, seed) bind(C, name="tc3_set_montecarlo")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: qg3ph(3)
            integer(c_int), value :: seed
        end function

        integer(c_int) function tc3_set_boundary(ctx, mfp_max) bind(C, name="tc3_set_boundary")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: mfp_max
        end function

        integer(c_int) function tc3_set_qmesh(ctx, dims, timereversal) bind(C, name="tc3_set_qmesh")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(in) :: dims(3)
            integer(c_int), value :: timereversal
        end function

        integer(c_int) function tc3_mesh_info(ctx, nsym, nfull, nirr, ntriplets, nrows) &
                bind(C, name="tc3_mesh_info")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: nsym, nfull, nirr, nrows
            integer(c_int64_t), intent(out) :: ntriplets
        end function

        integer(c_int) function tc3_harmonic(ctx, omega_min) bind(C, name="tc3_harmonic")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: omega_min
        end function

        integer(c_int) function tc3_get_harmonic(ctx, full, omega, vel, egv) bind(C, name="tc3_get_harmonic")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: full
            type(c_ptr), value :: omega, vel, egv
        end function

        integer(c_int) function tc3_scatter(ctx, nT, temperatures, precision, integrationtype, sigma, classical) &
                bind(C, name="tc3_scatter")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nT, precision, integrationtype, classical
            real(c_double), intent(in) :: temperatures(*)
            real(c_double), value :: sigma
        end function

        integer(c_int) function tc3_get_scattering(ctx, iT, linewidth, F0, scalar_mfp) &
                bind(C, name="tc3_get_scattering")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: iT
            type(c_ptr), value :: linewidth, F0, scalar_mfp
        end function

        integer(c_int) function tc3_kappa(ctx, maxiter, tol, classical, kappa_sma, kappa_iter, kappa_offdiag, niter) &
                bind(C, name="tc3_kappa")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: maxiter, classical
            real(c_double), value :: tol
            type(c_ptr), value :: kappa_sma, kappa_iter, kappa_offdiag, niter
        end function

        integer(c_int) function tc3_get_kappa_log(ctx, iT, maxlen, log, nlog) bind(C, name="tc3_get_kappa_log")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: iT, maxlen
            real(c_double), intent(out) :: log(7, *)
            integer(c_int), intent(out) :: nlog
        end function

        integer(c_int) function tc3_get_modes(ctx, iT, Fn, kappa_mode) bind(C, name="tc3_get_modes")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: iT
            type(c_ptr), value :: Fn, kappa_mode
        end function

        integer(c_int) function tc3_get_timings(ctx, t, n) bind(C, name="tc3_get_timings")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: t(*)
            integer(c_int), value :: n
        end function

        integer(c_int) function tc3_write_grid_hdf5(ctx, iT, path) bind(C, name="tc3_write_grid_hdf5")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            integer(c_int), value :: iT
            character(kind=c_char), intent(in) :: path(*)
        end function

        integer(c_int) function tc3_write_cumulative_hdf5(ctx, iT, path, mfppts, freqpts, dosintegrationtype, &
                                                          sigma, dossigma, unit, atomic_species) &
                bind(C, name="tc3_write_cumulative_hdf5")
            import :: c_int, c_ptr, c_char, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: iT, mfppts, freqpts, dosintegrationtype
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: sigma, dossigma
            character(kind=c_char), intent(in) :: unit(*), atomic_species(*)
        end function

        integer(c_int) function tc3_adjoint_gamma(ctx, dL_dlinewidth, precision, dL_dfc3) &
                bind(C, name="tc3_adjoint_gamma")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: dL_dlinewidth(*)
            integer(c_int), value :: precision
            real(c_double), intent(out) :: dL_dfc3(3, 3, 3, *)
        end function

        integer(c_int) function tc3_adjoint_kappa_rta(ctx, dL_dkappa, precision, dL_dfc3) &
                bind(C, name="tc3_adjoint_kappa_rta")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: dL_dkappa(3, 3, *)
            integer(c_int), value :: precision
            real(c_double), intent(out) :: dL_dfc3(3, 3, 3, *)
        end function

        integer(c_int) function tc3_jvp(ctx, dfc2, dfc3, dlinewidth, domega, dvel, dkappa_sma) &
                bind(C, name="tc3_jvp")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: dfc2(3, 3, *), dfc3(3, 3, 3, *)
            real(c_double), intent(out) :: dlinewidth(*), domega(*), dvel(3, *), dkappa_sma(3, 3, *)
        end function

        integer(c_int) function tc3_get_psi2(ctx, precision, n, triplets, psi2, qpoints) &
                bind(C, name="tc3_get_psi2")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: precision, n
            integer(c_int64_t), intent(in) :: triplets(*)
            real(c_double), intent(out) :: psi2(*)
            type(c_ptr), value :: qpoints
        end function
    end interface

contains

    ! Message of the last failure on ctx as a Fortran string.
    function tc3_error_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        interface
            integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
                import :: c_ptr, c_size_t
                type(c_ptr), value :: p
            end function
        end interface
        type(c_ptr) :: p
        p = tc3_last_error(ctx)
        n = int(c_strlen(p))
        call c_f_pointer(p, s, [n])
        allocate (character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function
end module tc3_api
