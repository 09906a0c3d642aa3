! mgp_pressuresolver_mod -- the Fortran-side mirror of MGLET-base's
! pressuresolver_mod (src/flow/pressuresolver_mod.F90:45) on top of libmgp.
!
! Same public names and call signatures for the time loop:
!   init_pressuresolver(...)    (src/flow/pressuresolver_mod.F90:47-145)
!   mgpoisl(u, v, w, p, dt, ittot, irk)             (:160-171)
!   finish_pressuresolver()                         (:148-157)
! Inside MGLET, init_pressuresolver takes its arguments from the grid
! module, the pointer module and fort7; here they are explicit so the module
! builds on its own (the MGLET-side glue fills them from those modules).
! Fields are the flat `field_t%arr` arrays in the
! reference layout; errors stop the program through `errr`-like handling.
MODULE mgp_pressuresolver_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE mgp_bindc_mod
    IMPLICIT NONE
    PRIVATE

    TYPE(C_PTR), SAVE :: handle = C_NULL_PTR
    TYPE(mgp_stats_t), SAVE, PUBLIC :: last_stats

    PUBLIC :: init_pressuresolver, mgpoisl, finish_pressuresolver

CONTAINS

    ! ityp: 0 "sip", 1 "sor", 2 "classic"; the integer arrays are the
    ! reference's GRIDINFO columns, itypboconds(1, :, :), ip3d/ip1d*.
    SUBROUTINE init_pressuresolver(ngrid, minlevel, maxlevel, kk, jj, ii, &
            level, nbrgrid, itypbc, ip3, ip1x, ip1y, ip1z, dx, dy, dz, ddx, &
            ddy, ddz, bp, ityp, ninner, nouter, nouter_min, epcorr, omega, &
            rho, ibtype, precision, eps_rel)
        INTEGER(C_INT), INTENT(in) :: ngrid, minlevel, maxlevel
        INTEGER(C_INT), INTENT(in), TARGET :: kk(:), jj(:), ii(:), level(:)
        INTEGER(C_INT), INTENT(in), TARGET :: nbrgrid(:, :), itypbc(:, :)
        INTEGER(C_INT64_T), INTENT(in), TARGET :: ip3(:), ip1x(:), ip1y(:), &
            ip1z(:)
        REAL(C_DOUBLE), INTENT(in), TARGET :: dx(:), dy(:), dz(:), ddx(:), &
            ddy(:), ddz(:), bp(:)
        INTEGER(C_INT), INTENT(in) :: ityp, ninner, nouter, nouter_min
        REAL(C_DOUBLE), INTENT(in) :: epcorr, omega, rho, eps_rel
        INTEGER(C_INT), INTENT(in) :: ibtype, precision

        TYPE(mgp_grids_t) :: g
        TYPE(mgp_config_t) :: cfg
        INTEGER(C_INT) :: rc

        g%ngrid = ngrid
        g%minlevel = minlevel
        g%maxlevel = maxlevel
        g%kk = C_LOC(kk)
        g%jj = C_LOC(jj)
        g%ii = C_LOC(ii)
        g%level = C_LOC(level)
        g%nbrgrid = C_LOC(nbrgrid)
        g%itypbc = C_LOC(itypbc)
        g%ip3 = C_LOC(ip3)
        g%ip1x = C_LOC(ip1x)
        g%ip1y = C_LOC(ip1y)
        g%ip1z = C_LOC(ip1z)

        cfg%ityp = ityp
        cfg%ninner = ninner
        cfg%nouter = nouter
        cfg%nouter_min = nouter_min
        cfg%epcorr = epcorr
        cfg%omega = omega
        cfg%rho = rho
        cfg%ibtype = ibtype
        cfg%precision = precision
        cfg%eps_rel = eps_rel

        rc = mgp_init(g, cfg, C_LOC(dx), C_LOC(dy), C_LOC(dz), C_LOC(ddx), &
            C_LOC(ddy), C_LOC(ddz), C_LOC(bp), handle)
        CALL check(rc, "init_pressuresolver")
    END SUBROUTINE init_pressuresolver


    SUBROUTINE mgpoisl(u, v, w, p, dt, ittot, irk)
        REAL(C_DOUBLE), INTENT(inout), TARGET, CONTIGUOUS :: u(:), v(:), &
            w(:), p(:)
        REAL(C_DOUBLE), INTENT(in) :: dt
        INTEGER(C_INT), INTENT(in) :: ittot, irk
        INTEGER(C_INT) :: rc

        rc = mgp_mgpoisl(handle, C_LOC(u), C_LOC(v), C_LOC(w), C_LOC(p), dt, &
            ittot, irk, last_stats)
        CALL check(rc, "mgpoisl")
    END SUBROUTINE mgpoisl


    SUBROUTINE finish_pressuresolver()
        INTEGER(C_INT) :: rc

        IF (.NOT. C_ASSOCIATED(handle)) RETURN
        rc = mgp_finish(handle)
        handle = C_NULL_PTR
        CALL check(rc, "finish_pressuresolver")
    END SUBROUTINE finish_pressuresolver


    SUBROUTINE check(rc, what)
        INTEGER(C_INT), INTENT(in) :: rc
        CHARACTER(len=*), INTENT(in) :: what
        IF (rc /= MGP_OK) THEN
            WRITE(*, '(A, ": libmgp status ", I0, ": ", A)') what, rc, &
                TRIM(mgp_error_message())
            ERROR STOP 1
        END IF
    END SUBROUTINE check
END MODULE mgp_pressuresolver_mod
