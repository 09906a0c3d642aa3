! mgp_bindc_mod -- ISO_C_BINDING interface of libmgp.so (include/mgp.h).
!
! The derived types mirror the C structs field by field (BIND(C) gives
! them the C layout); the functions are the C entry points.  Arrays are
! passed as C_LOC of the caller's contiguous Fortran arrays, the way the
! reference already binds zlib's crc32 (src/core/checksum_mod.F90:18-24).
MODULE mgp_bindc_mod
    USE, INTRINSIC :: ISO_C_BINDING
    IMPLICIT NONE
    PRIVATE

    INTEGER(C_INT), PARAMETER, PUBLIC :: MGP_OK = 0

    ! struct mgp_grids (GRIDINFO subset, src/core/gridio_mod.F90:15-37)
    TYPE, BIND(C), PUBLIC :: mgp_grids_t
        INTEGER(C_INT) :: ngrid = 0, minlevel = 0, maxlevel = 0
        TYPE(C_PTR) :: kk = C_NULL_PTR, jj = C_NULL_PTR, ii = C_NULL_PTR
        TYPE(C_PTR) :: level = C_NULL_PTR, nbrgrid = C_NULL_PTR
        TYPE(C_PTR) :: itypbc = C_NULL_PTR
        TYPE(C_PTR) :: ip3 = C_NULL_PTR, ip1x = C_NULL_PTR
        TYPE(C_PTR) :: ip1y = C_NULL_PTR, ip1z = C_NULL_PTR
        ! GRIDINFO iparent and i/j/kposition; C_NULL_PTR for one level
        TYPE(C_PTR) :: iparent = C_NULL_PTR, position = C_NULL_PTR
    END TYPE mgp_grids_t

    ! struct mgp_config (/flow/pressuresolver keys,
    ! src/flow/pressuresolver_mod.F90:57-142)
    TYPE, BIND(C), PUBLIC :: mgp_config_t
        INTEGER(C_INT) :: ityp = 0, ninner = 5, nouter = 9, nouter_min = 0
        ! omega: the reference's default_value=1.1 is a default-real literal
        ! converted with REAL(..., real64) (src/flow/pressuresolver_mod.F90:135)
        REAL(C_DOUBLE) :: epcorr = 0.0_C_DOUBLE
        REAL(C_DOUBLE) :: omega = REAL(1.1, C_DOUBLE)
        REAL(C_DOUBLE) :: rho = 1.0_C_DOUBLE
        INTEGER(C_INT) :: ibtype = 0, precision = 0
        REAL(C_DOUBLE) :: eps_rel = 0.0_C_DOUBLE
        INTEGER(C_INT) :: sor_all_levels = 0, check_every = 4
        ! SIP alfa; 0 = the reference's hard-coded default-real 0.92
        REAL(C_DOUBLE) :: alfa = 0.0_C_DOUBLE
        ! 1: mgpoisl stops after ftoc, the caller runs parent, bound_flow
        ! and connect of U, V, W, P itself
        INTEGER(C_INT) :: sync_by_caller = 0
    END TYPE mgp_config_t

    TYPE, BIND(C), PUBLIC :: mgp_stats_t
        INTEGER(C_INT) :: ipc = 0, nouter_done = 0, converged = 0
        REAL(C_DOUBLE) :: maxrhs0 = 0.0_C_DOUBLE, maxrhs = 0.0_C_DOUBLE
        REAL(C_DOUBLE) :: solve_ms = 0.0_C_DOUBLE
    END TYPE mgp_stats_t

    PUBLIC :: mgp_init, mgp_init_mask, mgp_mgpoisl, mgp_mgpoisl_device, &
        mgp_finish, mgp_last_error, mgp_set_iterations, mgp_error_message, &
        mgp_set_sdiv, mgp_connect_device, mgp_tstle4_device, mgp_rkstep_device, &
        mgp_bound_layout, mgp_set_bound, mgp_flow_bc_device, mgp_par_ftoc_device

    INTERFACE
        INTEGER(C_INT) FUNCTION mgp_init(grids, cfg, dx, dy, dz, ddx, ddy, &
                ddz, bp, handle) BIND(C, name="mgp_init")
            IMPORT :: C_INT, C_PTR, mgp_grids_t, mgp_config_t
            TYPE(mgp_grids_t), INTENT(in) :: grids
            TYPE(mgp_config_t), INTENT(in) :: cfg
            TYPE(C_PTR), VALUE :: dx, dy, dz, ddx, ddy, ddz, bp
            TYPE(C_PTR), INTENT(out) :: handle
        END FUNCTION mgp_init

        INTEGER(C_INT) FUNCTION mgp_init_mask(grids, cfg, dx, dy, dz, ddx, &
                ddy, ddz, blocked, handle) BIND(C, name="mgp_init_mask")
            IMPORT :: C_INT, C_PTR, mgp_grids_t, mgp_config_t
            TYPE(mgp_grids_t), INTENT(in) :: grids
            TYPE(mgp_config_t), INTENT(in) :: cfg
            TYPE(C_PTR), VALUE :: dx, dy, dz, ddx, ddy, ddz, blocked
            TYPE(C_PTR), INTENT(out) :: handle
        END FUNCTION mgp_init_mask

        INTEGER(C_INT) FUNCTION mgp_mgpoisl(handle, u, v, w, p, dt, ittot, &
                irk, stats) BIND(C, name="mgp_mgpoisl")
            IMPORT :: C_INT, C_PTR, C_DOUBLE, mgp_stats_t
            TYPE(C_PTR), VALUE :: handle, u, v, w, p
            REAL(C_DOUBLE), VALUE :: dt
            INTEGER(C_INT), VALUE :: ittot, irk
            TYPE(mgp_stats_t), INTENT(out) :: stats
        END FUNCTION mgp_mgpoisl

        INTEGER(C_INT) FUNCTION mgp_mgpoisl_device(handle, u, v, w, p, dt, &
                ittot, irk, stats) BIND(C, name="mgp_mgpoisl_device")
            IMPORT :: C_INT, C_PTR, C_DOUBLE, mgp_stats_t
            TYPE(C_PTR), VALUE :: handle, u, v, w, p
            REAL(C_DOUBLE), VALUE :: dt
            INTEGER(C_INT), VALUE :: ittot, irk
            TYPE(mgp_stats_t), INTENT(out) :: stats
        END FUNCTION mgp_mgpoisl_device

        INTEGER(C_INT) FUNCTION mgp_finish(handle) BIND(C, name="mgp_finish")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle
        END FUNCTION mgp_finish

        TYPE(C_PTR) FUNCTION mgp_last_error() BIND(C, name="mgp_last_error")
            IMPORT :: C_PTR
        END FUNCTION mgp_last_error

        INTEGER(C_INT) FUNCTION mgp_set_iterations(handle, ninner, nouter, &
                nouter_min, epcorr, eps_rel, check_every) &
                BIND(C, name="mgp_set_iterations")
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: handle
            INTEGER(C_INT), VALUE :: ninner, nouter, nouter_min, check_every
            REAL(C_DOUBLE), VALUE :: epcorr, eps_rel
        END FUNCTION mgp_set_iterations

        INTEGER(C_INT) FUNCTION mgp_set_sdiv(handle, sdiv) &
                BIND(C, name="mgp_set_sdiv")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle, sdiv
        END FUNCTION mgp_set_sdiv

        ! offsets: C_LOC of INTEGER(C_INT64_T) (6*ngrid) or C_NULL_PTR
        INTEGER(C_INT) FUNCTION mgp_bound_layout(handle, offsets, n) &
                BIND(C, name="mgp_bound_layout")
            IMPORT :: C_INT, C_PTR, C_INT64_T
            TYPE(C_PTR), VALUE :: handle, offsets
            INTEGER(C_INT64_T), INTENT(out) :: n
        END FUNCTION mgp_bound_layout

        INTEGER(C_INT) FUNCTION mgp_set_bound(handle, ubuf, vbuf, wbuf, pinf) &
                BIND(C, name="mgp_set_bound")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle, ubuf, vbuf, wbuf, pinf
        END FUNCTION mgp_set_bound

        ! fields: C_LOC of an array of nfield device pointers
        INTEGER(C_INT) FUNCTION mgp_connect_device(handle, level, layers, &
                corners, forward, fields, nfield) &
                BIND(C, name="mgp_connect_device")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle, fields
            INTEGER(C_INT), VALUE :: level, layers, corners, forward, nfield
        END FUNCTION mgp_connect_device

        ! parent, then bound_flow%bound, of one level (device fields)
        INTEGER(C_INT) FUNCTION mgp_flow_bc_device(handle, level, u, v, w, p) &
                BIND(C, name="mgp_flow_bc_device")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle, u, v, w, p
            INTEGER(C_INT), VALUE :: level
        END FUNCTION mgp_flow_bc_device

        ! par_ftoc_norm of one level (device fields); sum /= 0: plain sums
        INTEGER(C_INT) FUNCTION mgp_par_ftoc_device(handle, level, u, v, w, &
                sum) BIND(C, name="mgp_par_ftoc_device")
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: handle, u, v, w
            INTEGER(C_INT), VALUE :: level, sum
        END FUNCTION mgp_par_ftoc_device

        ! tstle4 (noib) on device fields; gradp may be C_NULL_PTR
        INTEGER(C_INT) FUNCTION mgp_tstle4_device(handle, uo, vo, wo, u, v, &
                w, p, g, gmol, ilesmodel, gradp) &
                BIND(C, name="mgp_tstle4_device")
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: handle, uo, vo, wo, u, v, w, p, g, gradp
            REAL(C_DOUBLE), VALUE :: gmol
            INTEGER(C_INT), VALUE :: ilesmodel
        END FUNCTION mgp_tstle4_device

        INTEGER(C_INT) FUNCTION mgp_rkstep_device(handle, p, dp, rhsp, frhs, &
                dtfu) BIND(C, name="mgp_rkstep_device")
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: handle, p, dp, rhsp
            REAL(C_DOUBLE), VALUE :: frhs, dtfu
        END FUNCTION mgp_rkstep_device
    END INTERFACE

CONTAINS

    ! Text of mgp_last_error() as a Fortran string
    FUNCTION mgp_error_message() RESULT(msg)
        CHARACTER(len=512) :: msg
        CHARACTER(kind=C_CHAR), POINTER :: cs(:)
        INTEGER :: n
        TYPE(C_PTR) :: p

        msg = ""
        p = mgp_last_error()
        IF (.NOT. C_ASSOCIATED(p)) RETURN
        CALL C_F_POINTER(p, cs, [512])
        DO n = 1, 512
            IF (cs(n) == C_NULL_CHAR) EXIT
            msg(n:n) = cs(n)
        END DO
    END FUNCTION mgp_error_message
END MODULE mgp_bindc_mod
