! mgp_driver -- a Fortran time-loop stand-in that calls the projection
! through mgp_pressuresolver_mod, exactly as MGLET's timeintegrate_flow
! calls mgpoisl (src/flow/timeintegration_mod.F90:127).
!
! Usage: mgp_driver <case.bin> <out.bin>
! case.bin (stream, native endianness): ngrid minlevel maxlevel n3d n1x n1y
! n1z (int64), kk jj ii level (int32 x ngrid), nbrgrid (int32 x 26 ngrid),
! itypbc (int32 x 6 ngrid), ip3 ip1x ip1y ip1z (int64 x ngrid, 1-based),
! dx dy dz ddx ddy ddz (float64), bp u v w p (float64 x n3d),
! ityp ninner nouter nouter_min ibtype precision (int32), epcorr omega rho
! eps_rel dt (float64).
! out.bin: u v w p (float64 x n3d), ipc nouter_done converged (int32),
! maxrhs0 maxrhs (float64).
PROGRAM mgp_driver
    USE, INTRINSIC :: ISO_C_BINDING
    USE mgp_pressuresolver_mod
    IMPLICIT NONE

    CHARACTER(len=4096) :: fin, fout
    INTEGER(C_INT64_T) :: hdr(7)
    INTEGER(C_INT) :: ngrid, minlevel, maxlevel
    INTEGER(C_INT), ALLOCATABLE :: kk(:), jj(:), ii(:), level(:)
    INTEGER(C_INT), ALLOCATABLE :: nbrgrid(:, :), itypbc(:, :)
    INTEGER(C_INT64_T), ALLOCATABLE :: ip3(:), ip1x(:), ip1y(:), ip1z(:)
    REAL(C_DOUBLE), ALLOCATABLE :: dx(:), dy(:), dz(:), ddx(:), ddy(:), ddz(:)
    REAL(C_DOUBLE), ALLOCATABLE :: bp(:), u(:), v(:), w(:), p(:)
    INTEGER(C_INT) :: ip(6)
    REAL(C_DOUBLE) :: rp(5)
    INTEGER :: lu

    CALL GET_COMMAND_ARGUMENT(1, fin)
    CALL GET_COMMAND_ARGUMENT(2, fout)

    OPEN(newunit=lu, file=TRIM(fin), access="stream", form="unformatted", &
        status="old", action="read")
    READ(lu) hdr
    ngrid = INT(hdr(1), C_INT)
    minlevel = INT(hdr(2), C_INT)
    maxlevel = INT(hdr(3), C_INT)
    ALLOCATE(kk(ngrid), jj(ngrid), ii(ngrid), level(ngrid))
    ALLOCATE(nbrgrid(26, ngrid), itypbc(6, ngrid))
    ALLOCATE(ip3(ngrid), ip1x(ngrid), ip1y(ngrid), ip1z(ngrid))
    ALLOCATE(dx(hdr(5)), ddx(hdr(5)), dy(hdr(6)), ddy(hdr(6)), dz(hdr(7)), &
        ddz(hdr(7)))
    ALLOCATE(bp(hdr(4)), u(hdr(4)), v(hdr(4)), w(hdr(4)), p(hdr(4)))
    READ(lu) kk, jj, ii, level, nbrgrid, itypbc, ip3, ip1x, ip1y, ip1z
    READ(lu) dx, dy, dz, ddx, ddy, ddz, bp, u, v, w, p
    READ(lu) ip, rp
    CLOSE(lu)

    CALL init_pressuresolver(ngrid, minlevel, maxlevel, kk, jj, ii, level, &
        nbrgrid, itypbc, ip3, ip1x, ip1y, ip1z, dx, dy, dz, ddx, ddy, ddz, &
        bp, ityp=ip(1), ninner=ip(2), nouter=ip(3), nouter_min=ip(4), &
        epcorr=rp(1), omega=rp(2), rho=rp(3), ibtype=ip(5), &
        precision=ip(6), eps_rel=rp(4))
    CALL mgpoisl(u, v, w, p, rp(5), 1_C_INT, 1_C_INT)
    CALL finish_pressuresolver()

    OPEN(newunit=lu, file=TRIM(fout), access="stream", form="unformatted", &
        status="replace", action="write")
    WRITE(lu) u, v, w, p
    WRITE(lu) last_stats%ipc, last_stats%nouter_done, last_stats%converged
    WRITE(lu) last_stats%maxrhs0, last_stats%maxrhs
    CLOSE(lu)
    WRITE(*, '("mgp_driver: ipc=", I0, " maxrhs=", ES12.5)') last_stats%ipc, &
        last_stats%maxrhs
END PROGRAM mgp_driver
