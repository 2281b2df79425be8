!-----------------------------------------------------------------------
! FASOLT_GPU: the FASOLT3 increment driver on the MI355X library.
!
! Keeps the reference's inputs and call sequence (gabortimar/CPFEM_code
! src/fasolt/Fasolt3.f90:11-68): dirname and fafcdf.dat (FSINIT,
! Fasolt3.f90:188-221), the .fsd input file named on stdin (RDDATA,
! Faslt_io_nowin.f90:9-146), nsave / isave from stdin, then per increment
! UPDATE_NODE, NRITER, UPDATE_GPTV and DoReorient, writing <name>-<inc>.frs,
! .slp, fafrep.txt and twins exactly as the reference does.  The hot path
! runs on the GPU through the C ABI (module CPFEM_CAPI); the state stays in
! HBM between calls and only the report scalars come back per iteration.
!
! Solver settings that the reference compiles in (SOR omega 1.2, 1000
! sweeps, stop on the relative update; Stif_sol.f90:173-210) can be changed
! in an optional file cpfem.dat in the same directory as fafcdf.dat:
!   omega  sor_kind(0 update, 1 residual)  sor_tol  sor_maxit  precision(0 FP64, 1 mixed)
! omega <= 0 lets the library choose omega and the sweep launch shape on the
! first system it solves (cpfem_autotune).
!
! Calibration (no counterpart in the reference): with an optional file
! calib.dat next to fafcdf.dat holding the five parameters (tau0 taus h0 q m),
! the stress component (1..6) and one target value per increment, every
! increment's constitutive inputs are recorded (cpfem_calib_record) and after
! the run the replay loss sum (curve - target)^2 and its adjoint gradient are
! written to calib.txt (points that do not converge excluded).
!-----------------------------------------------------------------------
MODULE HOST_STATE
    USE ISO_C_BINDING
    USE CPFEM_CAPI
    IMPLICIT NONE
    INTEGER, PARAMETER :: maxsvs = 63, maxss = 60, maxmds = 12
    INTEGER(C_INT32_T) :: nlmnt, nnod, ngps, nsv, ninc, ninc1, nxs, nfs, nmatl, nmpar, maxit, inc
    INTEGER(C_INT32_T) :: last_tflag = 0
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: le(:), lt(:), nd(:), np(:), ln(:,:), lg(:,:), ib(:)
    INTEGER(C_INT32_T), ALLOCATABLE :: twinsys(:), timestwinned(:)
    ! the reference's REAL(4) module arrays (FAS_COM / FAS_CPL, fas-com-mod.f90:13-80),
    ! passed to the library as they are (cpfem_create_r4, cpfem_set_state_r4, cpfem_get_state_r4)
    REAL(4), ALLOCATABLE, TARGET :: x(:), dx(:), fc(:), eps(:), dep(:), sv(:,:), dsv(:,:)
    REAL(4), ALLOCATABLE :: ep(:,:), de(:,:), st(:,:), ds(:,:), rt(:,:), dr(:,:)
    REAL(4), ALLOCATABLE :: srate(:,:), cumslip(:,:)
    REAL(4), ALLOCATABLE, TARGET :: tiem(:), x_spec(:,:,:), f_spec(:,:,:)
    INTEGER(4), TARGET :: nss(maxmds)
    REAL(4), TARGET :: e(9, maxmds), ssang(maxss, maxmds), ab(3, maxss, maxmds), an(3, maxss, maxmds)
    REAL(4), TARGET :: ampar(10, maxss, maxmds), smatrx(maxss, maxss, maxmds)
    REAL(4) :: amu, anu, aga, twinprob
    REAL(4) :: toler, alpha, afactor, efix, smalp, smalf, smald, smale, smalr
    INTEGER(4) :: twinlim
    TYPE(cpfem_control) :: ctl
    TYPE(C_PTR) :: ctx = C_NULL_PTR
    CHARACTER(len=100) :: dir = './'
    INTEGER, PARAMETER :: iomon = 51, ichan = 52, ochan = 53
END MODULE HOST_STATE

PROGRAM FASOLT_GPU
    USE HOST_STATE
    IMPLICIT NONE
    CHARACTER(len=4) :: text
    CHARACTER(len=100) :: f_name
    INTEGER(4) :: nsave, icurr, isave(1:100), rc, ntw, ierr, ccomp, nexcl
    LOGICAL :: calib
    REAL(C_DOUBLE) :: cpar(5), closs, cgrad(5)
    REAL(C_DOUBLE), ALLOCATABLE :: ctarget(:), ccurve(:)
    TYPE(cpfem_mesh_r4) :: mesh
    TYPE(cpfem_materials_r4) :: mtab

    CALL FSINIT
    CALL RDDATA(f_name)

    ! SETPTA (Fasolt3.f90:24) and the upload of the module arrays
    mesh%nlmnt = nlmnt; mesh%nnod = nnod; mesh%ngps = ngps; mesh%nsv = nsv
    mesh%le = C_LOC(le); mesh%lt = C_LOC(lt); mesh%nd = C_LOC(nd); mesh%np = C_LOC(np)
    mesh%ln = C_LOC(ln); mesh%lg = C_LOC(lg); mesh%ib = C_LOC(ib)
    mesh%ninc1 = ninc1; mesh%nxs = nxs; mesh%nfs = nfs
    mesh%tiem = C_LOC(tiem); mesh%x_spec = C_LOC(x_spec); mesh%f_spec = C_LOC(f_spec)
    mtab%nmatl = nmatl; mtab%twinlim = twinlim; mtab%twinprob = twinprob
    mtab%nss = C_LOC(nss); mtab%e = C_LOC(e); mtab%ssang = C_LOC(ssang)
    mtab%ab = C_LOC(ab); mtab%an = C_LOC(an); mtab%ampar = C_LOC(ampar); mtab%smatrx = C_LOC(smatrx)
    rc = cpfem_create_r4(mesh, mtab, ctl, ctx)
    IF (rc .NE. 0) CALL FAFERR(rc)
    rc = cpfem_set_state_r4(ctx, x, dx, eps, dep, sv, dsv, maxsvs, ep, de, st, ds, rt, dr, twinsys, timestwinned)
    IF (rc .NE. 0) CALL FAFERR(rc)

    WRITE(*, '(15X,''Number of increments is: '',I4)') ninc
    WRITE(*, '(/,10X,''Enter number of intermediate saves (max. 99):  '')', ADVANCE='NO')
    READ *, nsave
    IF (nsave .GT. 0) THEN
        WRITE(*, '(10X,''Enter intermediate increments to save at: '')', ADVANCE='NO')
        READ(*, *) (isave(icurr), icurr = 1, nsave)
    END IF

    ! optional calibration record
    calib = .FALSE.
    OPEN(UNIT=ichan, FILE=TRIM(dir)//'calib.dat', STATUS='OLD', IOSTAT=ierr)
    IF (ierr .EQ. 0) THEN
        ALLOCATE(ctarget(ninc), ccurve(ninc))
        READ(ichan, *) cpar
        READ(ichan, *) ccomp
        READ(ichan, *) ctarget
        CLOSE(ichan)
        rc = cpfem_calib_begin(ctx, ninc, 10, 1.0d-10, 60)
        calib = rc .EQ. 0
        IF (.NOT. calib) WRITE(*, '(10X,''calib.dat ignored: cpfem_calib_begin returned '',I4)') rc
    END IF

    OPEN(unit=55, file="twins", action="write", status="replace")
    DO inc = 1, ninc
        rc = cpfem_update_node(ctx)                 ! UPDATE_NODE
        IF (rc .NE. 0) CALL FAFERR(rc)
        CALL NRITER
        IF (calib) THEN
            rc = cpfem_calib_record(ctx, inc)
            IF (rc .NE. 0) CALL FAFERR(rc)
        END IF
        rc = cpfem_update_gptv(ctx)                 ! UPDATE_GPTV
        IF (rc .NE. 0) CALL FAFERR(rc)
        CALL REORIENT_LOG(ntw)                      ! DoReorient + twins log
        IF (nsave .GT. 0) THEN
            DO icurr = 1, nsave
                IF (inc .GE. isave(icurr)) THEN
                    WRITE(text, '(I4)') inc
                    CALL PULL_STATE
                    CALL WTDATA(TRIM(f_name)//'-'//TRIM(ADJUSTL(text))//'.frs')
                    CALL WRITESLIPS(TRIM(f_name)//'-'//TRIM(ADJUSTL(text))//'.slp')
                    isave(icurr) = ninc + 1
                END IF
            END DO
        END IF
    END DO
    CLOSE(55)
    WRITE(text, '(I4)') ninc
    CALL PULL_STATE
    CALL WTDATA(TRIM(f_name)//'-'//TRIM(ADJUSTL(text))//'.frs')
    CALL WRITESLIPS(TRIM(f_name)//'-'//TRIM(ADJUSTL(text))//'.slp')
    IF (calib) THEN
        rc = cpfem_calib_exclude(ctx, cpar, nexcl)
        IF (rc .NE. 0) CALL FAFERR(rc)
        rc = cpfem_calib_run(ctx, cpar, ctarget, ccomp - 1, ccurve, closs, cgrad, C_NULL_PTR)
        IF (rc .NE. 0) CALL FAFERR(rc)
        OPEN(UNIT=ochan, FILE=TRIM(dir)//'calib.txt', STATUS='REPLACE')
        WRITE(ochan, '(A,I8)') 'excluded points ', nexcl
        WRITE(ochan, '(A,ES24.16)') 'loss ', closs
        WRITE(ochan, '(A,5ES24.16)') 'gradient ', cgrad
        DO icurr = 1, ninc
            WRITE(ochan, '(I6,2ES24.16)') icurr, ccurve(icurr), ctarget(icurr)
        END DO
        CLOSE(ochan)
        rc = cpfem_calib_end(ctx)
    END IF
    rc = cpfem_destroy(ctx)
END PROGRAM FASOLT_GPU

!-----------------------------------------------------------------------
! FSINIT (Fasolt3.f90:188-221) plus the optional solver file cpfem.dat.
!-----------------------------------------------------------------------
SUBROUTINE FSINIT
    USE HOST_STATE
    IMPLICIT NONE
    CHARACTER(len=100) :: dir1, dir2, dir3, dir4
    INTEGER :: ierr
    REAL(8) :: om, tol
    INTEGER :: kind, mx, prec
    OPEN(ichan, FILE='dirname', STATUS='OLD', IOSTAT=ierr)
    IF (ierr .EQ. 0) THEN
        READ(ichan, *) dir1, dir2, dir3, dir4, dir
    ELSE
        WRITE(*, '("No dirname file found. Path is current directory.")')
        dir = './'
    END IF
    CLOSE(ichan)
    OPEN(UNIT=ichan, FILE=TRIM(dir)//'fafcdf.dat', STATUS='OLD')
    READ(ichan, *) toler, maxit, alpha, afactor
    READ(ichan, *) efix, smalp, smalf
    READ(ichan, *) smald, smale, smalr
    CLOSE(ichan)
    ctl%toler = toler; ctl%maxit = maxit; ctl%alpha = alpha; ctl%afactor = afactor
    ctl%efix = efix; ctl%smalp = smalp; ctl%smalf = smalf
    ctl%smald = smald; ctl%smale = smale; ctl%smalr = smalr
    ! the reference's compiled-in solver settings (Stif_sol.f90:173-210)
    ctl%sor_omega = 1.2d0; ctl%sor_kind = 0; ctl%sor_tol = toler; ctl%sor_maxit = 1000; ctl%precision = 0
    ctl%stall_tol = 0.0d0
    OPEN(UNIT=ichan, FILE=TRIM(dir)//'cpfem.dat', STATUS='OLD', IOSTAT=ierr)
    IF (ierr .EQ. 0) THEN
        READ(ichan, *) om, kind, tol, mx, prec
        ctl%sor_omega = om; ctl%sor_kind = kind; ctl%sor_tol = tol; ctl%sor_maxit = mx; ctl%precision = prec
        CLOSE(ichan)
    END IF
END SUBROUTINE FSINIT

!-----------------------------------------------------------------------
! RDDATA (Faslt_io_nowin.f90:9-146): read <name>.fsd (REAL(4) records).
!-----------------------------------------------------------------------
SUBROUTINE RDDATA(f_name)
    USE HOST_STATE
    IMPLICIT NONE
    CHARACTER(len=*) :: f_name
    CHARACTER(len=100) :: dir1, dir2, dir3, fpath
    INTEGER :: ierr, i, j, k, m, nde, npe, n
    INTEGER(4) :: i4a, i4b, i4c, i4d
    REAL(4) :: r6(6), r3(3), r15(15), rbuf(maxsvs)
    LOGICAL :: here
    dir3 = ''
    OPEN(ichan, FILE='dirname', STATUS='OLD', IOSTAT=ierr)
    IF (ierr .EQ. 0) THEN
        READ(ichan, *) dir1, dir2, dir3
    END IF
    CLOSE(ichan)
    WRITE(*, '(''Name of FASOLT input file (no extension): '')', ADVANCE='NO')
    DO i = 1, 3
        READ(*, *) f_name
        fpath = TRIM(dir3)//TRIM(f_name)//'.fsd'
        INQUIRE(FILE=fpath, EXIST=here)
        IF (here) EXIT
        IF (i .LT. 3) THEN
            WRITE(*, '("File not found. Try again.")')
        ELSE
            WRITE(*, '("File not found. Exiting.")')
            STOP
        END IF
    END DO
    OPEN(ichan, FILE=fpath, STATUS='OLD', FORM='UNFORMATTED', ACCESS='SEQUENTIAL', IOSTAT=ierr)
    IF (ierr .NE. 0) CALL FAFERR(501)
    READ(ichan) ninc
    READ(ichan) nlmnt, nnod, ngps, nsv
    ALLOCATE(le(nlmnt), lt(nlmnt), nd(nlmnt), np(nlmnt), ln(20, nlmnt), lg(8, nlmnt), ib(nnod))
    ln = 0; lg = 0
    ALLOCATE(x(3*nnod), dx(3*nnod), fc(3*nnod))
    ALLOCATE(eps(ngps), dep(ngps), sv(maxsvs, ngps), dsv(maxsvs, ngps))
    ALLOCATE(ep(6, ngps), de(6, ngps), st(6, ngps), ds(6, ngps), rt(3, ngps), dr(3, ngps))
    ALLOCATE(srate(maxss, ngps), cumslip(maxss, ngps), twinsys(ngps), timestwinned(ngps))
    sv = 0.0; dsv = 0.0; srate = 0.0; cumslip = 0.0; twinsys = 0; timestwinned = 0
    DO i = 1, nlmnt
        READ(ichan) i4a, i4b, nde, npe
        le(i) = i4a; lt(i) = i4b; nd(i) = nde; np(i) = npe
        READ(ichan) (ln(j, i), j = 1, nde)
        READ(ichan) (lg(j, i), j = 1, npe)
    END DO
    DO i = 1, nnod
        READ(ichan) i4a
        ib(i) = i4a
        READ(ichan) r6
        x(3*i-2:3*i) = r6(1:3); dx(3*i-2:3*i) = r6(4:6)
        READ(ichan) r3
        fc(3*i-2:3*i) = r3
    END DO
    DO i = 1, ngps
        READ(ichan) r6(1:2)
        eps(i) = r6(1); dep(i) = r6(2)
        READ(ichan) rbuf(1:nsv)
        sv(1:nsv, i) = rbuf(1:nsv)
        READ(ichan) rbuf(1:nsv)
        dsv(1:nsv, i) = rbuf(1:nsv)
        READ(ichan) r6; ep(:, i) = r6
        READ(ichan) r6; de(:, i) = r6
        READ(ichan) r6; st(:, i) = r6
        READ(ichan) r6; ds(:, i) = r6
        READ(ichan) r3; rt(:, i) = r3
        READ(ichan) r3; dr(:, i) = r3
    END DO
    READ(ichan) nmatl, nmpar
    IF (nmatl .GT. 12) CALL FAFERR(601)
    nss = 0; e = 0.0; ssang = 0.0; ab = 0.0; an = 0.0; ampar = 0.0; smatrx = 0.0
    DO m = 1, nmatl
        READ(ichan) n
        IF (n .GT. maxss) CALL FAFERR(602)
        IF (n .GT. maxsvs - 3) CALL FAFERR(603)
        nss(m) = n
        READ(ichan) twinlim          ! module scalars: the last set's values hold (Faslt_io_nowin.f90:83-86)
        READ(ichan) twinprob
        READ(ichan) e(:, m)
        READ(ichan) ssang(1:n, m)
        DO i = 1, n
            READ(ichan) r6
            ab(:, i, m) = r6(1:3); an(:, i, m) = r6(4:6)
        END DO
        DO i = 1, n
            READ(ichan) ampar(1:nmpar, i, m)
        END DO
        DO i = 1, n
            READ(ichan) smatrx(1:n, i, m)
        END DO
    END DO
    READ(ichan) amu, anu, aga
    READ(ichan) ninc1, nxs, nfs
    ALLOCATE(tiem(ninc1), x_spec(15, 99, ninc1), f_spec(3, 99, ninc1))
    x_spec = 0.0; f_spec = 0.0
    DO i = 1, ninc1
        READ(ichan) r3(1)
        tiem(i) = r3(1)
        DO j = 1, nxs
            READ(ichan) r15
            x_spec(:, j, i) = r15
        END DO
        DO j = 1, nfs
            READ(ichan) r3
            f_spec(:, j, i) = r3
        END DO
    END DO
    CLOSE(ichan)
    OPEN(UNIT=iomon, FILE='fafrep.txt')
    k = 0
END SUBROUTINE RDDATA

!-----------------------------------------------------------------------
! NRITER (Stif_sol.f90:10-88): the damped Newton loop, kernels on the GPU.
!-----------------------------------------------------------------------
SUBROUTINE NRITER
    USE HOST_STATE
    IMPLICIT NONE
    LOGICAL :: l_conv
    INTEGER(4) :: nit, rc
    INTEGER(C_INT32_T) :: tflag, sweeps
    REAL(C_DOUBLE) :: sumcpiter, sfc, sdx, alp2, err, fdx
    CALL ITHEAD(inc)
    nit = 0
    l_conv = .TRUE.
    DO WHILE (l_conv)
        nit = nit + 1
        rc = cpfem_elvals(ctx, inc, sumcpiter, tflag)       ! fc = 0, ss = 0, ELVALS
        IF (rc .NE. 0) CALL FAFERR(rc)
        rc = cpfem_suvals(ctx, inc, sfc)                    ! SUVALS + free-node force norm
        IF (rc .NE. 0) CALL FAFERR(rc)
        rc = cpfem_facsls(ctx, sweeps, err)                 ! FACSLS
        IF (rc .NE. 0) CALL FAFERR(rc)
        rc = cpfem_nr_update(ctx, sdx, alp2)                ! damping and update
        IF (rc .NE. 0) CALL FAFERR(rc)
        sumcpiter = sumcpiter / REAL(ngps, 8)
        CALL ITEREP(nit, REAL(sdx, 4), REAL(sfc, 4), REAL(sumcpiter, 4), REAL(alp2, 4))
        l_conv = (sdx .GE. toler) .AND. (nit .LT. maxit)
        last_tflag = tflag
    END DO
    rc = cpfem_cumslip(ctx, inc)
    IF (rc .NE. 0) CALL FAFERR(rc)
    rc = cpfem_power(ctx, fdx)
    WRITE(iomon, '(/,5X,''Power: '',G12.5)') REAL(fdx, 4) / REAL(tiem(ninc1), 4)
    WRITE(*, '(/,5X,''Power: '',G12.5)') REAL(fdx, 4) / REAL(tiem(ninc1), 4)
END SUBROUTINE NRITER

SUBROUTINE ITHEAD(inc)
    USE HOST_STATE, ONLY: iomon
    IMPLICIT NONE
    INTEGER(4) :: inc
    WRITE(iomon, '(//,25X,''Increment number'',I4,/,10X,''Iteration'',2X,''Error(X)'',2X,''Error(F)'',2X,''Err(CPI)'',2X,''Decelrn.'',/)') inc
    WRITE(*, '(//,25X,''Increment number'',I4,/,10X,''Iteration'',2X,''Error(X)'',2X,''Error(F)'',2X,''Err(CPI)'',2X,''Decelrn.'',/)') inc
END SUBROUTINE ITHEAD

SUBROUTINE ITEREP(nit, sdx, sfc, cper, alp2)
    USE HOST_STATE, ONLY: iomon
    IMPLICIT NONE
    INTEGER(4) :: nit
    REAL(4) :: sdx, sfc, cper, alp2
    WRITE(iomon, '(12X,I3,4X,5(1X,E9.2))') nit, sdx, sfc, cper, alp2
    WRITE(*, '(12X,I3,4X,5(1X,E9.2))') nit, sdx, sfc, cper, alp2
END SUBROUTINE ITEREP

!-----------------------------------------------------------------------
! DoReorient (Fasolt3.f90:79-178) on the GPU, with the twins log lines.
!-----------------------------------------------------------------------
SUBROUTINE REORIENT_LOG(ntw)
    USE HOST_STATE
    IMPLICIT NONE
    INTEGER(4) :: ntw, rc, j
    INTEGER(C_INT32_T), ALLOCATABLE :: tws(:)
    REAL(4), ALLOCATABLE :: e0(:,:)
    ntw = 0
    ! tflag of the last pass (Fasolt3.f90:43): pull the state only when twins are pending
    IF (last_tflag .EQ. 0) RETURN
    CALL PULL_STATE
    ALLOCATE(tws(ngps), e0(3, ngps))
    tws = twinsys
    e0 = sv(1:3, :)
    rc = cpfem_reorient(ctx, ntw)
    IF (rc .NE. 0) CALL FAFERR(rc)
    CALL PULL_STATE
    DO j = 1, ngps
        IF (tws(j) .NE. 0) WRITE(55, '(I8, I8, I8, 9999ES15.5)') inc, j, tws(j), e0(:, j), sv(1:3, j)
    END DO
    DEALLOCATE(tws, e0)
    last_tflag = 0
END SUBROUTINE REORIENT_LOG

SUBROUTINE PULL_STATE
    USE HOST_STATE
    IMPLICIT NONE
    INTEGER(4) :: rc
    rc = cpfem_get_state_r4(ctx, x, dx, fc, eps, dep, sv, dsv, maxsvs, ep, de, st, ds, rt, dr, srate, cumslip, &
                            twinsys, timestwinned)
    IF (rc .NE. 0) CALL FAFERR(rc)
END SUBROUTINE PULL_STATE

!-----------------------------------------------------------------------
! WTDATA (Faslt_io_nowin.f90:185-299): <name>-<inc>.frs, REAL(4) records.
!-----------------------------------------------------------------------
SUBROUTINE WTDATA(f_name)
    USE HOST_STATE
    IMPLICIT NONE
    CHARACTER(len=*) :: f_name
    INTEGER :: i, j, ierr, m, n
    OPEN(UNIT=ochan, FILE=f_name, FORM='UNFORMATTED', ACCESS='SEQUENTIAL', STATUS='UNKNOWN', IOSTAT=ierr)
    WRITE(*, *) f_name
    IF (ierr .NE. 0) CALL FAFERR(502)
    WRITE(ochan) inc
    WRITE(ochan) nlmnt, nnod, ngps, nsv
    DO i = 1, nlmnt
        WRITE(ochan) le(i), lt(i), nd(i), np(i)
        WRITE(ochan) (ln(j, i), j = 1, nd(i))
        WRITE(ochan) (lg(j, i), j = 1, np(i))
    END DO
    DO i = 1, nnod
        WRITE(ochan) ib(i)
        WRITE(ochan) x(3*i-2:3*i), dx(3*i-2:3*i)
        WRITE(ochan) fc(3*i-2:3*i)
    END DO
    DO i = 1, ngps
        WRITE(ochan) eps(i), dep(i)
        WRITE(ochan) sv(1:nsv, i)
        WRITE(ochan) dsv(1:nsv, i)
        WRITE(ochan) ep(:, i)
        WRITE(ochan) de(:, i)
        WRITE(ochan) st(:, i)
        WRITE(ochan) ds(:, i)
        WRITE(ochan) rt(:, i)
        WRITE(ochan) dr(:, i)
    END DO
    WRITE(ochan) nmatl, nmpar
    DO m = 1, nmatl
        n = nss(m)
        WRITE(ochan) n
        WRITE(ochan) twinlim
        WRITE(ochan) twinprob
        WRITE(ochan) e(:, m)
        WRITE(ochan) ssang(1:n, m)
        DO i = 1, n
            WRITE(ochan) ab(:, i, m), an(:, i, m)
        END DO
        DO i = 1, n
            WRITE(ochan) ampar(1:nmpar, i, m)
        END DO
        DO i = 1, n
            WRITE(ochan) smatrx(1:n, i, m)
        END DO
    END DO
    WRITE(ochan) amu, anu, aga
    WRITE(ochan) ninc1, nxs, nfs
    DO i = 1, ninc1
        WRITE(ochan) tiem(i)
        DO j = 1, nxs
            WRITE(ochan) x_spec(:, j, i)
        END DO
        DO j = 1, nfs
            WRITE(ochan) f_spec(:, j, i)
        END DO
    END DO
    CLOSE(ochan)
END SUBROUTINE WTDATA

! WriteSlips (Faslt_io_nowin.f90:156-175): cumulative slip and slip rates per point.
SUBROUTINE WRITESLIPS(f_name)
    USE HOST_STATE
    IMPLICIT NONE
    CHARACTER(len=*) :: f_name
    INTEGER :: i, j, n
    n = nss(1)
    WRITE(*, *) f_name
    OPEN(unit=54, file=f_name, action="write", status="replace")
    DO i = 1, ngps
        WRITE(54, '(9999ES15.5)') (cumslip(j, i), j = 1, n), (srate(j, i), j = 1, n)
    END DO
    CLOSE(54)
END SUBROUTINE WRITESLIPS

!-----------------------------------------------------------------------
! FAFERR (Fasolt3.f90:299-323): warn at <= 500, stop above, dump above 900.
!-----------------------------------------------------------------------
SUBROUTINE FAFERR(ier)
    USE HOST_STATE
    IMPLICIT NONE
    INTEGER(4) :: ier, ierr
    IF (ier .LT. 0) THEN
        WRITE(*, '(/,10X,'' *** FASOLT_GPU: library error '',I4,/)') ier
        STOP 3
    END IF
    IF (ier .LE. 500) THEN
        WRITE(iomon, '(/,10X,'' *** FAFNER:  WARNING, '',I3,/)') ier
        WRITE(*, '(/,10X,'' *** FAFNER:  WARNING, '',I3,/)') ier
    ELSE
        WRITE(iomon, '(/,10X,'' *** FAFNER:  FATAL ERROR, '',I3,/)') ier
        WRITE(*, '(/,10X,'' *** FAFNER:  FATAL ERROR, '',I3,/)') ier
        IF (ier .GT. 900 .AND. C_ASSOCIATED(ctx)) THEN
            CALL PULL_STATE
            CALL WTDATA('faferr.frs')
        END IF
        WRITE(*, '(10x,''any key to finish..'')', ADVANCE='NO')
        READ(*, *, IOSTAT=ierr)
        STOP 1
    END IF
END SUBROUTINE FAFERR
