!-----------------------------------------------------------------------
! Test harness around the compiled reference solver (gabortimar/CPFEM_code,
! src/fasolt/).  It is linked against the reference's own objects, built by
! oracle/Makefile into oracle/_ref/, and only ever used as a checker by the
! test-suite and the fixture scripts under tests/golden/.
!
! It loads the reference modules FAS_COM / FAS_CPL from a raw little-endian
! stream file written by tests/golden/refio.py, calls the reference routines
! and writes their outputs as a raw stream file.
!
!   ref_harness dstrss <in> <out>   DSTRSS per point (Sig_eps.f90:84)
!   ref_harness nrstep <in> <out>   fc=0, ss=0, ELVALS, SUVALS, then FACSLS
!                                   (the body of one NRITER pass,
!                                   Stif_sol.f90:38-53)
!   ref_harness incs   <in> <out>   ninc increments of the reference loop:
!                                   x+=dx, NRITER, commit (Fasolt3.f90:38-46)
!
! FAFERR (Fasolt3.f90:299) lives in the reference's main-program file, which
! cannot be linked beside this main program; the harness supplies a routine
! of that name that records the largest code instead of stopping.
!-----------------------------------------------------------------------
SUBROUTINE FAFERR(ier)
    IMPLICIT NONE
    INTEGER(4) ier
    INTEGER(4) maxerr
    COMMON /HARNERR/ maxerr
    IF (ier .GT. maxerr) maxerr = ier
END SUBROUTINE FAFERR

MODULE HARN_IO
    USE FAS_COM
    USE FAS_CPL
    IMPLICIT NONE
CONTAINS
    SUBROUTINE read_material(u)
        INTEGER, INTENT(IN) :: u
        INTEGER(4) :: n, m
        READ(u) nmatl, nmpar
        e = 0.; nss = 0; ab = 0.; an = 0.; ampar = 0.; smatrx = 0.; ssang = 0.
        DO m = 1, nmatl
            READ(u) nss(m), twinlim, twinprob
            n = nss(m)
            READ(u) e(1:9, m)
            READ(u) ssang(1:n, m)
            READ(u) ab(1:3, 1:n, m)
            READ(u) an(1:3, 1:n, m)
            READ(u) ampar(1:10, 1:n, m)
            READ(u) smatrx(1:n, 1:n, m)
        END DO
    END SUBROUTINE read_material

    SUBROUTINE read_control(u)
        INTEGER, INTENT(IN) :: u
        READ(u) toler, maxit, alpha, afactor
        READ(u) efix, smalp, smalf
        READ(u) smald, smale, smalr
    END SUBROUTINE read_control

    SUBROUTINE read_mesh(u)
        INTEGER, INTENT(IN) :: u
        INTEGER(4) :: i
        READ(u) nlmnt, nnod, ngps, nsv
        DO i = 1, nlmnt
            READ(u) le(i), lt(i), nd(i), np(i)
            READ(u) ln(1:20, i)
            READ(u) lg(1:8, i)
        END DO
        READ(u) ib(1:nnod)
        READ(u) x(1:3*nnod), dx(1:3*nnod)
        READ(u) eps(1:ngps), dep(1:ngps)
        READ(u) sv(1:maxsvs, 1:ngps), dsv(1:maxsvs, 1:ngps)
        READ(u) ep(1:6, 1:ngps), de(1:6, 1:ngps), st(1:6, 1:ngps), ds(1:6, 1:ngps)
        READ(u) rt(1:3, 1:ngps), dr(1:3, 1:ngps)
        READ(u) twinsys(1:ngps), timestwinned(1:ngps)
        READ(u) ninc1, nxs, nfs
        DO i = 1, ninc1
            READ(u) tiem(i)
            READ(u) x_spec(1:15, 1:99, i)
            READ(u) f_spec(1:3, 1:99, i)
        END DO
        srate(:, 1:ngps) = 0.
        cumslip(:, 1:ngps) = 0.
    END SUBROUTINE read_mesh

    SUBROUTINE write_state(u)
        INTEGER, INTENT(IN) :: u
        WRITE(u) x(1:3*nnod), dx(1:3*nnod), fc(1:3*nnod)
        WRITE(u) eps(1:ngps), dep(1:ngps)
        WRITE(u) sv(1:maxsvs, 1:ngps), dsv(1:maxsvs, 1:ngps)
        WRITE(u) ep(1:6, 1:ngps), de(1:6, 1:ngps), st(1:6, 1:ngps), ds(1:6, 1:ngps)
        WRITE(u) rt(1:3, 1:ngps), dr(1:3, 1:ngps)
        WRITE(u) srate(1:maxss, 1:ngps), cumslip(1:maxss, 1:ngps)
        WRITE(u) twinsys(1:ngps), timestwinned(1:ngps), tflag
    END SUBROUTINE write_state
END MODULE HARN_IO

PROGRAM REF_HARNESS
    USE FAS_COM
    USE FAS_CPL
    USE HARN_IO
    IMPLICIT NONE
    CHARACTER(len=256) :: mode, fin, fout
    INTEGER(4) :: n, g, iel_g, i, k, ndof, nsteps, istep
    INTEGER(4), ALLOCATABLE :: iel(:), errs(:)
    REAL(4) :: deff, dtv
    REAL(4) :: des(6), rr(3), sts(6), dss(6), tt(6), chi(6, 6)
    INTEGER(4) :: maxerr
    COMMON /HARNERR/ maxerr
    REAL(4) :: sumcpiter
    COMMON /CPITER/ sumcpiter
    INTEGER(4) :: iomon, ichan, ochan
    COMMON /iochan/ iomon, ichan, ochan

    CALL GET_COMMAND_ARGUMENT(1, mode)
    CALL GET_COMMAND_ARGUMENT(2, fin)
    CALL GET_COMMAND_ARGUMENT(3, fout)
    iomon = 51; ichan = 52; ochan = 53
    OPEN(iomon, FILE=TRIM(fout)//'.rep', STATUS='REPLACE')
    OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
    OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    maxerr = 0
    tflag = 0

    SELECT CASE (TRIM(mode))
    CASE ('dstrss')
        CALL read_material(10)
        CALL read_control(10)
        READ(10) n, dtv
        inc = 1
        tiem(1) = dtv
        ALLOCATE(iel(n), errs(n))
        sumcpiter = 0.
        DO g = 1, n
            READ(10) iel(g)
            READ(10) de(1:6, g), dr(1:3, g), st(1:6, g), ds(1:6, g)
            READ(10) sv(1:maxsvs, g), dsv(1:maxsvs, g)
            READ(10) twinsys(g), timestwinned(g)
        END DO
        srate(:, 1:n) = 0.
        DO g = 1, n
            maxerr = 0
            des = de(1:6, g); rr = dr(1:3, g); sts = st(1:6, g); dss = ds(1:6, g)
            deff = 0.
            CALL DSTRSS(iel(g), g, deff, des, des, sts, dss, tt, rr, chi)
            errs(g) = maxerr
            WRITE(11) tt, dss, chi, deff
            WRITE(11) dsv(1:maxsvs, g), srate(1:maxss, g)
            WRITE(11) twinsys(g), errs(g)
        END DO
        WRITE(11) sumcpiter, tflag

    CASE ('nrstep')
        CALL read_material(10)
        CALL read_control(10)
        CALL read_mesh(10)
        READ(10) inc
        ndof = 3*nnod
        CALL SETPTA
        sumcpiter = 0.
        fc(1:ndof) = 0.
        ss(1:maxofd, 1:ndof) = 0.
        CALL ELVALS
        CALL SUVALS
        WRITE(11) maxerr, sumcpiter
        WRITE(11) fc(1:ndof), ddx(1:ndof)
        WRITE(11) is(1:81, 1:ndof)
        WRITE(11) ss(1:81, 1:ndof)
        CALL write_state(11)
        CALL FACSLS(ndof)
        WRITE(11) maxerr
        WRITE(11) ddx(1:ndof)

    CASE ('incs')
        CALL read_material(10)
        CALL read_control(10)
        CALL read_mesh(10)
        READ(10) nsteps
        ndof = 3*nnod
        CALL SETPTA
        DO istep = 1, nsteps
            inc = istep
            x(1:ndof) = x(1:ndof) + dx(1:ndof)              ! UPDATE_NODE
            CALL NRITER
            DO i = 1, ngps                                    ! UPDATE_GPTV
                eps(i) = eps(i) + dep(i)
                rt(1:3, i) = rt(1:3, i) + dr(1:3, i)
                ep(1:6, i) = ep(1:6, i) + de(1:6, i)
                st(1:6, i) = st(1:6, i) + ds(1:6, i)
                DO k = 1, nsv
                    sv(k, i) = sv(k, i) + dsv(k, i)
                END DO
            END DO
            WRITE(11) maxerr
            CALL write_state(11)
        END DO
    CASE DEFAULT
        PRINT *, 'unknown mode ', TRIM(mode)
        STOP 2
    END SELECT
    CLOSE(11)
    CLOSE(10)
    CLOSE(iomon)
END PROGRAM REF_HARNESS
