!-----------------------------------------------------------------------
! ISO_C_BINDING interface to libcpfem.so (include/cpfem.h).
!
! The derived types mirror the C structs field for field; the routines are
! the C entry points.  Every function returns 0 or a FAFERR code
! (Fasolt3.f90:299-323); negative values are library errors.
!-----------------------------------------------------------------------
MODULE CPFEM_CAPI
    USE ISO_C_BINDING
    IMPLICIT NONE

    INTEGER, PARAMETER :: CPFEM_MAXSS = 60

    TYPE, BIND(C) :: cpfem_material
        INTEGER(C_INT32_T) :: nss, twinlim
        REAL(C_DOUBLE) :: twinprob
        REAL(C_DOUBLE) :: e(9)
        REAL(C_DOUBLE) :: ssang(CPFEM_MAXSS)
        REAL(C_DOUBLE) :: ab(3, CPFEM_MAXSS), an(3, CPFEM_MAXSS)
        REAL(C_DOUBLE) :: ampar(10, CPFEM_MAXSS)
        REAL(C_DOUBLE) :: smatrx(CPFEM_MAXSS, CPFEM_MAXSS)   ! smatrx(j,i) as in the reference
    END TYPE cpfem_material

    TYPE, BIND(C) :: cpfem_control
        REAL(C_DOUBLE) :: toler
        INTEGER(C_INT32_T) :: maxit, sor_maxit
        REAL(C_DOUBLE) :: alpha, afactor
        REAL(C_DOUBLE) :: efix, smalp, smalf
        REAL(C_DOUBLE) :: smald, smale, smalr
        REAL(C_DOUBLE) :: sor_omega, sor_tol
        INTEGER(C_INT32_T) :: sor_kind, precision
        REAL(C_DOUBLE) :: stall_tol
    END TYPE cpfem_control

    TYPE, BIND(C) :: cpfem_mesh
        INTEGER(C_INT32_T) :: nlmnt, nnod, ngps, nsv
        TYPE(C_PTR) :: le, lt, nd, np, ln, lg, ib
        INTEGER(C_INT32_T) :: ninc1, nxs, nfs
        TYPE(C_PTR) :: tiem, x_spec, f_spec
    END TYPE cpfem_mesh

    ! REAL(4) module arrays of the reference (fas-com-mod.f90:13-80)
    TYPE, BIND(C) :: cpfem_mesh_r4
        INTEGER(C_INT32_T) :: nlmnt, nnod, ngps, nsv
        TYPE(C_PTR) :: le, lt, nd, np, ln, lg, ib
        INTEGER(C_INT32_T) :: ninc1, nxs, nfs
        TYPE(C_PTR) :: tiem, x_spec, f_spec          ! REAL(4)
    END TYPE cpfem_mesh_r4

    TYPE, BIND(C) :: cpfem_materials_r4
        INTEGER(C_INT32_T) :: nmatl, twinlim
        REAL(C_FLOAT) :: twinprob
        TYPE(C_PTR) :: nss                           ! nss(maxmds)
        TYPE(C_PTR) :: e, ssang, ab, an, ampar, smatrx   ! e(9,maxmds), ssang(60,maxmds), ab/an(3,60,maxmds),
                                                     ! ampar(10,60,maxmds), smatrx(60,60,maxmds)
    END TYPE cpfem_materials_r4

    TYPE, BIND(C) :: cpfem_stats
        INTEGER(C_INT32_T) :: nit, sor_sweeps, sor_sweeps_total, err
        REAL(C_DOUBLE) :: sdx, sfc, sumcpiter, alp2, sor_error
        REAL(C_DOUBLE) :: t_const_ms, t_asm_ms, t_suvals_ms, t_sor_ms, t_update_ms, t_total_ms
        INTEGER(C_INT64_T) :: flops_const, transc_const, gp_updates, pts_continued, pts_wide
        REAL(C_DOUBLE) :: t_sweep_ms
        INTEGER(C_INT64_T) :: sweep_launches
        REAL(C_DOUBLE) :: sor_omega, t_tune_ms
        INTEGER(C_INT32_T) :: sor_wpb, tune_sweeps
        INTEGER(C_INT64_T) :: pts_987
        INTEGER(C_INT32_T) :: pts_987_last, graph_launches
        REAL(C_DOUBLE) :: t_const_pass_ms(4)
        REAL(C_DOUBLE) :: t_graph_ms
    END TYPE cpfem_stats

    INTERFACE
        INTEGER(C_INT) FUNCTION cpfem_create(mesh, mats, nmat, ctl, ctx) BIND(C, name='cpfem_create')
            IMPORT :: C_INT, C_INT32_T, C_PTR, cpfem_mesh, cpfem_material, cpfem_control
            TYPE(cpfem_mesh), INTENT(IN) :: mesh
            TYPE(cpfem_material), INTENT(IN) :: mats(*)
            INTEGER(C_INT32_T), VALUE :: nmat
            TYPE(cpfem_control), INTENT(IN) :: ctl
            TYPE(C_PTR), INTENT(OUT) :: ctx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_create_r4(mesh, mats, ctl, ctx) BIND(C, name='cpfem_create_r4')
            IMPORT :: C_INT, C_PTR, cpfem_mesh_r4, cpfem_materials_r4, cpfem_control
            TYPE(cpfem_mesh_r4), INTENT(IN) :: mesh
            TYPE(cpfem_materials_r4), INTENT(IN) :: mats
            TYPE(cpfem_control), INTENT(IN) :: ctl
            TYPE(C_PTR), INTENT(OUT) :: ctx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_set_state_r4(ctx, x, dx, eps, dep, sv, dsv, ldsv, ep, de, st, ds, rt, dr, &
                                                   twinsys, timestwinned) BIND(C, name='cpfem_set_state_r4')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_FLOAT
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_FLOAT), INTENT(IN) :: x(*), dx(*), eps(*), dep(*), sv(*), dsv(*)
            INTEGER(C_INT32_T), VALUE :: ldsv
            REAL(C_FLOAT), INTENT(IN) :: ep(*), de(*), st(*), ds(*), rt(*), dr(*)
            INTEGER(C_INT32_T), INTENT(IN) :: twinsys(*), timestwinned(*)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_get_state_r4(ctx, x, dx, fc, eps, dep, sv, dsv, ldsv, ep, de, st, ds, rt, dr, &
                                                   srate, cumslip, twinsys, timestwinned) &
                                                   BIND(C, name='cpfem_get_state_r4')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_FLOAT
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_FLOAT), INTENT(INOUT) :: x(*), dx(*), fc(*), eps(*), dep(*), sv(*), dsv(*)
            INTEGER(C_INT32_T), VALUE :: ldsv
            REAL(C_FLOAT), INTENT(INOUT) :: ep(*), de(*), st(*), ds(*), rt(*), dr(*), srate(*), cumslip(*)
            INTEGER(C_INT32_T), INTENT(INOUT) :: twinsys(*), timestwinned(*)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_destroy(ctx) BIND(C, name='cpfem_destroy')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_set_state(ctx, x, dx, eps, dep, sv, dsv, ldsv, ep, de, st, ds, rt, dr, &
                                                twinsys, timestwinned) BIND(C, name='cpfem_set_state')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(IN) :: x(*), dx(*), eps(*), dep(*), sv(*), dsv(*)
            INTEGER(C_INT32_T), VALUE :: ldsv
            REAL(C_DOUBLE), INTENT(IN) :: ep(*), de(*), st(*), ds(*), rt(*), dr(*)
            INTEGER(C_INT32_T), INTENT(IN) :: twinsys(*), timestwinned(*)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_get_state(ctx, x, dx, fc, eps, dep, sv, dsv, ldsv, ep, de, st, ds, rt, dr, &
                                                srate, cumslip, twinsys, timestwinned) BIND(C, name='cpfem_get_state')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(OUT) :: x(*), dx(*), fc(*), eps(*), dep(*), sv(*), dsv(*)
            INTEGER(C_INT32_T), VALUE :: ldsv
            REAL(C_DOUBLE), INTENT(OUT) :: ep(*), de(*), st(*), ds(*), rt(*), dr(*), srate(*), cumslip(*)
            INTEGER(C_INT32_T), INTENT(OUT) :: twinsys(*), timestwinned(*)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_update_node(ctx) BIND(C, name='cpfem_update_node')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_elvals(ctx, inc, sumcpiter, tflag) BIND(C, name='cpfem_elvals')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: inc
            REAL(C_DOUBLE), INTENT(OUT) :: sumcpiter
            INTEGER(C_INT32_T), INTENT(OUT) :: tflag
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_suvals(ctx, inc, sfc) BIND(C, name='cpfem_suvals')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: inc
            REAL(C_DOUBLE), INTENT(OUT) :: sfc
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_facsls(ctx, sweeps, error) BIND(C, name='cpfem_facsls')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), INTENT(OUT) :: sweeps
            REAL(C_DOUBLE), INTENT(OUT) :: error
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_nr_update(ctx, sdx, alp2) BIND(C, name='cpfem_nr_update')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(OUT) :: sdx, alp2
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_update_gptv(ctx) BIND(C, name='cpfem_update_gptv')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_reorient(ctx, n) BIND(C, name='cpfem_reorient')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), INTENT(OUT) :: n
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_set_tflag(ctx, tflag) BIND(C, name='cpfem_set_tflag')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: tflag
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_volume_average(ctx, avg) BIND(C, name='cpfem_volume_average')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(OUT) :: avg(14)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_cumslip(ctx, inc) BIND(C, name='cpfem_cumslip')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: inc
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_power(ctx, fdx) BIND(C, name='cpfem_power')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(OUT) :: fdx
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_get_stats(ctx, st) BIND(C, name='cpfem_get_stats')
            IMPORT :: C_INT, C_PTR, cpfem_stats
            TYPE(C_PTR), VALUE :: ctx
            TYPE(cpfem_stats), INTENT(OUT) :: st
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_autotune(ctx, flags, omega, waves_per_block) BIND(C, name='cpfem_autotune')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: flags
            REAL(C_DOUBLE), INTENT(OUT) :: omega
            INTEGER(C_INT32_T), INTENT(OUT) :: waves_per_block
        END FUNCTION

        ! order of the node colours within a sweep (set /= 0 applies order(1:n))
        INTEGER(C_INT) FUNCTION cpfem_sweep_order(ctx, order, n, set) BIND(C, name='cpfem_sweep_order')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), INTENT(INOUT) :: order(*)
            INTEGER(C_INT32_T), VALUE :: n, set
        END FUNCTION

        ! calibration: record of the load path, replay loss and its adjoint gradient
        INTEGER(C_INT) FUNCTION cpfem_calib_begin(ctx, ninc, ckpt_every, toler, maxit) &
                BIND(C, name='cpfem_calib_begin')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: ninc, ckpt_every, maxit
            REAL(C_DOUBLE), VALUE :: toler
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_record(ctx, inc) BIND(C, name='cpfem_calib_record')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: inc
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_exclude(ctx, params, excluded) BIND(C, name='cpfem_calib_exclude')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(IN) :: params(5)
            INTEGER(C_INT32_T), INTENT(OUT) :: excluded
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_run(ctx, params, target, comp, curve, loss, grad, stats) &
                BIND(C, name='cpfem_calib_run')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(IN) :: params(5), target(*)
            INTEGER(C_INT32_T), VALUE :: comp
            REAL(C_DOUBLE), INTENT(OUT) :: curve(*), loss, grad(5)
            TYPE(C_PTR), VALUE :: stats
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_options(ctx, flags) BIND(C, name='cpfem_calib_options')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: flags        ! 1: CPFEM_CALIB_EQUILIBRIUM
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_checkpoints(ctx, ckpt_every) BIND(C, name='cpfem_calib_checkpoints')
            IMPORT :: C_INT, C_INT32_T, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: ckpt_every
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_history(ctx, inc, rec) BIND(C, name='cpfem_calib_history')
            IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT32_T), VALUE :: inc
            REAL(C_DOUBLE), INTENT(OUT) :: rec(19, *)
        END FUNCTION

        INTEGER(C_INT) FUNCTION cpfem_calib_end(ctx) BIND(C, name='cpfem_calib_end')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
        END FUNCTION
    END INTERFACE
END MODULE CPFEM_CAPI
