/* ocfd_ec.h - C ABI of the MI355X residual + LU-SGS library (libocfd_ec.so).
 *
 * Drop-in for the OpenCFD-EC 3D v1.16a hot path.  A host that keeps the reference's control.ec
 * parameters, Block_TYPE arrays and BC_MSG_TYPE tables hands them over once, then calls one
 * entry point per iteration in place of NS_Time_advance_LU_SGS.  All citations are file:line
 * in the OpenCFD-EC source tree.
 *
 * Conventions
 *  - Every array argument is a plain pointer to Fortran (column-major) data with the reference's
 *    own bounds, given next to each field.  LAP = 4 (sub_modules.f90:20).
 *  - Every function returns 0 on success or one of the OCFD_E* codes; ocfd_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - A context owns one HIP stream; calls are not reentrant per context and return with the
 *    device work complete (the async/graph variants say otherwise).
 *  - No entry point takes ownership of caller memory.
 */
#ifndef OCFD_EC_H
#define OCFD_EC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCFD_OK 0
#define OCFD_EHIP 1          /* HIP runtime failure (message has the HIP error string) */
#define OCFD_EARG 2          /* invalid argument (null pointer, size, unknown block) */
#define OCFD_ENAN 3          /* NaN in the residual (comput_max_Res_onemesh, sub_time_advance.f90:358-364) */
#define OCFD_EUNSUPPORTED 4  /* option or boundary condition outside the hot path (sub_boundary.f90:48-51) */

typedef struct ocfd_ctx ocfd_ctx;

/* Global_Var scalars after set_const_para (sub_modules.f90:140-183; sub_read_parameter.f90:371-429).
 * A Fortran host fills the matching bind(C) derived type from module Global_Var. */
typedef struct ocfd_params {
    double Ma, Re, gamma, Cv, Cp, PrL, PrT, T_inf;
    double Twall;            /* already divided by T_inf; <= 0 means adiabatic */
    double AoA, A_alfa, A_beta, P_outlet;
    double CFL, dtmax, dtmin, dt_global, w_LU, MUT_MAX, Kt_inf, Wt_inf;
    double Ldmin, Ldmax, Lpmin, Lpmax, Lumax, LSAmax;   /* Lpmin/Lpmax already divided by gamma Ma^2 */
    double Ref_S, Ref_L, Centroid[3];
    double Periodic_d[3];    /* Periodic_dX, Periodic_dY, Periodic_dZ (ghost-node shift of bc -2/-3 faces) */
    int32_t NVAR, If_viscous, Iflag_turbulence_model, Iflag_Scheme, Iflag_Flux;
    int32_t IFlag_Reconstruction, Bound_Scheme, Iflag_local_dt, If_dtime_mesh;
    int32_t IF_Scheme_Positivity, IFLAG_LIMIT_FLOW, Cood_Y_UP;
    int32_t precision;       /* 0: all FP64 (the drop-in).  1: reconstruction, face fluxes (inviscid,
                                viscous) and the SA or SST face and source terms in FP32; metrics,
                                LU-SGS (Jacobians, diagonal), update, limiter and norms in FP64. */
    int32_t IF_InnerFlow;      /* IF_InnerFlow (sub_modules.f90:163).  1 turns bc 5/6 into the total-
                                  pressure inlet and the outlet of the inner-flow mode and changes
                                  the initial field (sub_boundary.f90:28-40; sub_init.f90:322): such
                                  a case is refused with OCFD_EUNSUPPORTED when it has a bc 5 or 6
                                  subface, and ocfd_init_flow refuses it. */
    int32_t IF_TurboMachinary; /* IF_TurboMachinary (sub_modules.f90:162).  1 (the turbomachinery
                                  solver) is outside this path: ocfd_create refuses it. */
} ocfd_params;

/* Geometry of one block as Comput_Goemetric_var leaves it (sub_geometry.f90:6-346). */
typedef struct ocfd_block_geom {
    const double *x, *y, *z;          /* x(0:nx+1, 0:ny+1, 0:nz+1) nodes with one ghost layer */
    const double *xc, *yc, *zc;       /* xc(1-LAP:nx+LAP-1, ...) cell centres */
    const double *vol;                /* vol(1:nx-1, 1:ny-1, 1:nz-1) */
    const double *dtime_mesh;         /* dtime_mesh(1:nx-1, 1:ny-1, 1:nz-1) */
    const double *dw;                 /* dw(1:nx-1, ...); may be NULL without a turbulence model */
    const double *Si, *Sj, *Sk;       /* Si(1:nx, 1:ny, 1:nz) face areas */
    const double *n[9];               /* ni1 ni2 ni3 nj1 nj2 nj3 nk1 nk2 nk3 (1:nx, 1:ny, 1:nz) */
    const double *met[4][9];          /* [s][q]: ix iy iz jx jy jz kx ky kz of set s (ix0.. ix1.. ix2.. ix3..),
                                         (1:nx, 1:ny, 1:nz); set 0 may be NULL without a turbulence model */
} ocfd_block_geom;

/* Setup (replaces nothing on the hot path; the host's init keeps running) */
int ocfd_create(const ocfd_params *p, ocfd_ctx **out);
int ocfd_destroy(ocfd_ctx *ctx);
const char *ocfd_last_error(void);

/* One block: its Block_TYPE geometry and its BC_MSG_TYPE table, nsub records of 21 ints in the
 * bc3d.inc order ib ie jb je kb ke bc face f_no ib1 ie1 jb1 je1 kb1 ke1 nb1 face1 f_no1 L1 L2 L3
 * (sub_modules.f90:63-68; sub_IO.f90:700-705).  block_no is the reference's Block_no, 1..N. */
int ocfd_add_block(ocfd_ctx *ctx, int32_t block_no, int32_t nx, int32_t ny, int32_t nz,
                   int32_t nsub, const int32_t *bc_msg21, const ocfd_block_geom *g);

/* Node coordinates only, x(1:nx,1:ny,1:nz); ocfd_setup_geometry then builds the ghost nodes,
 * the geometry, dtime_mesh and the wall distance on the device (Update_coordinate_buffer,
 * Comput_Goemetric_var, check_mesh_quality, comput_wall_dist: sub_update_buffer_mpi.f90:230-512,
 * sub_geometry.f90:6-515, sub_comput_dw.f90:43-168). */
int ocfd_add_block_nodes(ocfd_ctx *ctx, int32_t block_no, int32_t nx, int32_t ny, int32_t nz,
                         int32_t nsub, const int32_t *bc_msg21,
                         const double *x, const double *y, const double *z);
int ocfd_setup_geometry(ocfd_ctx *ctx);

/* State transfer.  U(NVAR, 1-LAP:nx+LAP-1, 1-LAP:ny+LAP-1, 1-LAP:nz+LAP-1), the layout of
 * Block_TYPE%U (sub_init.f90:469).  The state stays resident on the device between iterations. */
int ocfd_upload_U(ocfd_ctx *ctx, int32_t block_no, const double *U);
int ocfd_download_U(ocfd_ctx *ctx, int32_t block_no, double *U);

/* Fields of the last iteration in the reference layout (-1:nx+1, -1:ny+1, -1:nz+1):
 * OCFD_F_RES / OCFD_F_DU are (NVAR, -1:nx+1, ...), the others scalar. */
#define OCFD_F_RES 1
#define OCFD_F_DU 2
#define OCFD_F_DT 3
#define OCFD_F_MU 4
#define OCFD_F_MUT 5
#define OCFD_F_DW 6      /* wall distance dw (cells 1..n-1; zero elsewhere), e.g. for write_dw
                            (sub_IO.f90:573-644) after ocfd_setup_geometry computed it */
int ocfd_download_field(ocfd_ctx *ctx, int32_t block_no, int32_t field, double *out);

/* Init_flow, freestream branch (sub_init.f90:264-410): freestream on every cell, then BC and halo */
int ocfd_init_flow(ocfd_ctx *ctx);

/* NS_Time_advance(1) with Time_Method = LU-SGS (sub_time_advance.f90:6-23,75-118), repeated
 * niter times: residual, spectral radii and dt, LU-SGS sweeps, update, limit_flow, boundary
 * conditions, block halo, force_vt_kw.  Kstep and tt advance as in the reference. */
int ocfd_step_lusgs(ocfd_ctx *ctx, int32_t niter);

/* The same iteration at the granularity of the reference's routines:
 *   ocfd_residual      = Set_Un + Comput_Residual_one_mesh incl. du_LU_SGS (sub_Residual.f90:14-127)
 *   ocfd_update_limit  = U = Un + DU and limit_flow (sub_time_advance.f90:86-109)
 *   ocfd_bc            = Boundary_condition_onemesh (sub_boundary.f90:6-80)
 *   ocfd_halo          = update_buffer_onemesh (sub_update_buffer_mpi.f90:18-36)
 *   ocfd_force_vt_kw   = force_vt_kw (sub_time_advance.f90:27-61) */
int ocfd_residual(ocfd_ctx *ctx);
int ocfd_update_limit(ocfd_ctx *ctx);
int ocfd_bc(ocfd_ctx *ctx);
int ocfd_halo(ocfd_ctx *ctx);
int ocfd_force_vt_kw(ocfd_ctx *ctx);

/* comput_max_Res_onemesh (sub_time_advance.f90:330-380): Res_max(NVAR), Res_rms(NVAR).
 * Returns OCFD_ENAN when the residual holds a NaN. */
int ocfd_res_norms(ocfd_ctx *ctx, double *res_max, double *res_rms);

/* ocfd_res_norms without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The norms' scratch is a fixed device array of the library, about
 * 0.7 MB on every GPU it runs on, held until the process ends and shared by every call on that
 * GPU: calls that overlap on one GPU (several threads or contexts at once) race on it.  At most
 * 64 blocks per context (OCFD_EARG beyond).  Same results as ocfd_res_norms. */
int ocfd_res_norms_noalloc(ocfd_ctx *ctx, double *res_max, double *res_rms);

/* comput_force (sub_Post.f90:4-265): out[0..8] = Ft0(1:9) after scaling, out[9..11] = CL CD CS */
int ocfd_forces(ocfd_ctx *ctx, double *out12);

/* Block status: IF_OverLimit latch and the number of cells limit_flow_U replaced in the last
 * iteration (sub_limitflow.f90:82-172). */
int ocfd_block_status(ocfd_ctx *ctx, int32_t block_no, int32_t *overlimit, int32_t *nlimited);

/* Iteration counters (Mesh(1)%Kstep, Mesh(1)%tt). */
int ocfd_get_counters(ocfd_ctx *ctx, int32_t *kstep, double *tt);

/* Stages of one iteration, in order.  ocfd_step_lusgs runs them one after the other; with graphs
 * on, each stage is captured in its own hipGraph on first use and replayed afterwards.
 *   PRE         comput_duvtpc, get_viscous, Turbulence_model_SA/SST, Amut_boundary
 *               (sub_Residual.f90:62-80)
 *   INVISCID    flux_inviscous_i/j/k incl. the k-face boundary flux (sub_Residual.f90:86-88)
 *   VISCOUS_DT  flux_viscous_i/j/k, Res assembly, comput_Lijk, comput_dt (sub_Residual.f90:90-117)
 *   LU_FORWARD / LU_BACKWARD   du_LU_SGS sweeps (sub_LU_SGS.f90:20-126)
 *   UPDATE      U = Un + DU, limit_flow (sub_time_advance.f90:86-109)
 *   BC_HALO     Boundary_condition_onemesh, update_buffer_onemesh, force_vt_kw
 *               (sub_time_advance.f90:110-117) */
#define OCFD_STAGE_PRE 0
#define OCFD_STAGE_INVISCID 1
#define OCFD_STAGE_VISCOUS_DT 2
#define OCFD_STAGE_LU_FORWARD 3
#define OCFD_STAGE_LU_BACKWARD 4
#define OCFD_STAGE_UPDATE 5
#define OCFD_STAGE_BC_HALO 6
#define OCFD_NSTAGES 7

/* ocfd_step_lusgs with a HIP event at every stage boundary on the context's stream:
 * stage_ms[s] = device time of stage s summed over the niter iterations.  Instrumentation; the
 * reference has no counterpart (show_Wall_time, opencfd_ec3d_v1.16a.f90:235, times whole steps). */
int ocfd_step_timed(ocfd_ctx *ctx, int32_t niter, double *stage_ms);

/* Capture each stage of the iteration in a hipGraph on first use and replay it (1) or launch
 * eagerly (0). */
int ocfd_set_graph(ocfd_ctx *ctx, int32_t enable);

/* Schedule of the LU-SGS sweeps; all give the reference's DU bit for bit.
 *   OCFD_LU_PIPELINED        (default) one persistent launch per sweep: (j,k) tiles of 64
 *                            columns (8x8 unless set otherwise) march along i, three waves per
 *                            tile (one per neighbour direction); tile-edge DU exchanged through
 *                            sentinel-marked, device-coherent stores and loads
 *   OCFD_LU_HYPERPLANE       one launch per i+j+k hyperplane, the reference's own order
 *                            (sub_LU_SGS.f90:20-126)
 *   OCFD_LU_PIPELINED_FENCED one wave per 8x8 tile on the block layout, progress counters with
 *                            agent-scope release/acquire fences at every step (L2 write-back
 *                            and invalidate) */
#define OCFD_LU_PIPELINED 0
#define OCFD_LU_HYPERPLANE 1
#define OCFD_LU_PIPELINED_FENCED 2
int ocfd_set_lu_schedule(ocfd_ctx *ctx, int32_t schedule);

/* Tile shape of OCFD_LU_PIPELINED: tj x (64/tj) (j,k) columns per tile, tj = 4, 8 (default) or
 * 16.  Results do not depend on it (bit for bit); the sweep time does.  No counterpart in the
 * reference. */
int ocfd_set_lu_tile(ocfd_ctx *ctx, int32_t tj);

/* Failure-path switch of the pipelined LU-SGS sweeps of this context (not of a batch).  A sweep
 * wait normally gives up after 2^26 polls of an upstream tile (seconds); the step then returns
 * OCFD_EHIP, U is left at the last iteration completed before the failure (the update, limiter
 * and force_vt_kw kernels of that and every later iteration of the call leave it as it is, and
 * the later sweeps return at once), and the sweep tables are restored for the next call.  With on = 1 every wait that finds its producer behind gives up at its first poll, which
 * forces that path; on = 0 restores normal waiting.  No counterpart in the reference. */
int ocfd_set_lu_fail_waits(ocfd_ctx *ctx, int32_t on);

/* Launch mode.  on = 1: every kernel of the iteration is launched once per block, eagerly (no
 * graphs), and the LU-SGS sweeps keep 8x8 tiles (the autotuner refuses); on = 0 (the default):
 * one launch per kernel over all blocks, graphs per ocfd_set_graph, any tile shape.  The results
 * are the same bit for bit.  No counterpart in the reference (the per-block form is its block
 * loop, for comparison with the batched launches). */
int ocfd_set_per_block(ocfd_ctx *ctx, int32_t on);
int ocfd_get_lu_tile(ocfd_ctx *ctx, int32_t *tj);

/* Tile-shape autotuner of OCFD_LU_PIPELINED.  No counterpart in the reference.  Runs niter
 * ordinary iterations with each tile shape (tj = 4, 8, 16), times the two sweeps of each with HIP
 * events and keeps the fastest shape.  The solve advances by 3 niter iterations exactly as plain
 * iterations would advance it (every shape gives the same result bit for bit); counters and time
 * advance likewise.  best_tj (may be NULL) receives the shape kept, ms (may be NULL, 3 entries)
 * the device time of one forward plus backward sweep for tj = 4, 8, 16. */
int ocfd_autotune_lu(ocfd_ctx *ctx, int32_t niter, int32_t *best_tj, double *ms);

/* Batches.  No counterpart in the reference: several contexts (solves of different cases with
 * the same NVAR, schemes, flux, turbulence model, limiter, gamma and w_LU) advance together,
 * every kernel of the iteration covering the blocks of all of them in one launch; each context
 * keeps its own parameters (Mach, AoA, Re, ...), state, counters, norms and forces.  A context is
 * added after its geometry and state are set; while it is in a batch it is stepped through the
 * batch only: ocfd_step_lusgs, ocfd_step_timed, ocfd_residual, ocfd_update_limit, ocfd_bc,
 * ocfd_halo, ocfd_force_vt_kw, ocfd_init_flow and ocfd_autotune_lu refuse a member with OCFD_EARG
 * (ocfd_res_norms, ocfd_forces and the downloads work).  A context is in at most one batch.
 * ocfd_destroy of a member first removes it from its batch; ocfd_batch_destroy releases its
 * members, which may then be stepped alone.  Results are those of the context stepped alone, bit
 * for bit. */
typedef struct ocfd_batch ocfd_batch;
int ocfd_batch_create(ocfd_batch **out);
int ocfd_batch_destroy(ocfd_batch *batch);
int ocfd_batch_add(ocfd_batch *batch, ocfd_ctx *ctx);
int ocfd_batch_remove(ocfd_batch *batch, ocfd_ctx *ctx);
int ocfd_batch_step(ocfd_batch *batch, int32_t niter);
/* ocfd_step_timed for the batch's shared launches. */
int ocfd_batch_step_timed(ocfd_batch *batch, int32_t niter, double *stage_ms);

/* Self-check of the kernels' shared-reciprocal FP64 division (the IEEE division sequence without
 * its operand pre-scaling and special-value fix-up, used where several divisions share a divisor).
 * No counterpart in the reference.  Divides n hashed operand pairs (numerator exponents within
 * +-numerator_exponent, divisor exponents within +-40, signs and exact zeros included) both ways
 * on the current device and returns the number of quotients whose values differ. */
int ocfd_check_division(int64_t n, uint64_t seed, int32_t numerator_exponent, int64_t *mismatches);
int ocfd_batch_set_graph(ocfd_batch *batch, int32_t enable);
/* ocfd_set_lu_tile and ocfd_autotune_lu for the batch's shared launches (every member advances
 * by 3 niter iterations). */
int ocfd_batch_set_lu_tile(ocfd_batch *batch, int32_t tj);
int ocfd_batch_autotune_lu(ocfd_batch *batch, int32_t niter, int32_t *best_tj, double *ms);

/* Device memory held by the context's block arrays, in bytes (admission control). */
int ocfd_mem_bytes(ocfd_ctx *ctx, int64_t *bytes);

/* Wait for the context's stream. */
int ocfd_sync(ocfd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
