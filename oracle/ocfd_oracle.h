/* Public entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C FP64 restatement of the OpenCFD-EC residual + LU-SGS iteration and of the setup
 * it consumes.  Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; the MI355X library never links or calls it.
 *
 * orc_create(rp, ip): namelist values (before set_const_para)
 *   rp[0..31] = Ma Re gamma AoA AoS P_outlet PrL PrT T_inf Twall CFL dtmax dtmin dt_global w_LU
 *               MUT_MAX Kt_inf Wt_inf Ldmin Ldmax Lpmin Lpmax Lumax LSAmax Ref_S Ref_L
 *               Centroid(1:3) Periodic_dX Periodic_dY Periodic_dZ
 *   ip[0..11] = If_viscous Iflag_turbulence_model Iflag_Scheme Iflag_Flux IFlag_Reconstruction
 *               Bound_Scheme Iflag_local_dt If_dtime_mesh IF_Scheme_Positivity IFLAG_LIMIT_FLOW
 *               Cood_Y_UP Time_Method
 */
#ifndef OCFD_ORACLE_H
#define OCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx_s orc_ctx_t;

#ifndef ORC_INTERNAL_H
typedef struct orc_ctx_opaque orc_ctx;
#endif

orc_ctx *orc_create(const double *rp, const int *ip);
void orc_destroy(orc_ctx *c);
int orc_add_block(orc_ctx *c, int nx, int ny, int nz, const double *x, const double *y, const double *z);
int orc_add_subface(orc_ctx *c, int blk, const int *rng6, int bc, const int *link7);
int orc_get_inc(orc_ctx *c, int blk, int *out21);
int orc_setup(orc_ctx *c);
int orc_init_flow(orc_ctx *c);
int orc_set_interior_U(orc_ctx *c, int blk, const double *U);
int orc_bc_halo(orc_ctx *c);
int orc_step(orc_ctx *c);
int orc_residual(orc_ctx *c);
int orc_update(orc_ctx *c);
int orc_bc(orc_ctx *c);
int orc_halo_only(orc_ctx *c);
int orc_vt_kw(orc_ctx *c);
int orc_monitor(orc_ctx *c);
void orc_block_dims(orc_ctx *c, int blk, int *out9);
int orc_get_array(orc_ctx *c, int blk, const char *name, double *out);
int orc_set_array(orc_ctx *c, int blk, const char *name, const double *in);
int orc_get_state(orc_ctx *c, int *kstep, double *tt, double *res_max, double *res_rms, double *force12);
/* face-level hooks for closed-form checks: 1-D flux of rotated primitive states (rho, un, ut1, ut2,
 * p), and the left/right reconstruction of u(-3..4) around face i-1/2 */
int orc_face_flux(int iflux, const double *QL, const double *QR, double gamma, double *F);
int orc_face_recon(int scheme, const double *u8, double *LR);

#ifdef __cplusplus
}
#endif
#endif
