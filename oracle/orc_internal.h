/* Oracle: plain-C FP64 restatement of the OpenCFD-EC residual + LU-SGS iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the MI355X library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product path
 * never calls it.
 *
 * Every routine follows the reference operation order (Fortran left-to-right evaluation,
 * explicit parentheses below) so that results agree with the compiled reference bit for bit
 * wherever no libm transcendental is involved.  Citations are file:line in the reference tree.
 *
 * Array convention: every per-block array (cell, face, node) lives in one padded box
 * i in [-3, nx+3], j in [-3, ny+3], k in [-3, nz+3] (the reference's LAP=4 ghost range),
 * i fastest.  Reference index (i,j,k) maps to IX(b,i,j,k).
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define LAP 4
#define ORC_MAXV 7

typedef struct {
    int ib, ie, jb, je, kb, ke, bc, face, f_no;     /* sub_modules.f90:65 */
    int ib1, ie1, jb1, je1, kb1, ke1, nb1, face1, f_no1;
    int L1, L2, L3;
} orc_bcmsg;

typedef struct {
    int nx, ny, nz, NI, NJ, NK;
    size_t n;
    int nsub;
    orc_bcmsg *bc;
    int IF_OverLimit;
    int nneg_last;                 /* limiter count of the last limit_flow_U call */
    /* geometry (sub_modules.f90:76-84) */
    double *x, *y, *z, *xc, *yc, *zc, *vol, *Si, *Sj, *Sk;
    double *nv[9];                 /* ni1 ni2 ni3 nj1 nj2 nj3 nk1 nk2 nk3 */
    double *met[4][9];             /* set s: ix iy iz jx jy jz kx ky kz (s=0..3) */
    double *dw, *dtime_mesh;
    int *BcI[2], *BcJ[2], *BcK[2]; /* 2-D maps, see BCI/BCJ/BCK */
    /* flow (sub_modules.f90:86-95) */
    double *U[ORC_MAXV], *Un[ORC_MAXV], *Res[ORC_MAXV], *DU[ORC_MAXV];
    double *dt, *mu, *mut;
    double *Surf[6][3];            /* Surf1..6(.,.,1:3), 2-D maps */
    /* per-iteration temporaries (Flow_Var, opencfd_ec3d_v1.16a.f90:115-121) */
    double *d, *uu, *v, *w, *T, *p, *cc;
    double *Flux[ORC_MAXV];
    double *Lci, *Lcj, *Lck, *Lvi, *Lvj, *Lvk;
} orc_block;

typedef struct {
    /* namelist values after set_const_para (sub_read_parameter.f90:371-429) */
    double Ma, Re, gamma, Cv, Cp, PrL, PrT, T_inf, Twall, AoA, AoS, A_alfa, A_beta, P_outlet;
    double CFL, dtmax, dtmin, dt_global, w_LU, MUT_MAX, Kt_inf, Wt_inf;
    double Ldmin, Ldmax, Lpmin, Lpmax, Lumax, LSAmax, Ref_S, Ref_L, Centroid[3];
    double Periodic_dX, Periodic_dY, Periodic_dZ;
    int NVAR, If_viscous, turb, Scheme, Flux, Recon, Bound_Scheme, Iflag_local_dt;
    int If_dtime_mesh, IF_Scheme_Positivity, IFLAG_LIMIT_FLOW, Cood_Y_UP, Time_Method;
    int nblk, cap;
    orc_block *blk;
    long Num_Cell;
    int Kstep;
    double tt;
    double Res_max[ORC_MAXV], Res_rms[ORC_MAXV];
    double force[9 + 3];           /* Ft0(1:9) after scaling, then CL CD CS */
} orc_ctx;

#define IX(b, i, j, k) ((size_t)((i) + 3) + (size_t)(b)->NI * ((size_t)((j) + 3) + (size_t)(b)->NJ * (size_t)((k) + 3)))
/* 2-D face maps: BcI(j,k), BcJ(i,k), BcK(i,j), Surf1/4(j,k), Surf2/5(i,k), Surf3/6(i,j) */
#define JK(b, j, k) ((size_t)((j) + 3) + (size_t)(b)->NJ * (size_t)((k) + 3))
#define IK(b, i, k) ((size_t)((i) + 3) + (size_t)(b)->NI * (size_t)((k) + 3))
#define IJ(b, i, j) ((size_t)((i) + 3) + (size_t)(b)->NI * (size_t)((j) + 3))

#define PI_EC 3.1415926535897932
#define LIM_ZERO 1.0e-20

/* scheme / flux ids (sub_modules.f90:21-25) */
enum { S_UD1 = 0, S_NND2 = 1, S_UD3 = 2, S_MUSCL2U = 3, S_MUSCL2C = 4, S_MUSCL3 = 5, S_OMUSCL2 = 6,
       S_WENO5 = 7, S_UD5 = 8, S_WENO7 = 9, S_CD2 = 20 };
enum { F_SW = 1, F_HLL = 2, F_HLLC = 3, F_ROE = 4, F_VL = 5, F_AUSM = 6 };
enum { BCW = 2, BCSYM = 3, BCFAR = 4, BCIN = 5, BCOUT = 6, BCEXT = 401 };

/* setup (orc_setup.c) */
void orc_coordinate_ghosts(orc_ctx *c);
void orc_geometry(orc_ctx *c);
void orc_mesh_center_halo(orc_ctx *c);
void orc_mesh_quality(orc_ctx *c);
void orc_wall_distance(orc_ctx *c);
void orc_set_bck(orc_ctx *c);
void orc_halo_map(const orc_bcmsg *bc, int layers, int *kb, int *ke, int *kb1, int *ke1,
                  int *L, int *P, int *ks);

/* hot path (orc_step.c) */
void orc_boundary(orc_ctx *c);
void orc_halo(orc_ctx *c);
void orc_residual_all(orc_ctx *c);
void orc_update_limit(orc_ctx *c);
void orc_force_vt_kw(orc_ctx *c);
int orc_max_res(orc_ctx *c, int *where);
void orc_force(orc_ctx *c);

#endif
