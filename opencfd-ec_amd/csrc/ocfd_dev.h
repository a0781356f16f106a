// Device-side data model of one block: every per-block array (cell, face or node) lives in one
// padded box i in [-3, nx+3], j in [-3, ny+3], k in [-3, nz+3] (the reference's LAP=4 ghost
// range), i fastest, one array per variable (SoA).  Threads map to i so every access along a
// face or cell row is a coalesced 64-wide load.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ocfd {

constexpr int LAP = 4;
constexpr int MAXV = 7;

struct Par {                 // control parameters after set_const_para (sub_read_parameter.f90:371-429)
    double Ma, Re, gamma, Cv, Cp, PrL, PrT, T_inf, Twall, AoA, A_alfa, A_beta, P_outlet;
    double CFL, dtmax, dtmin, dt_global, w_LU, MUT_MAX, Kt_inf, Wt_inf;
    double Ldmin, Ldmax, Lpmin, Lpmax, Lumax, LSAmax;
    int NV, visc, turb, scheme, flux, recon, bscheme, local_dt, dtime_mesh, positivity, limit_flow;
};

struct Blk {
    int nx, ny, nz, NI, NJ, NK;
    double *U[MAXV], *Res[MAXV], *DU[MAXV];
    double *dt, *mu, *muf, *mut;        // mu: Sutherland; muf: mu after the turbulence-model (x*Re)/Re round trip
    double *prim[7];                     // rho u v w p T c (per iteration; limiter snapshot afterwards)
    double *Lc[3], *Lv[3];
    double *vol, *S[3], *nv[9], *met[4][9], *dw, *dtm, *x, *y, *z, *xc, *yc, *zc;
    double *F[3][5];                     // inviscid face flux, m = 1..5
    double *Fv[3][4];                    // viscous face flux, m = 2..5
    double *Fsa[3][2];                   // SA face terms: convection+diffusion flux, grad(vt).n S
    double *aux;                         // turbulence scratch (SST f1)
    uint8_t *bcf[3];                     // face-row flags: bit0 lower face physical, bit1 upper face physical
    double *surf[6][3];                  // Surf1..6 (wall viscous force of the i-/j-/k-/i+/j+/k+ faces)
    int *flags;                          // [0] IF_OverLimit latch, [1] limited cells this iteration, [2] SA cells above LSAmax
};

__device__ __host__ __forceinline__ size_t IX(const Blk& b, int i, int j, int k) {
    return (size_t)(i + 3) + (size_t)b.NI * ((size_t)(j + 3) + (size_t)b.NJ * (size_t)(k + 3));
}
__device__ __host__ __forceinline__ size_t JK(const Blk& b, int j, int k) { return (size_t)(j + 3) + (size_t)b.NJ * (size_t)(k + 3); }
__device__ __host__ __forceinline__ size_t IK(const Blk& b, int i, int k) { return (size_t)(i + 3) + (size_t)b.NI * (size_t)(k + 3); }
__device__ __host__ __forceinline__ size_t IJ(const Blk& b, int i, int j) { return (size_t)(i + 3) + (size_t)b.NI * (size_t)(j + 3); }

// one subface of the .inc table (sub_modules.f90:63-68)
struct Sub {
    int ib, ie, jb, je, kb, ke, bc, face, f_no;
    int ib1, ie1, jb1, je1, kb1, ke1, nb1, face1, f_no1;
    int L1, L2, L3;
};

// one tile of the pipelined LU-SGS sweep: cells j0..j0+tj-1, k0..k0+tk-1 of block `blk`, all i
struct LuTask {
    int blk, j0, k0, tj, tk;
    int dep_j, off_j, last_j;    // upstream tile along j: task index (-1: none), step offset, its last step
    int dep_k, off_k, last_k;    // upstream tile along k
    int pofs;                    // sweep-ordered storage of this tile: element (step q, lane) at pofs + 64 q + lane
    int pofs_j, tj_up;           // the j-upstream tile's storage offset and j extent
    int pofs_k, tk_up;           // the k-upstream tile's storage offset and k extent
};

// device tables of the pipelined sweeps of one context
struct LuPlan {
    const Blk* blks = nullptr;
    const LuTask* fwd = nullptr;
    const LuTask* bwd = nullptr;
    int n_fwd = 0, n_bwd = 0;
    int* prog_fwd = nullptr;     // last completed step per tile, -1 before the sweep
    int* prog_bwd = nullptr;
    int* err = nullptr;          // set when a wait exceeds its poll limit
    // sweep-ordered copies of the sweep's inputs and of DU (field f of element e at pk[f * P + e])
    double* pk = nullptr;
    long long P = 0;
    int nwin = 0;                // i-windows per tile of the pack/unpack transposes
};

}  // namespace ocfd
