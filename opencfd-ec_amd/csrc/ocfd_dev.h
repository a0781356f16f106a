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
    int mixed;                   // 1: reconstruction, face fluxes and the SA terms in FP32
};

struct Blk {
    int nx, ny, nz, NI, NJ, NK;
    double *U[MAXV], *Res[MAXV], *DU[MAXV];
    double *dt, *mu, *muf, *mut;        // mu: Sutherland; muf: mu after the turbulence-model (x*Re)/Re round trip
    double *prim[7];                     // rho u v w p T c (per iteration; limiter snapshot afterwards)
    double *Lc[3], *Lv[3];
    double *al[3];                       // du_LU_SGS's diagonal of the cell pass (one, or three for SA), for the sweeps
    double *vol, *S[3], *nv[9], *met[4][9], *dw, *dtm, *x, *y, *z, *xc, *yc, *zc;
    double *F[3][5];                     // inviscid face flux, m = 1..5
    double *Fv[3][4];                    // viscous face flux, m = 2..5
    double *Fsa[3][2];                   // SA face terms: convection+diffusion flux, grad(vt).n S; SST: k and omega fluxes
    // mixed precision: FP32 primitives (rho u v w p T) and FP32 face terms; the face arrays share
    // the memory of their FP64 counterparts
    float *primf[6], *Ff[3][5], *Fvf[3][4], *Fsaf[3][2];
    // mixed precision: FP32 copies of the geometry the FP32 kernels read (face normals, areas,
    // metrics, and per direction the node difference t1 of the inviscid flux's face frame), made
    // once per block from the FP64 arrays
    float *nvf[9], *Sf[3], *metf[4][9], *t1f[3][3];
    double *aux;                         // turbulence scratch (SST f1)
    double *kt, *wt;                     // SST: k and omega per unit mass, max(U, 0) / rho, on the box
    uint8_t *bcf[3];                     // face-row flags: bit0 lower face physical, bit1 upper face physical
    double *surf[6][3];                  // Surf1..6 (wall viscous force of the i-/j-/k-/i+/j+/k+ faces)
    int *flags;                          // [0] IF_OverLimit latch, [1] limited cells this iteration, [2] SA cells above LSAmax
    unsigned long long *samask;          // SA cells above LSAmax, one bit per cell in (i,j,k) order (SA only)
};

__device__ __host__ __forceinline__ size_t IX(const Blk& b, int i, int j, int k) {
    return (size_t)(i + 3) + (size_t)b.NI * ((size_t)(j + 3) + (size_t)b.NJ * (size_t)(k + 3));
}
__device__ __host__ __forceinline__ size_t JK(const Blk& b, int j, int k) { return (size_t)(j + 3) + (size_t)b.NJ * (size_t)(k + 3); }
__device__ __host__ __forceinline__ size_t IK(const Blk& b, int i, int k) { return (size_t)(i + 3) + (size_t)b.NI * (size_t)(k + 3); }
__device__ __host__ __forceinline__ size_t IJ(const Blk& b, int i, int j) { return (size_t)(i + 3) + (size_t)b.NI * (size_t)(j + 3); }

// what the residual norms read of a block
struct NormBlk {
    const double* res[MAXV];
    int nx, ny, nz, NI, NJ;
};

// one subface of the .inc table (sub_modules.f90:63-68)
struct Sub {
    int ib, ie, jb, je, kb, ke, bc, face, f_no;
    int ib1, ie1, jb1, je1, kb1, ke1, nb1, face1, f_no1;
    int L1, L2, L3;
};

// ---------------------------------------------------------------------------------------------
// Block sets: every block of a context (of every case it holds) goes through the same launch.
// A launch over an iteration space covers the tiles of all blocks; tile t belongs to block
// tile_blk[t] and is that block's (t - tile0[block])-th tile of 64 points (i fastest).

enum Space { SP_BOX = 0, SP_CELL, SP_FI, SP_FJ, SP_FK, SP_N0, SP_COUNT };

// extent and origin of an iteration space of one block
__device__ __host__ __forceinline__ void space_dims(int sp, const Blk& b, int& n0, int& n1, int& n2, int& o0, int& o1, int& o2) {
    switch (sp) {
    case SP_BOX: n0 = b.NI; n1 = b.NJ; n2 = b.NK; o0 = o1 = o2 = -3; break;
    case SP_CELL: n0 = b.nx - 1; n1 = b.ny - 1; n2 = b.nz - 1; o0 = o1 = o2 = 1; break;
    case SP_FI: n0 = b.nx; n1 = b.ny - 1; n2 = b.nz - 1; o0 = o1 = o2 = 1; break;
    case SP_FJ: n0 = b.nx - 1; n1 = b.ny; n2 = b.nz - 1; o0 = o1 = o2 = 1; break;
    case SP_FK: n0 = b.nx - 1; n1 = b.ny - 1; n2 = b.nz; o0 = o1 = o2 = 1; break;
    default: n0 = b.nx + 1; n1 = b.ny + 1; n2 = b.nz + 1; o0 = o1 = o2 = 0; break;
    }
}

struct BSet {
    const Blk* blk;
    const Par* par;
    const int* tile_blk;     // for the launch's space
    const int* tile0;
    const int* halt;         // nonzero *halt: an LU-SGS sweep of this iteration gave up; the update
                             // kernels then leave U as it is (may be null)
};

__device__ __forceinline__ bool halted(const BSet& S) { return S.halt && *(volatile const int*)S.halt != 0; }

// Workgroups of a one-dimensional grid are handed to the 8 XCDs in turn (workgroup t to XCD
// t mod 8).  Renumbering so that XCD x runs the x-th eighth of the tiles in order keeps the
// j/k neighbour rows of a stencil, which are also tiles of the same launch, in that XCD's L2;
// the last n mod 8 tiles keep their numbers (a bijection on [0, n)).
__device__ __forceinline__ int xcd_tile(int t, int n) {
    const int full = n & ~7;
    return t < full ? (t & 7) * (full >> 3) + (t >> 3) : t;
}

constexpr int ROW_CHUNK = 16;

// point `threadIdx.x` of tile t of the launch's iteration space
template <int SP>
__device__ __forceinline__ bool locate_tile(const BSet& S, int t, int& bi, int& i, int& j, int& k) {
    bi = S.tile_blk[t];
    int n0, n1, n2, o0, o1, o2;
    space_dims(SP, S.blk[bi], n0, n1, n2, o0, o1, o2);
    // the block's points in i-fastest order, 64 per tile: a tile runs on into the next row when
    // a row is shorter or not a multiple of 64 (cut blocks of 16-cell rows keep every lane busy)
    const int a = (t - S.tile0[bi]) * 64 + (int)threadIdx.x;
    if (a >= n0 * n1 * n2) return false;
    const int r = a / n0;
    i = o0 + (a - r * n0);
    // rows (j, k) in chunks of ROW_CHUNK planes along k, k fastest inside a chunk: the rows a
    // stencil reads at j +- 1 are ROW_CHUNK rows away and those at k +- 1 (k +- 4 for the
    // reconstruction of the k faces) the neighbouring rows, so both stay in the L2 of the XCD that
    // runs the tile (the renumbering above keeps a launch's neighbouring tiles on one XCD)
    const int full = n2 / ROW_CHUNK * ROW_CHUNK;
    if (r < n1 * full) {
        const int c = r / (n1 * ROW_CHUNK), rr = r - c * (n1 * ROW_CHUNK);
        j = o1 + rr / ROW_CHUNK;
        k = o2 + c * ROW_CHUNK + rr % ROW_CHUNK;
    } else {
        const int rem = n2 - full, rr = r - n1 * full;
        j = o1 + rr / rem;
        k = o2 + full + rr % rem;
    }
    return true;
}

template <int SP>
__device__ __forceinline__ bool locate(const BSet& S, int& bi, int& i, int& j, int& k) {
    return locate_tile<SP>(S, xcd_tile((int)blockIdx.x, (int)gridDim.x), bi, i, j, k);
}

#define OCFD_LOCATE(SP)                                  \
    int bi_, i, j, k;                                    \
    if (!locate<(SP)>(S, bi_, i, j, k)) return;          \
    const Blk& b = S.blk[bi_];                           \
    const Par& p = S.par[bi_];

// one subface of one block in a subface launch: n0 x n1 points, tiles of 64 along n0
enum SubKind { SUB_WALL = 0, SUB_FAR, SUB_SYM, SUB_EXT, SUB_AMUT, SUB_KBC };
struct SubTask {
    int blk, kind, arg, n0, n1, tile0;
    Sub B;
};
struct SubSet {
    const Blk* blk;
    const Par* par;
    const SubTask* t;
    const int* tile_task;
};

#define OCFD_SUB_LOCATE                                                      \
    const SubTask& T_ = S.t[S.tile_task[blockIdx.x]];                        \
    const Blk& b = S.blk[T_.blk];                                            \
    const Par& p = S.par[T_.blk];                                            \
    const Sub& B = T_.B;                                                     \
    const int lt_ = (int)blockIdx.x - T_.tile0, T0_ = (T_.n0 + 63) / 64;    \
    const int a = (lt_ % T0_) * 64 + (int)threadIdx.x, c = lt_ / T0_;

// block-to-block ghost copy of one inner subface (Umessage_send_mpi's index map)
struct HaloMap { int kb[3], kb1[3], ke1[3], L[3], P[3], ks[3]; };
struct HaloTask {
    int src, dst, n, tile0;      // n points, tiles of 256
    HaloMap h;
};
struct HaloSet {
    const Blk* blk;
    const HaloTask* t;
    const int* tile_task;
};

// one tile of the pipelined LU-SGS sweep: cells j0..j0+tj-1, k0..k0+tk-1 of block `blk`, all i
struct LuTask {
    int blk, j0, k0, tj, tk;
    int dep_j, off_j, last_j;    // upstream tile along j: task index (-1: none), step offset, its last step
    int dep_k, off_k, last_k;    // upstream tile along k
    long long pofs;              // sweep-ordered storage of this tile: element (step q, field f, lane) at
                                 // pofs + (q NF + f) 64 + lane
    long long pofs_j;            // the j-upstream tile's storage offset
    long long pofs_k;            // the k-upstream tile's storage offset
    int tj_up, tk_up;            // the upstream tiles' j and k extents
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
    // sweep-ordered copies of the sweep's inputs and of DU (layout below), P elements
    double* pk = nullptr;
    long long P = 0;
    const int2* win = nullptr;   // (forward tile, first i) of every i-window of PK_W cells of the tiles
    int nwins = 0;
    int tj = 8;                  // tile width in j (tiles are tj x 64/tj (j,k) columns)
};

// lanes of a pipelined LU-SGS tile (one 64-wide wave per tile and sweep direction)
constexpr int LU_NT = 64;

// The sweep-ordered storage.  For a tile, step q (the cells its lanes visit at forward step q; the
// backward sweep visits them at step nsteps-1-q) holds NF fields of 64 consecutive doubles, one
// per lane: element (q, f, lane) at pofs + (q NF + f) 64 + lane_pos(lane).  A step's inputs therefore sit in
// one contiguous run that a lane addresses from one pointer with constant offsets.  AL holds the
// LU-SGS diagonal: alfa of components 1..5 (and of 6, 7 for SST) and, for SA only, the SA
// component's alfa of the forward (AL+1) and of the backward sweep (AL+2), which sum the viscous
// radii in different orders.  The buffer
// has LU_PAD_STEPS steps of padding at both ends, so loads a few steps past either end of a tile
// (prefetches of steps that do not exist, at most the larger tile side + 2 steps before an
// upstream tile's first step) stay inside it.
template <int NV> struct Pk {
    static constexpr int U = 0, R = NV, DUF = 2 * NV, DUB = 3 * NV, LC = 4 * NV, LV = 4 * NV + 3, AL = 4 * NV + 6,
                         S = 4 * NV + 9, N = 4 * NV + 12, NF = 4 * NV + 21;
};
// Position of lane (jl, kl) = lane % TJ, lane / TJ of a TJ x (64/TJ) tile inside a step's run of 64
// elements: first the tile's j-edge columns (jl = TJ-1, then jl = 0), then the rest of its k-edge
// rows (kl = TK-1, then kl = 0), then the interior row by row.  The lanes a downstream tile reads
// across an edge therefore share few 128-byte lines (8x8 tiles: one line per field for the
// j-edge and two for the k-edge, against four and one in row order).
template <int TJ>
__device__ __host__ __forceinline__ int lane_pos(int lane) {
    constexpr int TK = 64 / TJ;
    const int jl = lane % TJ, kl = lane / TJ;
    if (jl == TJ - 1) return kl;
    if (jl == 0) return TK + kl;
    if (kl == TK - 1) return 2 * TK + (jl - 1);
    if (kl == 0) return 2 * TK + (TJ - 2) + (jl - 1);
    return 2 * TK + 2 * (TJ - 2) + (kl - 1) * (TJ - 2) + (jl - 1);
}
// the lane at position pos of a step's run (the inverse of lane_pos)
template <int TJ>
__device__ __host__ __forceinline__ int pos_lane(int pos) {
    constexpr int TK = 64 / TJ;
    int jl, kl;
    if (pos < TK) { jl = TJ - 1; kl = pos; }
    else if (pos < 2 * TK) { jl = 0; kl = pos - TK; }
    else if (pos < 2 * TK + (TJ - 2)) { kl = TK - 1; jl = pos - 2 * TK + 1; }
    else if (pos < 2 * TK + 2 * (TJ - 2)) { kl = 0; jl = pos - 2 * TK - (TJ - 2) + 1; }
    else { const int r = pos - 2 * TK - 2 * (TJ - 2); kl = r / (TJ - 2) + 1; jl = r % (TJ - 2) + 1; }
    return jl + TJ * kl;
}
constexpr int LU_PAD_STEPS = 24;
constexpr int PK_W = 32;      // i-window of one workgroup of the sweep-order transposes and cell pass

// Before a sweep every DU element it writes holds this NaN (a payload no arithmetic produces); a
// consumer in a downstream tile reloads an element until it is something else.
constexpr unsigned long long LU_SENTINEL = 0x7FF4C0DEDEADBEEFull;
__device__ __forceinline__ bool is_sentinel(double v) { return (unsigned long long)__double_as_longlong(v) == LU_SENTINEL; }
__device__ __forceinline__ double sentinel_value() { return __longlong_as_double((long long)LU_SENTINEL); }

}  // namespace ocfd
