// LU-SGS sweeps (du_LU_SGS, sub_LU_SGS.f90:4-129) as one persistent, pipelined launch per sweep.
//
// The forward sweep of a cell needs the final DU of its i-1, j-1 and k-1 neighbours (the
// backward sweep: i+1, j+1, k+1), and nothing else that changes during the sweep.  Any order that
// honours those three dependencies gives the reference's hyperplane result bit for bit, because
// every cell evaluates the same expression in the same order once its neighbours are final.
//
// Schedule: the (j,k) cross-section of a block is cut into 8x8 tiles; one 64-lane wave owns a
// tile and marches along i, lane (jl,kl) handling cell i = q - jl - kl + 1 at local step q (a
// hyperplane inside the tile).  Within the tile the j/k neighbours come through LDS (the lane
// that computed them one step earlier) and the i neighbour stays in registers.  Across tiles the
// upstream tile publishes its last completed step in a progress counter (agent-scope release)
// and the downstream tile waits (agent-scope acquire) for the step that produced the cells it
// reads.  Tiles are numbered so that every upstream tile has a lower workgroup index; workgroups
// are dispatched in index order, so every wait is on a workgroup already running or finished,
// and every wait is also bounded by a poll limit that raises an error flag instead of hanging.
#include <hip/hip_runtime.h>

#include <vector>

#include "ocfd_dev.h"
#include "ocfd_launch.h"
#include "ocfd_phys.h"

namespace ocfd {

constexpr int LU_TJ = 8, LU_TK = 8, LU_NT = LU_TJ * LU_TK;
constexpr long LU_POLL_LIMIT = 1L << 26;

// Per-lane record shared through LDS with the downstream lanes of the tile.
template <int NV>
struct LuSlot {
    double U[NV], DU[NV];
    double Lc[2], Lv[2], S[2];   // j and k direction values of the cell (S: backward sweep only)
};

template <int NV, bool FWD>
__global__ void __launch_bounds__(LU_NT) k_lusgs_sweep(const Blk* __restrict__ blks, const LuTask* __restrict__ tasks,
                                                       int* __restrict__ prog, int* __restrict__ prog_other,
                                                       int n_other, int* __restrict__ err, Par p) {
    __shared__ LuSlot<NV> slot[2][LU_NT];
    const int w = blockIdx.x, lane = threadIdx.x;
    // this launch re-arms the counters of the other sweep (same tile count) for its next launch
    if (lane == 0 && w < n_other) prog_other[w] = -1;

    const LuTask T = tasks[w];
    const Blk& b = blks[T.blk];
    const int jl = lane % LU_TJ, kl = lane / LU_TJ;
    const bool live = jl < T.tj && kl < T.tk;
    // tile-local offsets along the sweep direction
    const int jo = FWD ? jl : T.tj - 1 - jl, ko = FWD ? kl : T.tk - 1 - kl;
    const int j = T.j0 + jl, k = T.k0 + kl;
    const int nxm = b.nx - 1;
    const size_t sj = (size_t)b.NI, sk = (size_t)b.NI * b.NJ;
    const double wlu = p.w_LU, gamma = p.gamma;
    const bool visc = p.visc == 1;
    const int nsteps = nxm + T.tj + T.tk - 2;
    int known_j = -1, known_k = -1;

    // i-direction neighbour (the previous cell of this lane)
    double pU[NV], pDU[NV], pLc = 0.0, pLv = 0.0, pS = 0.0;
#pragma unroll
    for (int m = 0; m < NV; m++) { pU[m] = 0.0; pDU[m] = 0.0; }

    for (int q = 0; q < nsteps; q++) {
        const int buf = q & 1;
        // ---- wait for the upstream tiles (lane 0 polls, the barrier releases the wave) ----
        if (lane == 0) {
            int need_j = T.dep_j >= 0 ? min(q + T.off_j, T.last_j) : -1;
            int need_k = T.dep_k >= 0 ? min(q + T.off_k, T.last_k) : -1;
            bool fresh = false;
            long polls = 0;
            while (known_j < need_j || known_k < need_k) {
                if (known_j < need_j) known_j = __hip_atomic_load(prog + T.dep_j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (known_k < need_k) known_k = __hip_atomic_load(prog + T.dep_k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                fresh = true;
                if (known_j >= need_j && known_k >= need_k) break;
                if (++polls > LU_POLL_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    known_j = known_k = -2;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (fresh) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0) return;

        const int i = FWD ? q - jo - ko + 1 : nxm - (q - jo - ko);
        const bool act = live && i >= 1 && i <= nxm;
        double oU[NV], oDU[NV], oLc1 = 0.0, oLc2 = 0.0, oLv1 = 0.0, oLv2 = 0.0, oS1 = 0.0, oS2 = 0.0;
        if (act) {
            const size_t s = IX(b, i, j, k);
            double Lc0 = b.Lc[0][s], Lc1 = b.Lc[1][s], Lc2 = b.Lc[2][s];
            double Lv0 = b.Lv[0][s], Lv1 = b.Lv[1][s], Lv2 = b.Lv[2][s];
            double vol = b.vol[s];
            double sumLc = (Lc0 + Lc1) + Lc2;
            double sumLv = (Lv0 + Lv1) + Lv2;
            double alfa0 = vol / b.dt[s] + wlu * sumLc;
            double alfa[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) alfa[m] = alfa0;
            if (visc) {
#pragma unroll
                for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 2.0 * sumLv;
                if (NV == 6) {
                    if (FWD) alfa[5] = alfa[5] + sumLv;
                    else alfa[5] = ((alfa[5] + Lv0) + Lv1) + Lv2;
                }
            }
#pragma unroll
            for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 0.0 * vol;
            double acc[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) acc[m] = FWD ? b.Res[m][s] : 0.0;
#pragma unroll
            for (int dir = 0; dir < 3; dir++) {
                const int pos = dir == 0 ? i : (dir == 1 ? j : k);
                const int lim = dir == 0 ? nxm : (dir == 1 ? b.ny - 1 : b.nz - 1);
                const bool has = FWD ? pos != 1 : pos != lim;
                double du[NV];
                if (has) {
                    double Un[NV], DUn[NV], Lcn, Lvn, Sn;
                    const int src = dir == 0 ? 0 : (dir == 1 ? (jo > 0 ? 1 : 2) : (ko > 0 ? 1 : 2));
                    if (src == 0) {            // previous cell of this lane
#pragma unroll
                        for (int m = 0; m < NV; m++) { Un[m] = pU[m]; DUn[m] = pDU[m]; }
                        Lcn = pLc; Lvn = pLv; Sn = pS;
                    } else if (src == 1) {     // lane of the same tile, one step earlier
                        const int nl = dir == 1 ? (FWD ? lane - 1 : lane + 1) : (FWD ? lane - LU_TJ : lane + LU_TJ);
                        const LuSlot<NV>& o = slot[buf ^ 1][nl];
                        const int d = dir - 1;
#pragma unroll
                        for (int m = 0; m < NV; m++) { Un[m] = o.U[m]; DUn[m] = o.DU[m]; }
                        Lcn = o.Lc[d]; Lvn = o.Lv[d]; Sn = o.S[d];
                    } else {                   // upstream tile, through global memory
                        const size_t t = FWD ? s - (dir == 1 ? sj : sk) : s + (dir == 1 ? sj : sk);
#pragma unroll
                        for (int m = 0; m < NV; m++) { Un[m] = b.U[m][t]; DUn[m] = b.DU[m][t]; }
                        Lcn = b.Lc[dir][t]; Lvn = b.Lv[dir][t]; Sn = FWD ? 0.0 : b.S[dir][t];
                    }
#pragma unroll
                    for (int m = 5; m < NV; m++) if (Un[m] < 0.0) Un[m] = 0.0;     // limit_vt
                    double DF[NV];
                    comput_dfn<NV>(DF, Un, DUn, b.nv[3 * dir][s], b.nv[3 * dir + 1][s], b.nv[3 * dir + 2][s], gamma);
                    const double wLc = wlu * Lcn;
                    if (FWD) {
                        const double Sd = b.S[dir][s];
#pragma unroll
                        for (int m = 0; m < NV; m++) du[m] = 0.5 * (DF[m] * Sd + wLc * DUn[m]);
                    } else {
#pragma unroll
                        for (int m = 0; m < NV; m++) du[m] = -0.5 * (DF[m] * Sn - wLc * DUn[m]);
                    }
                    if (visc) {
#pragma unroll
                        for (int m = 0; m < NV; m++) du[m] = du[m] + Lvn * DUn[m];
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < NV; m++) du[m] = 0.0;
                }
#pragma unroll
                for (int m = 0; m < NV; m++) acc[m] = acc[m] + du[m];
            }
#pragma unroll
            for (int m = 0; m < NV; m++) {
                oU[m] = b.U[m][s];
                oDU[m] = FWD ? acc[m] / alfa[m] : b.DU[m][s] + acc[m] / alfa[m];
                b.DU[m][s] = oDU[m];
            }
            oLc1 = Lc1; oLc2 = Lc2; oLv1 = Lv1; oLv2 = Lv2;
            if (!FWD) { oS1 = b.S[1][s]; oS2 = b.S[2][s]; }
            // this cell is the i neighbour of the lane's next cell
#pragma unroll
            for (int m = 0; m < NV; m++) { pU[m] = oU[m]; pDU[m] = oDU[m]; }
            pLc = Lc0; pLv = Lv0; pS = FWD ? 0.0 : b.S[0][s];
        }
        LuSlot<NV>& me = slot[buf][lane];
        if (act) {
#pragma unroll
            for (int m = 0; m < NV; m++) { me.U[m] = oU[m]; me.DU[m] = oDU[m]; }
            me.Lc[0] = oLc1; me.Lc[1] = oLc2; me.Lv[0] = oLv1; me.Lv[1] = oLv2; me.S[0] = oS1; me.S[1] = oS2;
        }
        __syncthreads();
        if (lane == 0) __hip_atomic_store(prog + w, q, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------------------------
// The same schedule with a lighter exchange between tiles.  Only the DU values a downstream tile
// reads (the tile's last j row and last k row) cross workgroups; they are written and read with
// device-coherent 64-bit atomics (relaxed, agent scope), which bypass the per-XCD caches, so no
// L2 write-back or invalidate is needed.  The producer waits for its stores to complete before
// it publishes its progress (every PUB steps); the consumer reads the progress with the same
// kind of atomic.  Everything a step needs that does not change during the sweep is loaded one
// step ahead, and the cross-tile DU of the next step is fetched right after the barrier.

__device__ __forceinline__ void st_coherent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coherent(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// per-step inputs of one lane that do not change during the sweep
template <int NV>
struct LuOwn {
    double R[NV];        // Res (forward) or the forward DU (backward)
    double U[NV];
    double Lc[3], Lv[3], S[3], n[9];
    double vol, dt;
};
// the upstream-tile neighbour of one direction
template <int NV>
struct LuRemote {
    double U[NV], DU[NV];
    double Lc, Lv, S;
};

// du contribution of one upstream neighbour (sub_LU_SGS.f90:42-60 forward, :94-119 backward)
template <int NV, bool FWD>
__device__ __forceinline__ void lu_neighbour(double* du, const double* Un_in, const double* DUn, double Lcn, double Lvn,
                                             double Sn, const LuOwn<NV>& o, int dir, double wlu, double gamma, bool visc) {
    double Un[NV], DF[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) Un[m] = Un_in[m];
#pragma unroll
    for (int m = 5; m < NV; m++) if (Un[m] < 0.0) Un[m] = 0.0;     // limit_vt
    comput_dfn<NV>(DF, Un, DUn, o.n[3 * dir], o.n[3 * dir + 1], o.n[3 * dir + 2], gamma);
    const double wLc = wlu * Lcn;
    if (FWD) {
        const double Sd = o.S[dir];
#pragma unroll
        for (int m = 0; m < NV; m++) du[m] = 0.5 * (DF[m] * Sd + wLc * DUn[m]);
    } else {
#pragma unroll
        for (int m = 0; m < NV; m++) du[m] = -0.5 * (DF[m] * Sn - wLc * DUn[m]);
    }
    if (visc) {
#pragma unroll
        for (int m = 0; m < NV; m++) du[m] = du[m] + Lvn * DUn[m];
    }
}

// Fields of the sweep-ordered storage: for a tile, element (q, lane) holds the cell that lane
// handles at forward step q (the backward sweep visits the same cells at step nsteps-1-q), so
// every step reads and writes 64 consecutive doubles per field.
template <int NV> struct Pk {
    static constexpr int U = 0, R = NV, DU = 2 * NV, LC = 3 * NV, LV = 3 * NV + 3, DT = 3 * NV + 6, VOL = 3 * NV + 7,
                         S = 3 * NV + 8, N = 3 * NV + 11, NF = 3 * NV + 20;
};

template <int NV>
__device__ __forceinline__ double* box_field(const Blk& b, int f) {
    if (f < NV) return b.U[f];
    f -= NV;
    if (f < NV) return b.Res[f];
    f -= NV;
    if (f < NV) return b.DU[f];
    f -= NV;
    if (f < 3) return b.Lc[f];
    f -= 3;
    if (f < 3) return b.Lv[f];
    f -= 3;
    if (f == 0) return b.dt;
    if (f == 1) return b.vol;
    f -= 2;
    if (f < 3) return b.S[f];
    f -= 3;
    return b.nv[f];
}

constexpr int PK_W = 32;      // i-window of one transpose workgroup
constexpr int PK_T = 256;

// box <-> sweep order for fields [f0,f1) and [f2,f3) of one tile and one i-window, through LDS so
// that both sides are read and written in contiguous runs
template <int NV, bool TO_PACKED>
__global__ void __launch_bounds__(PK_T) k_lu_transpose(const Blk* __restrict__ blks, const LuTask* __restrict__ tiles,
                                                       double* __restrict__ pk, long long P, int f0, int f1, int f2, int f3) {
    __shared__ double buf[LU_NT][PK_W + 1];
    const LuTask T = tiles[blockIdx.x];
    const Blk& b = blks[T.blk];
    const int nxm = b.nx - 1;
    const int i0 = 1 + (int)blockIdx.y * PK_W;
    if (i0 > nxm) return;
    const int W = min(PK_W, nxm - i0 + 1);
    const int nsteps = nxm + T.tj + T.tk - 2;
    const int t = threadIdx.x;
    for (int f = f0; f < f3; f++) {
        if (f == f1) { f = f2 - 1; continue; }
        double* box = box_field<NV>(b, f);
        double* pf = pk + (size_t)f * (size_t)P + T.pofs;
        if (TO_PACKED) {
            for (int e = t; e < LU_NT * PK_W; e += PK_T) {
                int r = e / PK_W, c = e % PK_W, jl = r % LU_TJ, kl = r / LU_TJ;
                if (jl < T.tj && kl < T.tk && c < W) buf[r][c] = box[IX(b, i0 + c, T.j0 + jl, T.k0 + kl)];
            }
            __syncthreads();
            for (int e = t; e < (W + LU_TJ + LU_TK - 2) * LU_NT; e += PK_T) {
                int q = i0 - 1 + e / LU_NT, lane = e % LU_NT, jl = lane % LU_TJ, kl = lane / LU_TJ;
                int i = q - jl - kl + 1;
                if (jl < T.tj && kl < T.tk && i >= i0 && i < i0 + W && q < nsteps) pf[(size_t)q * LU_NT + lane] = buf[lane][i - i0];
            }
        } else {
            for (int e = t; e < (W + LU_TJ + LU_TK - 2) * LU_NT; e += PK_T) {
                int q = i0 - 1 + e / LU_NT, lane = e % LU_NT, jl = lane % LU_TJ, kl = lane / LU_TJ;
                int i = q - jl - kl + 1;
                if (jl < T.tj && kl < T.tk && i >= i0 && i < i0 + W && q < nsteps) buf[lane][i - i0] = pf[(size_t)q * LU_NT + lane];
            }
            __syncthreads();
            for (int e = t; e < LU_NT * PK_W; e += PK_T) {
                int r = e / PK_W, c = e % PK_W, jl = r % LU_TJ, kl = r / LU_TJ;
                if (jl < T.tj && kl < T.tk && c < W) box[IX(b, i0 + c, T.j0 + jl, T.k0 + kl)] = buf[r][c];
            }
        }
        __syncthreads();
    }
}

template <int NV, bool FWD>
__device__ __forceinline__ void lu_load_own_pk(const double* pk, long long P, size_t e, LuOwn<NV>& o) {
    using F = Pk<NV>;
#pragma unroll
    for (int m = 0; m < NV; m++) {
        o.R[m] = pk[(size_t)((FWD ? F::R : F::DU) + m) * P + e];
        o.U[m] = pk[(size_t)(F::U + m) * P + e];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        o.Lc[d] = pk[(size_t)(F::LC + d) * P + e];
        o.Lv[d] = pk[(size_t)(F::LV + d) * P + e];
        o.S[d] = pk[(size_t)(F::S + d) * P + e];
    }
#pragma unroll
    for (int q = 0; q < 9; q++) o.n[q] = pk[(size_t)(F::N + q) * P + e];
    o.vol = pk[(size_t)F::VOL * P + e];
    o.dt = pk[(size_t)F::DT * P + e];
}

template <int NV, bool FWD>
__device__ __forceinline__ void lu_load_remote_pk(const double* pk, long long P, size_t e, int dir, LuRemote<NV>& r) {
    using F = Pk<NV>;
#pragma unroll
    for (int m = 0; m < NV; m++) {
        r.U[m] = pk[(size_t)(F::U + m) * P + e];
        r.DU[m] = ld_coherent(pk + (size_t)(F::DU + m) * P + e);
    }
    r.Lc = pk[(size_t)(F::LC + dir) * P + e];
    r.Lv = pk[(size_t)(F::LV + dir) * P + e];
    r.S = FWD ? 0.0 : pk[(size_t)(F::S + dir) * P + e];
}

template <int NV, bool FWD>
__global__ void __launch_bounds__(LU_NT) k_lusgs_pipe(const Blk* __restrict__ blks, const LuTask* __restrict__ tasks,
                                                      double* __restrict__ pk, long long P,
                                                      int* __restrict__ prog, int* __restrict__ prog_other,
                                                      int n_other, int* __restrict__ err, Par p) {
    using F = Pk<NV>;
    __shared__ LuSlot<NV> slot[2][LU_NT];
    __shared__ int bail;
    const int w = blockIdx.x, lane = threadIdx.x;
    if (lane == 0 && w < n_other) prog_other[w] = -1;

    const LuTask T = tasks[w];
    const Blk& b = blks[T.blk];
    const int jl = lane % LU_TJ, kl = lane / LU_TJ;
    const bool live = jl < T.tj && kl < T.tk;
    const int jo = FWD ? jl : T.tj - 1 - jl, ko = FWD ? kl : T.tk - 1 - kl;
    const int j = T.j0 + jl, k = T.k0 + kl;
    const int nxm = b.nx - 1;
    const double wlu = p.w_LU, gamma = p.gamma;
    const bool visc = p.visc == 1;
    const int nsteps = nxm + T.tj + T.tk - 2;
    const bool has_j = FWD ? j != 1 : j != b.ny - 1, has_k = FWD ? k != 1 : k != b.nz - 1;
    const bool rem_j = live && has_j && jo == 0, rem_k = live && has_k && ko == 0;
    const bool pub_lane = live && (jo == T.tj - 1 || ko == T.tk - 1);
    const bool deps = T.dep_j >= 0 || T.dep_k >= 0;
    // the upstream lanes this lane reads across the tile edges
    const int jlu = FWD ? T.tj_up - 1 : 0, klu = FWD ? T.tk_up - 1 : 0;
    int known_j = -1, known_k = -1, got_j = -1, got_k = -1;
    if (lane == 0) bail = 0;

    auto cell_i = [&](int q) { return FWD ? q - jo - ko + 1 : nxm - (q - jo - ko); };
    auto own_e = [&](int q) { return (size_t)T.pofs + (size_t)(FWD ? q : nsteps - 1 - q) * LU_NT + lane; };
    auto need_j = [&](int q) { return T.dep_j >= 0 ? min(q + T.off_j, T.last_j) : -1; };
    auto need_k = [&](int q) { return T.dep_k >= 0 ? min(q + T.off_k, T.last_k) : -1; };
    auto wait_for = [&](int q) {
        long polls = 0;
        while (known_j < need_j(q) || known_k < need_k(q)) {
            if (known_j < need_j(q)) known_j = __hip_atomic_load(prog + T.dep_j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (known_k < need_k(q)) known_k = __hip_atomic_load(prog + T.dep_k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (known_j >= need_j(q) && known_k >= need_k(q)) break;
            if (++polls > LU_POLL_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bail = 1;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };
    auto load_remote = [&](int q, LuRemote<NV>& rj, LuRemote<NV>& rk) {
        const int iq = cell_i(q);
        if (live && iq >= 1 && iq <= nxm && q < nsteps) {
            if (rem_j) lu_load_remote_pk<NV, FWD>(pk, P, (size_t)T.pofs_j + (size_t)(iq - 1 + jlu + kl) * LU_NT + jlu + LU_TJ * kl, 1, rj);
            if (rem_k) lu_load_remote_pk<NV, FWD>(pk, P, (size_t)T.pofs_k + (size_t)(iq - 1 + jl + klu) * LU_NT + jl + LU_TJ * klu, 2, rk);
        }
    };

    LuOwn<NV> own, nxt;
    LuRemote<NV> rj, rk, rjn, rkn;
    double pU[NV], pDU[NV], pLc = 0.0, pLv = 0.0, pS = 0.0;
#pragma unroll
    for (int m = 0; m < NV; m++) { pU[m] = 0.0; pDU[m] = 0.0; }
    {
        int i0 = cell_i(0);
        if (live && i0 >= 1 && i0 <= nxm) lu_load_own_pk<NV, FWD>(pk, P, own_e(0), own);
        if (deps && lane == 0) wait_for(min(1, nsteps - 1));
        __syncthreads();
        if (bail) return;
        load_remote(0, rj, rk);
    }

    for (int q = 0; q < nsteps; q++) {
        const int buf = q & 1;
        const int i = cell_i(q), i1 = cell_i(q + 1);
        const bool act = live && i >= 1 && i <= nxm;
        const bool act1 = live && i1 >= 1 && i1 <= nxm && q + 1 < nsteps;
        if (act1) lu_load_own_pk<NV, FWD>(pk, P, own_e(q + 1), nxt);
        load_remote(q + 1, rjn, rkn);
        const bool look = deps && lane == 0 && q + 2 < nsteps && (known_j < need_j(q + 2) || known_k < need_k(q + 2));
        if (look) {
            if (T.dep_j >= 0) got_j = __hip_atomic_load(prog + T.dep_j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (T.dep_k >= 0) got_k = __hip_atomic_load(prog + T.dep_k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        double oDU[NV];
        if (act) {
            double sumLc = (own.Lc[0] + own.Lc[1]) + own.Lc[2];
            double sumLv = (own.Lv[0] + own.Lv[1]) + own.Lv[2];
            double alfa0 = own.vol / own.dt + wlu * sumLc;
            double alfa[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) alfa[m] = alfa0;
            if (visc) {
#pragma unroll
                for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 2.0 * sumLv;
                if (NV == 6) {
                    if (FWD) alfa[5] = alfa[5] + sumLv;
                    else alfa[5] = ((alfa[5] + own.Lv[0]) + own.Lv[1]) + own.Lv[2];
                }
            }
#pragma unroll
            for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 0.0 * own.vol;
            double du0[NV], du1[NV], du2[NV];
            const bool has_i = FWD ? i != 1 : i != nxm;
            if (has_i) lu_neighbour<NV, FWD>(du0, pU, pDU, pLc, pLv, pS, own, 0, wlu, gamma, visc);
            else {
#pragma unroll
                for (int m = 0; m < NV; m++) du0[m] = 0.0;
            }
            // (indices clamped: they are only read when the neighbour lane exists)
            const LuSlot<NV>& oj = slot[buf ^ 1][min(max(FWD ? lane - 1 : lane + 1, 0), LU_NT - 1)];
            const LuSlot<NV>& ok = slot[buf ^ 1][min(max(FWD ? lane - LU_TJ : lane + LU_TJ, 0), LU_NT - 1)];
            if (has_j) {
                if (rem_j) lu_neighbour<NV, FWD>(du1, rj.U, rj.DU, rj.Lc, rj.Lv, rj.S, own, 1, wlu, gamma, visc);
                else lu_neighbour<NV, FWD>(du1, oj.U, oj.DU, oj.Lc[0], oj.Lv[0], oj.S[0], own, 1, wlu, gamma, visc);
            } else {
#pragma unroll
                for (int m = 0; m < NV; m++) du1[m] = 0.0;
            }
            if (has_k) {
                if (rem_k) lu_neighbour<NV, FWD>(du2, rk.U, rk.DU, rk.Lc, rk.Lv, rk.S, own, 2, wlu, gamma, visc);
                else lu_neighbour<NV, FWD>(du2, ok.U, ok.DU, ok.Lc[1], ok.Lv[1], ok.S[1], own, 2, wlu, gamma, visc);
            } else {
#pragma unroll
                for (int m = 0; m < NV; m++) du2[m] = 0.0;
            }
#pragma unroll
            for (int m = 0; m < NV; m++) {
                double acc = FWD ? own.R[m] : 0.0;
                acc = ((acc + du0[m]) + du1[m]) + du2[m];
                oDU[m] = FWD ? acc / alfa[m] : own.R[m] + acc / alfa[m];
            }
        }
        // every memory operation issued before this point is complete: the previous step's DU
        // stores, and this step's loads.  Publish the previous step, then store this one.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0 && q >= 1) __hip_atomic_store(prog + w, q - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (act) {
            const size_t e = own_e(q);
#pragma unroll
            for (int m = 0; m < NV; m++) {
                double* dst = pk + (size_t)(F::DU + m) * P + e;
                if (pub_lane) st_coherent(dst, oDU[m]);
                else *dst = oDU[m];
            }
            LuSlot<NV>& me = slot[buf][lane];
#pragma unroll
            for (int m = 0; m < NV; m++) { me.U[m] = own.U[m]; me.DU[m] = oDU[m]; }
            me.Lc[0] = own.Lc[1]; me.Lc[1] = own.Lc[2]; me.Lv[0] = own.Lv[1]; me.Lv[1] = own.Lv[2];
            me.S[0] = own.S[1]; me.S[1] = own.S[2];
#pragma unroll
            for (int m = 0; m < NV; m++) { pU[m] = own.U[m]; pDU[m] = oDU[m]; }
            pLc = own.Lc[0]; pLv = own.Lv[0]; pS = own.S[0];
        }
        if (look) { known_j = max(known_j, got_j); known_k = max(known_k, got_k); }
        if (deps && lane == 0 && q + 2 < nsteps) wait_for(q + 2);
        __syncthreads();
        if (bail) return;
        own = nxt;
        rj = rjn;
        rk = rkn;
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) __hip_atomic_store(prog + w, nsteps - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------------------
// host side

// Tiles of every block for one sweep direction, numbered so that upstream tiles come first.
// The sweep-ordered storage of a tile does not depend on the direction: tiles are laid out block
// after block in forward order, each taking (nsteps x 64) elements.
std::vector<LuTask> lusgs_tasks(const std::vector<Blk>& blks, bool forward, long long* total) {
    std::vector<LuTask> out;
    long long pbase = 0;
    for (size_t bi = 0; bi < blks.size(); bi++) {
        const Blk& b = blks[bi];
        int nj = b.ny - 1, nk = b.nz - 1;
        int NJt = (nj + LU_TJ - 1) / LU_TJ, NKt = (nk + LU_TK - 1) / LU_TK;
        int base = (int)out.size();
        auto size_j = [&](int J) { return std::min(LU_TJ, nj - J * LU_TJ); };
        auto size_k = [&](int K) { return std::min(LU_TK, nk - K * LU_TK); };
        // storage offsets in forward tile order
        std::vector<long long> pofs((size_t)NJt * NKt);
        for (int K = 0; K < NKt; K++) for (int J = 0; J < NJt; J++) {
            pofs[J + NJt * K] = pbase;
            pbase += (long long)((b.nx - 1) + size_j(J) + size_k(K) - 2) * LU_NT;
        }
        auto P = [&](int J, int K) { return (int)pofs[J + NJt * K]; };
        // position of tile (J,K) in this block's list
        auto idx = [&](int J, int K) { return forward ? J + NJt * K : (NJt - 1 - J) + NJt * (NKt - 1 - K); };
        std::vector<LuTask> blk_tasks((size_t)NJt * NKt);
        for (int K = 0; K < NKt; K++) for (int J = 0; J < NJt; J++) {
            LuTask t{};
            t.blk = (int)bi; t.j0 = 1 + J * LU_TJ; t.k0 = 1 + K * LU_TK; t.tj = size_j(J); t.tk = size_k(K);
            t.pofs = P(J, K);
            int Ju = forward ? J - 1 : J + 1, Ku = forward ? K - 1 : K + 1;
            t.dep_j = t.dep_k = -1;
            if (Ju >= 0 && Ju < NJt) {
                t.dep_j = base + idx(Ju, K);
                t.off_j = size_j(Ju) - 1;
                t.last_j = (b.nx - 1) + size_j(Ju) + t.tk - 3;
                t.pofs_j = P(Ju, K);
                t.tj_up = size_j(Ju);
            }
            if (Ku >= 0 && Ku < NKt) {
                t.dep_k = base + idx(J, Ku);
                t.off_k = size_k(Ku) - 1;
                t.last_k = (b.nx - 1) + t.tj + size_k(Ku) - 3;
                t.pofs_k = P(J, Ku);
                t.tk_up = size_k(Ku);
            }
            blk_tasks[idx(J, K)] = t;
        }
        out.insert(out.end(), blk_tasks.begin(), blk_tasks.end());
    }
    if (total) *total = pbase;
    return out;
}

template <int NV>
static void sweep_nv(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st) {
    if (!fenced) {
        using F = Pk<NV>;
        dim3 tgrid((unsigned)L.n_fwd, (unsigned)L.nwin);
        if (forward) {
            // this iteration's inputs into sweep order: U, Res, Lc, Lv, dt
            hipLaunchKernelGGL((k_lu_transpose<NV, true>), tgrid, dim3(PK_T), 0, st, L.blks, L.fwd, L.pk, L.P,
                               F::U, F::DU, F::LC, F::VOL);
            hipLaunchKernelGGL((k_lusgs_pipe<NV, true>), dim3((unsigned)L.n_fwd), dim3(LU_NT), 0, st,
                               L.blks, L.fwd, L.pk, L.P, L.prog_fwd, L.prog_bwd, L.n_bwd, L.err, p);
        } else {
            hipLaunchKernelGGL((k_lusgs_pipe<NV, false>), dim3((unsigned)L.n_bwd), dim3(LU_NT), 0, st,
                               L.blks, L.bwd, L.pk, L.P, L.prog_bwd, L.prog_fwd, L.n_fwd, L.err, p);
            // DU back into the block layout
            hipLaunchKernelGGL((k_lu_transpose<NV, false>), tgrid, dim3(PK_T), 0, st, L.blks, L.fwd, L.pk, L.P,
                               F::DU, F::LC, F::LC, F::LC);
        }
        return;
    }
    if (forward)
        hipLaunchKernelGGL((k_lusgs_sweep<NV, true>), dim3((unsigned)L.n_fwd), dim3(LU_NT), 0, st,
                           L.blks, L.fwd, L.prog_fwd, L.prog_bwd, L.n_bwd, L.err, p);
    else
        hipLaunchKernelGGL((k_lusgs_sweep<NV, false>), dim3((unsigned)L.n_bwd), dim3(LU_NT), 0, st,
                           L.blks, L.bwd, L.prog_bwd, L.prog_fwd, L.n_fwd, L.err, p);
}

void launch_lusgs_pipelined(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st) {
    if (p.NV == 5) sweep_nv<5>(L, p, forward, fenced, st);
    else if (p.NV == 6) sweep_nv<6>(L, p, forward, fenced, st);
    else sweep_nv<7>(L, p, forward, fenced, st);
}

int lusgs_fields(int NV) { return 3 * NV + 20; }
int lusgs_windows(int max_cells_i) { return (max_cells_i + PK_W - 1) / PK_W; }

// the geometry the sweeps read (vol, S, n), into sweep order once per plan
template <int NV>
static void pack_geometry_nv(const LuPlan& L, hipStream_t st) {
    using F = Pk<NV>;
    hipLaunchKernelGGL((k_lu_transpose<NV, true>), dim3((unsigned)L.n_fwd, (unsigned)L.nwin), dim3(PK_T), 0, st,
                       L.blks, L.fwd, L.pk, L.P, F::VOL, F::NF, F::NF, F::NF);
}

void launch_lusgs_pack_geometry(const LuPlan& L, int NV, hipStream_t st) {
    if (NV == 5) pack_geometry_nv<5>(L, st);
    else if (NV == 6) pack_geometry_nv<6>(L, st);
    else pack_geometry_nv<7>(L, st);
}

}  // namespace ocfd
