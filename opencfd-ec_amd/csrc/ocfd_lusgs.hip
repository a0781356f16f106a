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

template <int NV, bool FWD>
__device__ __forceinline__ void lu_load_own(const Blk& b, size_t s, LuOwn<NV>& o) {
#pragma unroll
    for (int m = 0; m < NV; m++) { o.R[m] = FWD ? b.Res[m][s] : b.DU[m][s]; o.U[m] = b.U[m][s]; }
#pragma unroll
    for (int d = 0; d < 3; d++) { o.Lc[d] = b.Lc[d][s]; o.Lv[d] = b.Lv[d][s]; o.S[d] = b.S[d][s]; }
#pragma unroll
    for (int q = 0; q < 9; q++) o.n[q] = b.nv[q][s];
    o.vol = b.vol[s];
    o.dt = b.dt[s];
}

template <int NV, bool FWD>
__device__ __forceinline__ void lu_load_remote(const Blk& b, size_t t, int dir, LuRemote<NV>& r) {
#pragma unroll
    for (int m = 0; m < NV; m++) { r.U[m] = b.U[m][t]; r.DU[m] = ld_coherent(b.DU[m] + t); }
    r.Lc = b.Lc[dir][t];
    r.Lv = b.Lv[dir][t];
    r.S = FWD ? 0.0 : b.S[dir][t];
}

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

template <int NV, bool FWD, int PUB>
__global__ void __launch_bounds__(LU_NT) k_lusgs_pipe(const Blk* __restrict__ blks, const LuTask* __restrict__ tasks,
                                                      int* __restrict__ prog, int* __restrict__ prog_other,
                                                      int n_other, int* __restrict__ err, Par p) {
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
    const size_t sj = (size_t)b.NI, sk = (size_t)b.NI * b.NJ;
    const double wlu = p.w_LU, gamma = p.gamma;
    const bool visc = p.visc == 1;
    const int nsteps = nxm + T.tj + T.tk - 2;
    // neighbour sources of the j and k directions: none (block face), same tile (LDS) or upstream tile
    const bool has_j = FWD ? j != 1 : j != b.ny - 1, has_k = FWD ? k != 1 : k != b.nz - 1;
    const bool rem_j = live && has_j && jo == 0, rem_k = live && has_k && ko == 0;
    // edge lanes whose DU a downstream tile reads
    const bool pub_lane = live && (jo == T.tj - 1 || ko == T.tk - 1);
    const bool deps = T.dep_j >= 0 || T.dep_k >= 0;
    int known_j = -1, known_k = -1;
    if (lane == 0) bail = 0;

    auto cell_i = [&](int q) { return FWD ? q - jo - ko + 1 : nxm - (q - jo - ko); };
    // lane 0: wait until the upstream tiles have published what step q reads
    auto wait_for = [&](int q) {
        int need_j = T.dep_j >= 0 ? min(q + T.off_j, T.last_j) : -1;
        int need_k = T.dep_k >= 0 ? min(q + T.off_k, T.last_k) : -1;
        long polls = 0;
        while (known_j < need_j || known_k < need_k) {
            if (known_j < need_j) known_j = __hip_atomic_load(prog + T.dep_j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (known_k < need_k) known_k = __hip_atomic_load(prog + T.dep_k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (known_j >= need_j && known_k >= need_k) break;
            if (++polls > LU_POLL_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bail = 1;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };

    LuOwn<NV> own, nxt;
    LuRemote<NV> rj, rk;
    double pU[NV], pDU[NV], pLc = 0.0, pLv = 0.0, pS = 0.0;
#pragma unroll
    for (int m = 0; m < NV; m++) { pU[m] = 0.0; pDU[m] = 0.0; }

    // step 0 inputs
    {
        int i0 = cell_i(0);
        if (live && i0 >= 1 && i0 <= nxm) lu_load_own<NV, FWD>(b, IX(b, i0, j, k), own);
        if (deps && lane == 0) wait_for(0);
        __syncthreads();
        if (bail) return;
        if (live && i0 >= 1 && i0 <= nxm) {
            size_t s = IX(b, i0, j, k);
            if (rem_j) lu_load_remote<NV, FWD>(b, FWD ? s - sj : s + sj, 1, rj);
            if (rem_k) lu_load_remote<NV, FWD>(b, FWD ? s - sk : s + sk, 2, rk);
        }
    }

    for (int q = 0; q < nsteps; q++) {
        const int buf = q & 1;
        const int i = cell_i(q), i1 = cell_i(q + 1);
        const bool act = live && i >= 1 && i <= nxm;
        const bool act1 = live && i1 >= 1 && i1 <= nxm && q + 1 < nsteps;
        // next step's static inputs
        if (act1) lu_load_own<NV, FWD>(b, IX(b, i1, j, k), nxt);
        double oDU[NV];
        if (act) {
            const size_t s = IX(b, i, j, k);
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
            // neighbours from this lane and this tile first, the upstream tile's last
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
            if (has_j && !rem_j) lu_neighbour<NV, FWD>(du1, oj.U, oj.DU, oj.Lc[0], oj.Lv[0], oj.S[0], own, 1, wlu, gamma, visc);
            if (has_k && !rem_k) lu_neighbour<NV, FWD>(du2, ok.U, ok.DU, ok.Lc[1], ok.Lv[1], ok.S[1], own, 2, wlu, gamma, visc);
            if (rem_j) lu_neighbour<NV, FWD>(du1, rj.U, rj.DU, rj.Lc, rj.Lv, rj.S, own, 1, wlu, gamma, visc);
            if (rem_k) lu_neighbour<NV, FWD>(du2, rk.U, rk.DU, rk.Lc, rk.Lv, rk.S, own, 2, wlu, gamma, visc);
            if (!has_j) {
#pragma unroll
                for (int m = 0; m < NV; m++) du1[m] = 0.0;
            }
            if (!has_k) {
#pragma unroll
                for (int m = 0; m < NV; m++) du2[m] = 0.0;
            }
#pragma unroll
            for (int m = 0; m < NV; m++) {
                double acc = FWD ? own.R[m] : 0.0;
                acc = ((acc + du0[m]) + du1[m]) + du2[m];
                oDU[m] = FWD ? acc / alfa[m] : own.R[m] + acc / alfa[m];
                if (pub_lane) st_coherent(b.DU[m] + s, oDU[m]);
                else b.DU[m][s] = oDU[m];
            }
            LuSlot<NV>& me = slot[buf][lane];
#pragma unroll
            for (int m = 0; m < NV; m++) { me.U[m] = own.U[m]; me.DU[m] = oDU[m]; }
            me.Lc[0] = own.Lc[1]; me.Lc[1] = own.Lc[2]; me.Lv[0] = own.Lv[1]; me.Lv[1] = own.Lv[2];
            me.S[0] = own.S[1]; me.S[1] = own.S[2];
            // this cell is the i neighbour of the lane's next cell
#pragma unroll
            for (int m = 0; m < NV; m++) { pU[m] = own.U[m]; pDU[m] = oDU[m]; }
            pLc = own.Lc[0]; pLv = own.Lv[0]; pS = own.S[0];
        }
        // publish: the edge stores of this wave complete before the counter moves
        if ((q + 1) % PUB == 0 || q + 1 == nsteps) {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (lane == 0) __hip_atomic_store(prog + w, q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (q + 1 < nsteps && deps && lane == 0) wait_for(q + 1);
        __syncthreads();
        if (bail) return;
        if (act1) {
            const size_t s1 = IX(b, i1, j, k);
            if (rem_j) lu_load_remote<NV, FWD>(b, FWD ? s1 - sj : s1 + sj, 1, rj);
            if (rem_k) lu_load_remote<NV, FWD>(b, FWD ? s1 - sk : s1 + sk, 2, rk);
            own = nxt;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// host side

// Tiles of every block for one sweep direction, numbered so that upstream tiles come first.
std::vector<LuTask> lusgs_tasks(const std::vector<Blk>& blks, bool forward) {
    std::vector<LuTask> out;
    for (size_t bi = 0; bi < blks.size(); bi++) {
        const Blk& b = blks[bi];
        int nj = b.ny - 1, nk = b.nz - 1;
        int NJt = (nj + LU_TJ - 1) / LU_TJ, NKt = (nk + LU_TK - 1) / LU_TK;
        int base = (int)out.size();
        auto size_j = [&](int J) { return std::min(LU_TJ, nj - J * LU_TJ); };
        auto size_k = [&](int K) { return std::min(LU_TK, nk - K * LU_TK); };
        // position of tile (J,K) in this block's list
        auto idx = [&](int J, int K) { return forward ? J + NJt * K : (NJt - 1 - J) + NJt * (NKt - 1 - K); };
        std::vector<LuTask> blk_tasks((size_t)NJt * NKt);
        for (int K = 0; K < NKt; K++) for (int J = 0; J < NJt; J++) {
            LuTask t{};
            t.blk = (int)bi; t.j0 = 1 + J * LU_TJ; t.k0 = 1 + K * LU_TK; t.tj = size_j(J); t.tk = size_k(K);
            int Ju = forward ? J - 1 : J + 1, Ku = forward ? K - 1 : K + 1;
            t.dep_j = t.dep_k = -1;
            if (Ju >= 0 && Ju < NJt) {
                t.dep_j = base + idx(Ju, K);
                t.off_j = size_j(Ju) - 1;
                t.last_j = (b.nx - 1) + size_j(Ju) + t.tk - 3;
            }
            if (Ku >= 0 && Ku < NKt) {
                t.dep_k = base + idx(J, Ku);
                t.off_k = size_k(Ku) - 1;
                t.last_k = (b.nx - 1) + t.tj + size_k(Ku) - 3;
            }
            blk_tasks[idx(J, K)] = t;
        }
        out.insert(out.end(), blk_tasks.begin(), blk_tasks.end());
    }
    return out;
}

template <int NV>
static void sweep_nv(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st) {
    if (!fenced) {
        constexpr int PUB = 1;
        if (forward)
            hipLaunchKernelGGL((k_lusgs_pipe<NV, true, PUB>), dim3((unsigned)L.n_fwd), dim3(LU_NT), 0, st,
                               L.blks, L.fwd, L.prog_fwd, L.prog_bwd, L.n_bwd, L.err, p);
        else
            hipLaunchKernelGGL((k_lusgs_pipe<NV, false, PUB>), dim3((unsigned)L.n_bwd), dim3(LU_NT), 0, st,
                               L.blks, L.bwd, L.prog_bwd, L.prog_fwd, L.n_fwd, L.err, p);
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

}  // namespace ocfd
