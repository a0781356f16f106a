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

constexpr int LU_TJ = 8, LU_TK = 8;
static_assert(LU_TJ * LU_TK == LU_NT, "the fenced sweep's tiles are 8 x 8");
constexpr long LU_POLL_LIMIT = 1L << 26;

// err[0]: the error flag a wait sets when it gives up.  err[2]: nonzero makes every wait that
// finds its producer behind give up at its first poll (the host's failure-path test).  The poll
// limit itself stays a compile-time constant: a limit read at run time changes the register
// allocation of the team sweep and costs it a sixth of its speed.
__device__ __forceinline__ bool lu_fail_waits(const int* err) {
    return __hip_atomic_load(err + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

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
    __shared__ int gave_up;
    const int w = blockIdx.x, lane = threadIdx.x;
    // this launch re-arms the counters of the other sweep (same tile count) for its next launch
    if (lane == 0 && w < n_other) prog_other[w] = -1;
    // a sweep of this iteration (or an earlier one of the same launch sequence) gave up: nothing to do
    if (lane == 0) gave_up = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (gave_up != 0) return;

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
                if (++polls > LU_POLL_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                    lu_fail_waits(err)) {
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
// Team sweep: the same tiles and the same in-tile hyperplanes, three waves per tile.  Wave d
// (d = 0, 1, 2 for i, j, k) owns the contribution of the upstream neighbour in direction d.  At
// step q each wave first completes the DU of the cell it reads, the one its lane's neighbour
// computed at step q - 1 (its own previous cell for d = 0, the lane across the j or k step for
// d = 1, 2): DU = (((Res + du_i) + du_j) + du_k) / alfa in the reference's order (backward sweep:
// DU + ((du_i + du_j) + du_k) / alfa) from the three waves' terms of step q - 1 in LDS.  It then
// evaluates the neighbour term of its direction for its own cell and writes it to LDS.  One
// workgroup barrier per step separates the terms of a step from their use in the next; the LDS
// terms are double-buffered by step parity.  Wave 0 stores each completed DU of its own cell.
//
// Cross-tile values travel through the sweep-ordered DU arrays themselves.  Before a sweep every
// element of the array it writes holds a sentinel NaN (a payload no arithmetic produces); the
// producer overwrites an element read by a neighbour tile with a device-coherent 64-bit store and
// the consumer reloads it, device-coherent, until the sentinel is gone.  No progress counter and
// no wait for store completion sit on the producer's path.  The forward sweep writes DUF and the
// backward sweep DUB; the pack of the next sweeps' inputs puts the sentinel back into both.
// Tiles are numbered upstream first and workgroups start in index order, so every wait is on a
// tile that is running or done; every wait is also bounded by a poll limit that raises the error
// flag.

constexpr int LT_NT = 3 * LU_NT;
__device__ __forceinline__ void st_coherent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coherent(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}


// value of field f of the sweep-ordered storage at box index s (not for DUF/DUB, whose packing
// puts the sentinel into them)
template <int NV>
__device__ __forceinline__ double pack_value(const Blk& b, int f, size_t s) {
    using F = Pk<NV>;
    if (f < F::R) return b.U[f][s];
    if (f < F::DUF) return b.Res[f - F::R][s];
    if (f < F::LV) return b.Lc[f - F::LC][s];
    if (f < F::AL) return b.Lv[f - F::LV][s];
    if (f < F::S) return b.al[f - F::AL][s];
    if (f < F::N) return b.S[f - F::S][s];
    return b.nv[f - F::N][s];
}

constexpr int PK_T = 256;

// box -> sweep order for fields [f0,f1) and then from f2 on (one per grid z); one tile, one
// i-window and one field per workgroup, through LDS so both sides move in contiguous runs.  Neighbouring i-windows of a tile
// share the step rows at their boundary (each writes some lanes of them), so a tile's windows run
// back to back on one XCD, where the L2 merges those partial lines before they go to memory.
// A DU field's workgroup reads nothing and puts the sentinel into the elements of its window's
// cells that a downstream tile may wait on, in whole lines: a line written in part costs a
// read-modify-write of the memory beneath it (storing only the edge lanes of each step, before
// they led the run, made the pack of the 60.8 M-cell swept wing 3.0 ms slower, against 1.4 ms
// for whole steps; the U workgroups storing the sentinel beside each U, instead of workgroups of
// their own, 1.7 ms).
template <int NV, int TJ>
__global__ void __launch_bounds__(PK_T) k_lu_transpose(const Blk* __restrict__ blks, const LuTask* __restrict__ tiles,
                                                       double* __restrict__ pk, const int2* __restrict__ wins, int f0, int f1, int f2) {
    __shared__ double buf[LU_NT][PK_W + 1];
    const int2 wn = wins[xcd_tile((int)blockIdx.x, (int)gridDim.x)];     // (tile, i-window), windows fastest
    const LuTask T = tiles[wn.x];
    const Blk& b = blks[T.blk];
    const int nxm = b.nx - 1;
    const int i0 = wn.y;
    const int W = min(PK_W, nxm - i0 + 1);
    const int nsteps = nxm + T.tj + T.tk - 2;
    const int t = threadIdx.x;
    constexpr int NF = Pk<NV>::NF;
    const int z = (int)blockIdx.z;
    const int f = z < f1 - f0 ? f0 + z : f2 + (z - (f1 - f0));
    double* pf = pk + (size_t)T.pofs + (size_t)f * LU_NT;     // element (q, f, lane) at pf[q NF 64 + lane]
    if (f >= Pk<NV>::DUF && f < Pk<NV>::DUB + NV) {
        // a full tile's edge lanes lead the run (lane_pos): whole lines up to the last of them;
        // a ragged tile's edges lie elsewhere, so all of its positions
        constexpr int NSENT = (2 * (LU_NT / TJ) + 2 * (TJ - 2) + 15) / 16 * 16;
        const int npos = T.tj == TJ && T.tk == LU_NT / TJ ? NSENT : LU_NT;
        for (int e = t; e < (W + TJ + LU_NT / TJ - 2) * npos; e += PK_T) {
            int q = i0 - 1 + e / npos, pos = e % npos, lane = pos_lane<TJ>(pos), jl = lane % TJ, kl = lane / TJ;
            int i = q - jl - kl + 1;
            if (i >= i0 && i < i0 + W && q < nsteps) pf[(size_t)q * NF * LU_NT + pos] = sentinel_value();
        }
        return;
    }
    for (int e = t; e < LU_NT * PK_W; e += PK_T) {
        int r = e / PK_W, c = e % PK_W, jl = r % TJ, kl = r / TJ;
        if (jl < T.tj && kl < T.tk && c < W) buf[r][c] = pack_value<NV>(b, f, IX(b, i0 + c, T.j0 + jl, T.k0 + kl));
    }
    __syncthreads();
    for (int e = t; e < (W + TJ + LU_NT / TJ - 2) * LU_NT; e += PK_T) {
        int q = i0 - 1 + e / LU_NT, lane = e % LU_NT, jl = lane % TJ, kl = lane / TJ;
        int i = q - jl - kl + 1;
        if (jl < T.tj && kl < T.tk && i >= i0 && i < i0 + W && q < nsteps) pf[(size_t)q * NF * LU_NT + lane_pos<TJ>(lane)] = buf[lane][i - i0];
    }
}

// sweep order -> block arrays for the backward sweep's DU (component m per grid z), for a host that
// reads DU: the storage holds it from the end of the backward sweep until the next pack
template <int NV, int TJ>
__global__ void __launch_bounds__(PK_T) k_lu_du_to_box(const Blk* __restrict__ blks, const LuTask* __restrict__ tiles,
                                                       const double* __restrict__ pk, const int2* __restrict__ wins) {
    __shared__ double buf[LU_NT][PK_W + 1];
    const int2 wn = wins[xcd_tile((int)blockIdx.x, (int)gridDim.x)];
    const LuTask T = tiles[wn.x];
    const Blk& b = blks[T.blk];
    const int nxm = b.nx - 1;
    const int i0 = wn.y;
    const int W = min(PK_W, nxm - i0 + 1);
    const int nsteps = nxm + T.tj + T.tk - 2;
    const int t = threadIdx.x, m = (int)blockIdx.z;
    constexpr int NF = Pk<NV>::NF;
    const double* pf = pk + (size_t)T.pofs + (size_t)(Pk<NV>::DUB + m) * LU_NT;
    for (int e = t; e < (W + TJ + LU_NT / TJ - 2) * LU_NT; e += PK_T) {
        int q = i0 - 1 + e / LU_NT, lane = e % LU_NT, jl = lane % TJ, kl = lane / TJ;
        int i = q - jl - kl + 1;
        if (jl < T.tj && kl < T.tk && i >= i0 && i < i0 + W && q < nsteps) buf[lane][i - i0] = pf[(size_t)q * NF * LU_NT + lane_pos<TJ>(lane)];
    }
    __syncthreads();
    double* box = b.DU[m];
    for (int e = t; e < LU_NT * PK_W; e += PK_T) {
        int r = e / PK_W, c = e % PK_W, jl = r % TJ, kl = r / TJ;
        if (jl < T.tj && kl < T.tk && c < W) box[IX(b, i0 + c, T.j0 + jl, T.k0 + kl)] = buf[r][c];
    }
}

// the sentinel into every DUF and DUB element (and the padding)
template <int NV>
__global__ void k_fill_sentinel(double* p, long long n) {
    using F = Pk<NV>;
    const long long pad = (long long)LU_PAD_STEPS * F::NF * LU_NT;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int f = (int)(((e - pad) / LU_NT) % F::NF);
        if (e < pad || e >= n - pad || (f >= F::DUF && f < F::DUB + NV)) p[e] = sentinel_value();
    }
}


// inputs of one step of wave D: its own cell's face normal and (forward) area in direction D;
// the neighbour's U, Lc, Lv and (backward) S; and the Res (forward) or forward DU (backward) with
// the diagonal of the cell whose DU the wave completes at this step
template <int NV, bool FWD, int D>
struct TeamIn {
    double n[3], S;
    double U[NV], Lc, Lv, Sn;
    double R[NV], al[2];
};

// eo: own element of the step, en: the neighbour's, es: the element of the cell whose DU is
// completed; field f of an element at e[f * 64]
template <int NV, bool FWD, int D>
__device__ __forceinline__ void team_load(const double* __restrict__ eo, const double* __restrict__ en,
                                          const double* __restrict__ es, TeamIn<NV, FWD, D>& t) {
    using F = Pk<NV>;
#pragma unroll
    for (int c = 0; c < 3; c++) t.n[c] = eo[(F::N + 3 * D + c) * LU_NT];
    t.S = FWD ? eo[(F::S + D) * LU_NT] : 0.0;
#pragma unroll
    for (int m = 0; m < NV; m++) t.U[m] = en[(F::U + m) * LU_NT];
    t.Lc = en[(F::LC + D) * LU_NT];
    t.Lv = en[(F::LV + D) * LU_NT];
    t.Sn = FWD ? 0.0 : en[(F::S + D) * LU_NT];
#pragma unroll
    for (int m = 0; m < NV; m++) t.R[m] = es[((FWD ? F::R : F::DUF) + m) * LU_NT];
    t.al[0] = es[F::AL * LU_NT];
    t.al[1] = NV == 6 ? es[(F::AL + (FWD ? 1 : 2)) * LU_NT] : 0.0;
}

// the upstream tile's DU of a cell (may still be the sentinel).  An ordinary load, which the
// L2 of this XCD may serve: within a sweep an element only ever holds the sentinel or its final
// value, and every kernel starts with the L2 invalidated, so a cached copy is either current or a
// sentinel, and a sentinel sends the lane to the device-coherent reloads of team_wait_remote.
// Against device-coherent loads here, the sweeps of a 257x97x33 C-grid take 8 % less time.
template <int NV, bool FWD>
__device__ __forceinline__ void team_load_du(const double* __restrict__ e, double* DU) {
    using F = Pk<NV>;
#pragma unroll
    for (int m = 0; m < NV; m++) DU[m] = e[((FWD ? F::DUF : F::DUB) + m) * LU_NT];
}

// du of one upstream neighbour (sub_LU_SGS.f90:42-60 forward, :94-119 backward)
template <int NV, bool FWD>
__device__ __forceinline__ void team_neighbour(double* du, const double* Un_in, const double* DUn, double Lcn, double Lvn,
                                               double Sn, const double* n, double Sd, double wlu, double gamma, bool visc) {
    double Un[NV], DF[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) Un[m] = Un_in[m];
#pragma unroll
    for (int m = 5; m < NV; m++) if (Un[m] < 0.0) Un[m] = 0.0;     // limit_vt
    comput_dfn_nr<NV>(DF, Un, DUn, n[0], n[1], n[2], gamma);
    const double wLc = wlu * Lcn;
    if (FWD) {
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

// reload an upstream DU until the producer has replaced the sentinel; true when the wait gives up
template <int NV, bool FWD>
__device__ __forceinline__ bool team_wait_remote(const double* __restrict__ er, double* DU, int* __restrict__ err) {
    using F = Pk<NV>;
    long polls = 0;
    while (true) {
        bool ready = true;
#pragma unroll
        for (int m = 0; m < NV; m++) ready = ready && !is_sentinel(DU[m]);
        if (ready) return false;
        if (++polls > LU_POLL_LIMIT ||
            ((polls & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) ||
            (polls == 1 && lu_fail_waits(err))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int m = 0; m < NV; m++)
            if (is_sentinel(DU[m])) DU[m] = ld_coherent(er + ((FWD ? F::DUF : F::DUB) + m) * LU_NT);
    }
}

// a for the lanes whose bit of mk is set, else b: one v_cndmask per half, so the choice cannot
// become a masked overwrite of a register that a load may still be filling
__device__ __forceinline__ double pick(unsigned long long mk, double a, double b) {
    const unsigned long long ua = (unsigned long long)__double_as_longlong(a), ub = (unsigned long long)__double_as_longlong(b);
    unsigned lo, hi;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(lo) : "v"((unsigned)ub), "v"((unsigned)ua), "s"(mk));
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(hi) : "v"((unsigned)(ub >> 32)), "v"((unsigned)(ua >> 32)), "s"(mk));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ void team_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// LDS of a tile team: the records of the previous step (double-buffered by step parity) that
// the downstream lanes read, one array per value so a wave's 64 lanes touch 64 consecutive
// doubles, and the three waves' neighbour terms of the current step
template <int NV>
struct TeamShared {
    double dus[2][3][NV][LU_NT];      // [step parity][direction][component][lane]
    int bail;
};

// One wave's part of a tile's sweep.  The loop is unrolled by three with three register sets for
// the prefetched inputs, so a step's loads are issued two full steps before their use; an upstream
// tile's DU is loaded one step before its use.  Every global load and store is issued by all
// lanes (lanes without a cell address elements nobody reads, inside the tile's storage or the
// buffer's padding), which keeps the wait counts the compiler derives independent of the lane
// masks.  Each lane walks its streams of elements with one pointer each, advanced by one step of
// NF x 64 elements per step.
template <int NV, bool FWD, int D, int TJ>
__device__ void team_sweep(TeamShared<NV>& sh, const Blk& b, const LuTask& T, double* __restrict__ pk,
                           int* __restrict__ err, const Par& p, int lane) {
    using F = Pk<NV>;
    constexpr long STEP = (long)F::NF * LU_NT;                 // elements per step of a tile
    const int jl = lane % TJ, kl = lane / TJ;
    const bool live = jl < T.tj && kl < T.tk;
    const int jo = FWD ? jl : T.tj - 1 - jl, ko = FWD ? kl : T.tk - 1 - kl;
    const int j = T.j0 + jl, k = T.k0 + kl;
    const int nxm = b.nx - 1;
    const double wlu = p.w_LU, gamma = p.gamma;
    const bool visc = p.visc == 1;
    const int nsteps = nxm + T.tj + T.tk - 2;
    // the neighbour of direction D: another lane of the tile, or the upstream tile
    const bool has_d = D == 0 ? true : D == 1 ? (FWD ? j != 1 : j != b.ny - 1) : (FWD ? k != 1 : k != b.nz - 1);
    const bool inside = D == 0 ? true : D == 1 ? jo > 0 : ko > 0;
    const bool remote = D > 0 && live && has_d && !inside;
    const unsigned long long rmask = __builtin_amdgcn_ballot_w64(remote);
    // the lane whose cell of the previous step this lane reads (lanes without one read their own)
    const int src = D == 0 || !inside ? lane : D == 1 ? (FWD ? lane - 1 : lane + 1) : (FWD ? lane - TJ : lane + TJ);

    // cell i of this lane at step q is i0 + dir q
    const int i0 = FWD ? 1 - jo - ko : nxm + jo + ko;
    const int dir = FWD ? 1 : -1;
    // this lane's element at step 0 and per step
    const long base0 = (long)T.pofs + (FWD ? 0L : (long)(nsteps - 1) * STEP);
    const long own0 = base0 + lane_pos<TJ>(lane);
    const long dstep = FWD ? STEP : -STEP;
    // the element of lane src one step back
    const long src0 = base0 - dstep + lane_pos<TJ>(src);
    // the upstream element: the upstream tile stores by forward step, the lane across the edge is
    // jlu + 8 kl (j) or jl + 8 klu (k); lanes without one follow their own stream
    long rem0 = own0;
    if (remote) {
        const int jlu = FWD ? T.tj_up - 1 : 0, klu = FWD ? T.tk_up - 1 : 0;
        const long long pofs_r = D == 1 ? T.pofs_j : T.pofs_k;
        const int lane_r = D == 1 ? jlu + TJ * kl : jl + TJ * klu;
        // forward step of the upstream cell at this lane's step 0 (forward: i - 1 + jl_r + kl_r)
        const int s0 = FWD ? (D == 1 ? jlu : klu) : nxm + (D == 1 ? T.tk : T.tj) - 2;
        rem0 = (long)pofs_r + (long)s0 * STEP + lane_pos<TJ>(lane_r);
    }
    const long nb0 = remote ? rem0 : src0;
    auto own_p = [&](int q) { return pk + own0 + (long)q * dstep; };
    auto src_p = [&](int q) { return pk + src0 + (long)q * dstep; };
    auto nb_p = [&](int q) { return pk + nb0 + (long)q * dstep; };
    auto rem_p = [&](int q) { return pk + rem0 + (long)q * dstep; };
    auto load = [&](int q, TeamIn<NV, FWD, D>& t) { team_load<NV, FWD, D>(own_p(q), nb_p(q), src_p(q), t); };
    auto active = [&](int q) { const int i = i0 + dir * q; return live && q < nsteps && i >= 1 && i <= nxm; };

    // DU of the cell of lane src at step q - 1, from the terms of that step
    auto complete = [&](int q, const TeamIn<NV, FWD, D>& in, double* DUs) {
        const int pb = (q & 1) ^ 1;
        // components 1-5 (and 7) share one diagonal, the SA component has its own
        const double r0 = recip_nr(in.al[0]), r1 = NV == 6 ? recip_nr(in.al[1]) : 0.0;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            const bool own = NV == 6 && m == 5;
            const double al = in.al[own ? 1 : 0], ra = own ? r1 : r0;
            double acc = FWD ? in.R[m] : 0.0;
            acc = ((acc + sh.dus[pb][0][m][src]) + sh.dus[pb][1][m][src]) + sh.dus[pb][2][m][src];
            DUs[m] = FWD ? div_nr(acc, al, ra) : in.R[m] + div_nr(acc, al, ra);
        }
    };
    // wave 0 stores the DU of its own cell of step q - 1 (device-coherent: the elements a
    // neighbour tile reads need it, and one kind of store for all lanes)
    auto store = [&](int q, const double* DUs) {
        if (D == 0 && q >= 1) {
            double* e = own_p(q - 1);
#pragma unroll
            for (int m = 0; m < NV; m++) st_coherent(e + ((FWD ? F::DUF : F::DUB) + m) * LU_NT, DUs[m]);
        }
    };

    // one step: complete the read cell's DU, then this direction's term of the own cell, a
    // barrier; with check, true when the workgroup gives up.  Every third step checks (an LDS
    // read after the barrier that the other steps skip): a wave that gave up has set the error
    // flag, so the waits of the steps in between give up within 256 polls and their values are
    // discarded
    auto step = [&](int q, const TeamIn<NV, FWD, D>& in, double* rdu, bool check) -> bool {
        double DUs[NV];
        complete(q, in, DUs);
        store(q, DUs);
        const int i = i0 + dir * q;
        if (active(q)) {
            const bool has = D == 0 ? (FWD ? i != 1 : i != nxm) : has_d;
            if (D > 0 && remote) {
                // wait off the common path, so the wait counts in front of it do not cover the
                // loads issued for later steps
                bool ready = true;
#pragma unroll
                for (int m = 0; m < NV; m++) ready = ready && !is_sentinel(rdu[m]);
                if (!ready && team_wait_remote<NV, FWD>(rem_p(q), rdu, err)) sh.bail = 1;
            }
            double DUn[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) DUn[m] = D > 0 ? pick(rmask, rdu[m], DUs[m]) : DUs[m];
            double du[NV];
            team_neighbour<NV, FWD>(du, in.U, DUn, in.Lc, in.Lv, in.Sn, in.n, in.S, wlu, gamma, visc);
#pragma unroll
            for (int m = 0; m < NV; m++) sh.dus[q & 1][D][m][lane] = has ? du[m] : 0.0;
        }
        team_barrier();
        return check && sh.bail != 0;
    };

    TeamIn<NV, FWD, D> iA, iB, iC;
    double rA[NV], rB[NV], rC[NV];
    load(0, iA);
    load(1, iB);
    load(2, iC);
    if constexpr (D > 0) team_load_du<NV, FWD>(rem_p(0), rA);
    for (int q = 0; q < nsteps; q += 3) {
        if constexpr (D > 0) team_load_du<NV, FWD>(rem_p(q + 1), rB);
        if (step(q, iA, rA, false)) return;
        load(q + 3, iA);
        if (q + 1 >= nsteps) break;
        if constexpr (D > 0) team_load_du<NV, FWD>(rem_p(q + 2), rC);
        if (step(q + 1, iB, rB, false)) return;
        load(q + 4, iB);
        if (q + 2 >= nsteps) break;
        if constexpr (D > 0) team_load_du<NV, FWD>(rem_p(q + 3), rA);
        if (step(q + 2, iC, rC, true)) return;
        load(q + 5, iC);
    }
    // the DU of the last step
    if constexpr (D == 0) {
        TeamIn<NV, FWD, D> last;
        load(nsteps, last);
        double DUs[NV];
        complete(nsteps, last, DUs);
        store(nsteps, DUs);
    }
}

template <int NV, bool FWD, int TJ>
__global__ void __launch_bounds__(LT_NT) k_lusgs_team(const Blk* __restrict__ blks, const LuTask* __restrict__ tasks,
                                                      double* __restrict__ pk, int* __restrict__ err, Par p) {
    __shared__ TeamShared<NV> sh;
    const int wave = threadIdx.x / LU_NT, lane = threadIdx.x % LU_NT;
    // a sweep earlier in the launch sequence gave up waiting: this one does nothing either (its
    // DU stays unused, the update kernels leave U as it is)
    if (threadIdx.x == 0) sh.bail = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (sh.bail != 0) return;
    const LuTask T = tasks[blockIdx.x];
    const Blk& b = blks[T.blk];
    if (wave == 0) team_sweep<NV, FWD, 0, TJ>(sh, b, T, pk, err, p, lane);
    else if (wave == 1) team_sweep<NV, FWD, 1, TJ>(sh, b, T, pk, err, p, lane);
    else team_sweep<NV, FWD, 2, TJ>(sh, b, T, pk, err, p, lane);
}

// ---------------------------------------------------------------------------------------------
// host side

// Tiles of every block for one sweep direction, numbered so that upstream tiles come first.
// The sweep-ordered storage of a tile does not depend on the direction: tiles are laid out block
// after block in forward order, each taking (nsteps x 64) elements.
std::vector<LuTask> lusgs_tasks(const std::vector<Blk>& blks, bool forward, int NF, int TJ, long long* total) {
    const int TK = LU_NT / TJ;
    std::vector<LuTask> out;
    const long long pad = (long long)LU_PAD_STEPS * NF * LU_NT;
    long long pbase = pad;
    for (size_t bi = 0; bi < blks.size(); bi++) {
        const Blk& b = blks[bi];
        int nj = b.ny - 1, nk = b.nz - 1;
        int NJt = (nj + TJ - 1) / TJ, NKt = (nk + TK - 1) / TK;
        int base = (int)out.size();
        auto size_j = [&](int J) { return std::min(TJ, nj - J * TJ); };
        auto size_k = [&](int K) { return std::min(TK, nk - K * TK); };
        // storage offsets in forward tile order
        std::vector<long long> pofs((size_t)NJt * NKt);
        for (int K = 0; K < NKt; K++) for (int J = 0; J < NJt; J++) {
            pofs[J + NJt * K] = pbase;
            pbase += (long long)((b.nx - 1) + size_j(J) + size_k(K) - 2) * NF * LU_NT;
        }
        auto P = [&](int J, int K) { return pofs[J + NJt * K]; };
        // position of tile (J,K) in this block's list
        auto idx = [&](int J, int K) { return forward ? J + NJt * K : (NJt - 1 - J) + NJt * (NKt - 1 - K); };
        std::vector<LuTask> blk_tasks((size_t)NJt * NKt);
        for (int K = 0; K < NKt; K++) for (int J = 0; J < NJt; J++) {
            LuTask t{};
            t.blk = (int)bi; t.j0 = 1 + J * TJ; t.k0 = 1 + K * TK; t.tj = size_j(J); t.tk = size_k(K);
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
    if (total) *total = pbase + pad;
    return out;
}

template <int NV, int TJ>
static void sweep_nv(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st) {
    if (!fenced) {
        if (forward) {
            hipLaunchKernelGGL((k_lusgs_team<NV, true, TJ>), dim3((unsigned)L.n_fwd), dim3(LT_NT), 0, st,
                               L.blks, L.fwd, L.pk, L.err, p);
        } else {
            // the update takes DU from the sweep-ordered storage
            hipLaunchKernelGGL((k_lusgs_team<NV, false, TJ>), dim3((unsigned)L.n_bwd), dim3(LT_NT), 0, st,
                               L.blks, L.bwd, L.pk, L.err, p);
        }
        return;
    }
    // the fenced sweep's plans always use 8x8 tiles
    if (forward)
        hipLaunchKernelGGL((k_lusgs_sweep<NV, true>), dim3((unsigned)L.n_fwd), dim3(LU_NT), 0, st,
                           L.blks, L.fwd, L.prog_fwd, L.prog_bwd, L.n_bwd, L.err, p);
    else
        hipLaunchKernelGGL((k_lusgs_sweep<NV, false>), dim3((unsigned)L.n_bwd), dim3(LU_NT), 0, st,
                           L.blks, L.bwd, L.prog_bwd, L.prog_fwd, L.n_fwd, L.err, p);
}

// this iteration's sweep inputs into sweep order: U, Res, Lc, Lv and the diagonal (one field, or
// three for SA), and the sentinel into the DU elements the sweeps wait on
template <int NV, int TJ>
static void pack_nv(const LuPlan& L, const Par& p, hipStream_t st) {
    using F = Pk<NV>;
    constexpr int NAL = NV == 6 ? 3 : 1;
    constexpr int FEND = F::AL + NAL;
    hipLaunchKernelGGL((k_lu_transpose<NV, TJ>), dim3((unsigned)L.nwins, 1u, (unsigned)(FEND - F::U)),
                       dim3(PK_T), 0, st, L.blks, L.fwd, L.pk, L.win, F::U, FEND, FEND);
}

// the geometry the sweeps read (S, n), into sweep order once per plan
template <int NV, int TJ>
static void pack_geometry_nv(const LuPlan& L, const Par& p, hipStream_t st) {
    using F = Pk<NV>;
    hipLaunchKernelGGL((k_lu_transpose<NV, TJ>), dim3((unsigned)L.nwins, 1u, (unsigned)(F::NF - F::S)),
                       dim3(PK_T), 0, st, L.blks, L.fwd, L.pk, L.win, F::S, F::NF, F::NF);
}

// NVAR and the plan's tile width as template arguments
template <template <int, int> class K, typename... A>
static void by_shape(int NV, int tj, A&&... a) {
    switch (NV * 100 + tj) {
    case 504: K<5, 4>::run(a...); break;
    case 508: K<5, 8>::run(a...); break;
    case 516: K<5, 16>::run(a...); break;
    case 604: K<6, 4>::run(a...); break;
    case 608: K<6, 8>::run(a...); break;
    case 616: K<6, 16>::run(a...); break;
    case 704: K<7, 4>::run(a...); break;
    case 708: K<7, 8>::run(a...); break;
    default: K<7, 16>::run(a...); break;
    }
}
template <int NV, int TJ> struct SweepK { static void run(const LuPlan& L, const Par& p, bool f, bool fe, hipStream_t st) { sweep_nv<NV, TJ>(L, p, f, fe, st); } };
template <int NV, int TJ> struct PackK { static void run(const LuPlan& L, const Par& p, hipStream_t st) { pack_nv<NV, TJ>(L, p, st); } };
template <int NV, int TJ> struct GeoK { static void run(const LuPlan& L, const Par& p, hipStream_t st) { pack_geometry_nv<NV, TJ>(L, p, st); } };
template <int NV, int TJ> struct DuBoxK {
    static void run(const LuPlan& L, hipStream_t st) {
        hipLaunchKernelGGL((k_lu_du_to_box<NV, TJ>), dim3((unsigned)L.nwins, 1u, (unsigned)NV), dim3(PK_T), 0, st,
                           L.blks, L.fwd, L.pk, L.win);
    }
};

void launch_lusgs_pipelined(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st) {
    by_shape<SweepK>(p.NV, fenced ? 8 : L.tj, L, p, forward, fenced, st);
}

bool lusgs_tile_ok(int tj) { return tj == 4 || tj == 8 || tj == 16; }

int lusgs_fields(int NV) { return 4 * NV + 21; }
std::vector<int2> lusgs_window_list(const std::vector<Blk>& blks, const std::vector<LuTask>& fwd) {
    std::vector<int2> out;
    for (size_t t = 0; t < fwd.size(); t++) {
        const int nxm = blks[fwd[t].blk].nx - 1;
        for (int i0 = 1; i0 <= nxm; i0 += PK_W) out.push_back(make_int2((int)t, i0));
    }
    return out;
}

// the sentinel into DUF and DUB (also after a sweep gave up waiting)
template <int NV>
static void reset_nv(const LuPlan& L, hipStream_t st) {
    using F = Pk<NV>;
    hipLaunchKernelGGL(k_fill_sentinel<NV>, dim3(1024), dim3(256), 0, st, L.pk, L.P);
}

void launch_lusgs_pack_geometry(const LuPlan& L, const Par& p, hipStream_t st) {
    if (L.nwins > 0) by_shape<GeoK>(p.NV, L.tj, L, p, st);
}
void launch_lusgs_pack(const LuPlan& L, const Par& p, hipStream_t st) {
    if (L.nwins > 0) by_shape<PackK>(p.NV, L.tj, L, p, st);
}
void launch_lusgs_du_to_box(const LuPlan& L, int NV, hipStream_t st) {
    if (L.nwins > 0) by_shape<DuBoxK>(NV, L.tj, L, st);
}

void launch_lusgs_reset(const LuPlan& L, int NV, hipStream_t st) {
    if (NV == 5) reset_nv<5>(L, st);
    else if (NV == 6) reset_nv<6>(L, st);
    else reset_nv<7>(L, st);
}

}  // namespace ocfd
