// Hot-path kernels of one LU-SGS iteration on MI355X (gfx950).  FP64, no contraction (the
// library is built with -ffp-contract=off so results match the reference CPU build bit for bit
// wherever no transcendental is involved).  Each kernel cites the reference routine it replaces.
#include <type_traits>

#include "ocfd_dev.h"
#include "ocfd_launch.h"
#include "ocfd_phys.h"

namespace ocfd {

#define GRID1(n, bs) dim3((unsigned)(((n) + (bs) - 1) / (bs))), dim3(bs)

// ---------------------------------------------------------------------------------------------
// get_viscous's Sutherland viscosity of temperature T (sub_Residual.f90:1355-1383); muf keeps the
// value left after the turbulence model's mu*Re ... mu/Re round trip (sub_turbulence_SA.f90:41,262)
__device__ __forceinline__ void visc_cell(const Blk& b, const Par& p, size_t s, double T) {
    double Tsb = 110.4 / p.T_inf;
    double mu = ((1.0 / p.Re) * (1.0 + Tsb)) * sqrt((T * T) * T) / (Tsb + T);
    b.mu[s] = mu;
    b.muf[s] = (p.turb == 2 || p.turb == 3) ? (mu * p.Re) / p.Re : mu;
}

// comput_duvtpc (sub_Residual.f90:1417-1456): primitives on the whole LAP box; for a viscous case
// also get_viscous on the interior cells, from T in registers
__global__ void __launch_bounds__(64) k_prims(BSet S) {
    OCFD_LOCATE(SP_BOX)
    size_t s = IX(b, i, j, k);
    double p00 = 1.0 / (p.gamma * p.Ma * p.Ma);
    double d = b.U[0][s];
    const double rd = recip_nr(d);
    double u = div_nr(b.U[1][s], d, rd), v = div_nr(b.U[2][s], d, rd), w = div_nr(b.U[3][s], d, rd);
    double T = (b.U[4][s] - (0.5 * d) * ((u * u + v * v) + w * w)) / (p.Cv * d);
    b.prim[0][s] = d; b.prim[1][s] = u; b.prim[2][s] = v; b.prim[3][s] = w;
    b.prim[4][s] = (p00 * d) * T; b.prim[5][s] = T; b.prim[6][s] = sqrt(T) / p.Ma;
    if (p.turb == 3) {
        // SST: k and omega per unit mass, read by the turbulence stencils (sub_turbulence_SST.f90)
        const double kv = b.U[5][s], wv = b.U[6][s];
        b.kt[s] = (kv < 0.0 ? 0.0 : kv) / d;
        b.wt[s] = (wv < 0.0 ? 0.0 : wv) / d;
    }
    if (p.mixed) {
        b.primf[0][s] = (float)d; b.primf[1][s] = (float)u; b.primf[2][s] = (float)v; b.primf[3][s] = (float)w;
        b.primf[4][s] = (float)((p00 * d) * T); b.primf[5][s] = (float)T;
    }
    if (p.visc == 1 && i >= 1 && i <= b.nx - 1 && j >= 1 && j <= b.ny - 1 && k >= 1 && k <= b.nz - 1) visc_cell(b, p, s, T);
}

// mu and muf of the first cells of each face copied into its first ghost layer; one task per block
// and face, tiles of 64 cells along the face's first tangential direction
__global__ void __launch_bounds__(64) k_visc_ghost(SubSet S) {
    OCFD_SUB_LOCATE
    (void)p; (void)B;
    if (a >= T_.n0) return;
    const int f = T_.arg, u = a + 1, v = c + 1;
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    size_t d, s;
    if (f < 2) { d = IX(b, f == 0 ? 0 : nx, u, v); s = IX(b, f == 0 ? 1 : nx - 1, u, v); }
    else if (f < 4) { d = IX(b, u, f == 2 ? 0 : ny, v); s = IX(b, u, f == 2 ? 1 : ny - 1, v); }
    else { d = IX(b, u, v, f == 4 ? 0 : nz); s = IX(b, u, v, f == 4 ? 1 : nz - 1); }
    b.mu[d] = b.mu[s];
    b.muf[d] = b.muf[s];
}

// ---------------------------------------------------------------------------------------------
// Spalart-Allmaras (sub_turbulence_SA.f90:8-273).  U6 is read through limit_vt's clamp
// (sub_time_advance.f90:420-428): negative values act as zero.
__device__ __forceinline__ double vt_of(const Blk& b, size_t s) { double v = b.U[5][s]; return v < 0.0 ? 0.0 : v; }
__device__ __forceinline__ double pw3(double x) { return (x * x) * x; }
__device__ __forceinline__ double pw6(double x) { return ((((x * x) * x) * x) * x) * x; }   // flang expansion order

// Storage of the compute type T of the residual: FP64 (the drop-in) or FP32 (the mixed mode's
// reconstruction, fluxes and turbulence source, whose face fluxes the FP64 assembly reads).
template <typename T> __device__ __forceinline__ const T* prim_of(const Blk& b, int m);
template <> __device__ __forceinline__ const double* prim_of<double>(const Blk& b, int m) { return b.prim[m]; }
template <> __device__ __forceinline__ const float* prim_of<float>(const Blk& b, int m) { return b.primf[m]; }
// the geometry in the compute type: for FP32 the copies converted once per block, which hold the
// values a conversion of the FP64 arrays at every read gives
template <typename T> __device__ __forceinline__ const T* nv_of(const Blk& b, int q);
template <> __device__ __forceinline__ const double* nv_of<double>(const Blk& b, int q) { return b.nv[q]; }
template <> __device__ __forceinline__ const float* nv_of<float>(const Blk& b, int q) { return b.nvf[q]; }
template <typename T> __device__ __forceinline__ const T* S_of(const Blk& b, int d);
template <> __device__ __forceinline__ const double* S_of<double>(const Blk& b, int d) { return b.S[d]; }
template <> __device__ __forceinline__ const float* S_of<float>(const Blk& b, int d) { return b.Sf[d]; }
template <typename T> __device__ __forceinline__ const T* met_of(const Blk& b, int m, int q);
template <> __device__ __forceinline__ const double* met_of<double>(const Blk& b, int m, int q) { return b.met[m][q]; }
template <> __device__ __forceinline__ const float* met_of<float>(const Blk& b, int m, int q) { return b.metf[m][q]; }
__device__ __forceinline__ void put_F(const Blk& b, int d, int m, size_t s, double v) { b.F[d][m][s] = v; }
__device__ __forceinline__ void put_F(const Blk& b, int d, int m, size_t s, float v) { b.Ff[d][m][s] = v; }
__device__ __forceinline__ void put_Fv(const Blk& b, int d, int m, size_t s, double v) { b.Fv[d][m][s] = v; }
__device__ __forceinline__ void put_Fv(const Blk& b, int d, int m, size_t s, float v) { b.Fvf[d][m][s] = v; }
__device__ __forceinline__ void put_Fsa(const Blk& b, int d, int m, size_t s, double v) { b.Fsa[d][m][s] = v; }
__device__ __forceinline__ void put_Fsa(const Blk& b, int d, int m, size_t s, float v) { b.Fsaf[d][m][s] = v; }
template <typename T> __device__ __forceinline__ double get_F(const Blk& b, int d, int m, size_t s) {
    if constexpr (sizeof(T) == 8) return b.F[d][m][s]; else return (double)b.Ff[d][m][s];
}
template <typename T> __device__ __forceinline__ double get_Fv(const Blk& b, int d, int m, size_t s) {
    if constexpr (sizeof(T) == 8) return b.Fv[d][m][s]; else return (double)b.Fvf[d][m][s];
}
template <typename T> __device__ __forceinline__ double get_Fsa(const Blk& b, int d, int m, size_t s) {
    if constexpr (sizeof(T) == 8) return b.Fsa[d][m][s]; else return (double)b.Fsaf[d][m][s];
}

// mu_t over 0..n with both limit_mut passes (sub_Residual.f90:1387-1408) folded in
__global__ void __launch_bounds__(64) k_sa_mut(BSet S) {
    OCFD_LOCATE(SP_N0)
    size_t s = IX(b, i, j, k);
    const double Cv1_3 = pw3(7.1);
    double mu_sa = b.mu[s] * p.Re;
    double d = b.prim[0][s], vt = vt_of(b, s);
    double X = d * vt / mu_sa;
    double fv1 = pw3(X) / (pw3(X) + Cv1_3);
    double m1 = (fv1 * d) * vt;
    if (m1 < 0.0) m1 = 0.0;
    if (p.MUT_MAX >= 0.0 && m1 > p.MUT_MAX * mu_sa) m1 = p.MUT_MAX * mu_sa;
    double m2 = m1 / p.Re;
    double muf = (b.mu[s] * p.Re) / p.Re;
    if (m2 < 0.0) m2 = 0.0;
    if (p.MUT_MAX >= 0.0 && m2 > p.MUT_MAX * muf) m2 = p.MUT_MAX * muf;
    b.mut[s] = m2;
}

// SA constants (sub_turbulence_SA.f90:18-33)
struct SAc {
    static constexpr double sigma = 2.0 / 3.0, Cb1 = 0.1355, Cb2 = 0.622, k = 0.41;
    static constexpr double Cw2 = 0.3, Cw3 = 2.0, Ct3 = 1.2, Ct4 = 0.5;
};

// nu-tilde face flux of the SA equation on face (fi,fj,fk) of direction DIR (I-1/2 convention):
// fv = (-upwind convection + (1+cb2)/sigma (vt+mu/rho)/Re grad(vt).n) S and f2 = grad(vt).n S
// (sub_turbulence_SA.f90:60-174)
template <int DIR, typename T>
__device__ __forceinline__ void sa_face(const Blk& b, const Par& p, int fi, int fj, int fk, T& fv, T& f2) {
    constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
    const T c_v0 = T((0.5 * (1.0 + SAc::Cb2)) / SAc::sigma);
    const T Re = T(p.Re);
    const T* d = prim_of<T>(b, 0);
    const T *uu = prim_of<T>(b, 1), *vv = prim_of<T>(b, 2), *ww = prim_of<T>(b, 3);
    size_t f = IX(b, fi, fj, fk), fm = IX(b, fi - di, fj - dj, fk - dk);
    T s1x = nv_of<T>(b, 3 * DIR)[f], s1y = nv_of<T>(b, 3 * DIR + 1)[f], s1z = nv_of<T>(b, 3 * DIR + 2)[f];
    T gi, gj, gk;
    auto VT = [&](int a, int bb, int cc) { return T(vt_of(b, IX(b, a, bb, cc))); };
    if (DIR == 0) {
        gi = VT(fi, fj, fk) - VT(fi - 1, fj, fk);
        gj = T(0.25) * (((VT(fi, fj + 1, fk) - VT(fi, fj - 1, fk)) + VT(fi - 1, fj + 1, fk)) - VT(fi - 1, fj - 1, fk));
        gk = T(0.25) * (((VT(fi, fj, fk + 1) - VT(fi, fj, fk - 1)) + VT(fi - 1, fj, fk + 1)) - VT(fi - 1, fj, fk - 1));
    } else if (DIR == 1) {
        gi = T(0.25) * (((VT(fi + 1, fj, fk) - VT(fi - 1, fj, fk)) + VT(fi + 1, fj - 1, fk)) - VT(fi - 1, fj - 1, fk));
        gj = VT(fi, fj, fk) - VT(fi, fj - 1, fk);
        gk = T(0.25) * (((VT(fi, fj, fk + 1) - VT(fi, fj, fk - 1)) + VT(fi, fj - 1, fk + 1)) - VT(fi, fj - 1, fk - 1));
    } else {
        gi = T(0.25) * (((VT(fi + 1, fj, fk) - VT(fi - 1, fj, fk)) + VT(fi + 1, fj, fk - 1)) - VT(fi - 1, fj, fk - 1));
        gj = T(0.25) * (((VT(fi, fj + 1, fk) - VT(fi, fj - 1, fk)) + VT(fi, fj + 1, fk - 1)) - VT(fi, fj - 1, fk - 1));
        gk = VT(fi, fj, fk) - VT(fi, fj, fk - 1);
    }
    auto M = [&](int q) { return met_of<T>(b, DIR + 1, q)[f]; };
    T vtx = (gi * M(0) + gj * M(3)) + gk * M(6);
    T vty = (gi * M(1) + gj * M(4)) + gk * M(7);
    T vtz = (gi * M(2) + gj * M(5)) + gk * M(8);
    T v0 = c_v0 * (((T(vt_of(b, f)) + (T(b.mu[f]) * Re) / d[f]) + T(vt_of(b, fm))) + (T(b.mu[fm]) * Re) / d[fm]);
    T vn1 = (uu[fm] * s1x + vv[fm] * s1y) + ww[fm] * s1z;
    T vn2 = (uu[f] * s1x + vv[f] * s1y) + ww[f] * s1z;
    T vfi = T(0.5) * ((vn1 + fabs(vn1)) * T(vt_of(b, fm)) + (vn2 - fabs(vn2)) * T(vt_of(b, f)));
    T gn = (vtx * s1x + vty * s1y) + vtz * s1z;
    T Sf = S_of<T>(b, DIR)[f];
    fv = (-vfi + (v0 / Re) * gn) * Sf;
    f2 = gn * Sf;
}

// Res(6) of cell s from its six stored face fluxes (i, j, k in that order) and the source in the
// CFL3D form (sub_turbulence_SA.f90:103,138,174,181-249)
template <typename T>
__device__ __forceinline__ double sa_residual(const Blk& b, const Par& p, int i, int j, int k, size_t s) {
    const double Cw1 = SAc::Cb1 / (SAc::k * SAc::k) + (1.0 + SAc::Cb2) / SAc::sigma;
    const double Cv1_3 = pw3(7.1), Cw3_6 = pw6(SAc::Cw3);
    const double Re = p.Re;
    const double* d = b.prim[0];
    const T* dt_ = prim_of<T>(b, 0);
    const T *ut = prim_of<T>(b, 1), *vtt = prim_of<T>(b, 2), *wt = prim_of<T>(b, 3);
    double vt_c = vt_of(b, s), mu_c = b.mu[s] * Re;
    double v0c = ((vt_c + mu_c / d[s]) * SAc::Cb2) / SAc::sigma;
    const size_t sp[3] = {s + 1, s + (size_t)b.NI, s + (size_t)b.NI * b.NJ};
    double res = 0.0;
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        double fv0 = get_Fsa<T>(b, dir, 0, s), fv1 = get_Fsa<T>(b, dir, 0, sp[dir]);
        double f20 = get_Fsa<T>(b, dir, 1, s), f21 = get_Fsa<T>(b, dir, 1, sp[dir]);
        if (dir == 0) res = (fv1 - fv0) - (v0c / Re) * (f21 - f20);
        else res = ((res + fv1) - fv0) - (v0c / Re) * (f21 - f20);
    }
    auto UU = [&](int a, int bb, int cc) { return ut[IX(b, a, bb, cc)]; };
    auto VV = [&](int a, int bb, int cc) { return vtt[IX(b, a, bb, cc)]; };
    auto WW = [&](int a, int bb, int cc) { return wt[IX(b, a, bb, cc)]; };
    T ui = UU(i + 1, j, k) - UU(i - 1, j, k), vi = VV(i + 1, j, k) - VV(i - 1, j, k), wi = WW(i + 1, j, k) - WW(i - 1, j, k);
    T uj = UU(i, j + 1, k) - UU(i, j - 1, k), vj = VV(i, j + 1, k) - VV(i, j - 1, k), wj = WW(i, j + 1, k) - WW(i, j - 1, k);
    T uk = UU(i, j, k + 1) - UU(i, j, k - 1), vk = VV(i, j, k + 1) - VV(i, j, k - 1), wk = WW(i, j, k + 1) - WW(i, j, k - 1);
    auto M0 = [&](int q) { return met_of<T>(b, 0, q)[s]; };
    T ix = M0(0), iy = M0(1), iz = M0(2), jx = M0(3), jy = M0(4), jz = M0(5), kx = M0(6), ky = M0(7), kz = M0(8);
    T vx = (vi * ix + vj * jx) + vk * kx, wx = (wi * ix + wj * jx) + wk * kx;
    T uy = (ui * iy + uj * jy) + uk * ky, wy = (wi * iy + wj * jy) + wk * ky;
    T uz = (ui * iz + uj * jz) + uk * kz, vz = (vi * iz + vj * jz) + vk * kz;
    T S = sqrt((sq(wy - vz) + sq(uz - wx)) + sq(vx - uy));
    T dw = T(b.dw[s]);
    T X = dt_[s] * T(vt_c) / T(mu_c);
    T fv1 = (X * X * X) / ((X * X * X) + T(Cv1_3));
    T fv2 = T(1) - X / (T(1) + X * fv1);
    T ft2 = T(SAc::Ct3) * exp(-((T(SAc::Ct4) * X) * X));
    T S1 = S + (fv2 * T(vt_c)) / (T(Re) * sq(T(SAc::k) * dw));
    S1 = S1 > T(0) ? S1 : T(0);
    T r = T(vt_c) / (((((T(Re) * S1) * T(SAc::k)) * T(SAc::k)) * dw) * dw);
    if (r > T(10)) r = T(10);
    T g = r + T(SAc::Cw2) * ((((((r * r) * r) * r) * r) * r) - r);
    T fw = g * pow((T(1) + T(Cw3_6)) / ((((((g * g) * g) * g) * g) * g) + T(Cw3_6)), T(1.0 / 6.0));
    T Q3 = (T(SAc::Cb1) * ((T(1) - ft2) * fv2 + ft2)) / T(SAc::k * SAc::k) - T(Cw1) * fw;
    T Q_SA = ((T(SAc::Cb1) * (T(1) - ft2)) * S1) * T(vt_c) + (Q3 * sq(T(vt_c) / dw)) / T(Re);
    return res + (double)(Q_SA * T(b.vol[s]));
}

// Amut_boundary (sub_turbulence_SA.f90:282-368) on one wall subface
__device__ __forceinline__ void d_amut(const Blk& b, const Sub& B, int a, int c) {
    switch (B.face) {
    case 1: case 4: { int j = B.jb + a, k = B.kb + c; if (j > B.je - 1 || k > B.ke - 1) return;
        int ig = B.face == 1 ? 0 : B.ie, ii = B.face == 1 ? 1 : B.ie - 1; b.mut[IX(b, ig, j, k)] = -b.mut[IX(b, ii, j, k)]; break; }
    case 2: case 5: { int i = B.ib + a, k = B.kb + c; if (i > B.ie - 1 || k > B.ke - 1) return;
        int jg = B.face == 2 ? 0 : B.je, jj = B.face == 2 ? 1 : B.je - 1; b.mut[IX(b, i, jg, k)] = -b.mut[IX(b, i, jj, k)]; break; }
    default: { int i = B.ib + a, j = B.jb + c; if (i > B.ie - 1 || j > B.je - 1) return;
        int kg = B.face == 3 ? 0 : B.ke, kk = B.face == 3 ? 1 : B.ke - 1; b.mut[IX(b, i, j, kg)] = -b.mut[IX(b, i, j, kk)]; break; }
    }
}

// ---------------------------------------------------------------------------------------------
// Menter SST (sub_turbulence_SST.f90:7-348).  f1 on the last face of each direction reads the
// never-written cell (nx,.,.)/(.,ny,.)/(.,.,nz); here that storage holds zero.
// k and omega per unit mass, max(U, 0) / rho, as k_prims stores them for the iteration
__device__ __forceinline__ double kt_of(const Blk& b, size_t s) { return b.kt[s]; }
__device__ __forceinline__ double wt_of(const Blk& b, size_t s) { return b.wt[s]; }

// cell pass: mu_t, f1, sources into Res(6:7), in the compute type T (FP64, or FP32 in the mixed
// mode, where the velocities come from the FP32 primitives); stored in FP64
template <typename T>
__global__ void __launch_bounds__(64) k_sst_cell(BSet S) {
    OCFD_LOCATE(SP_CELL)
    const T sigma_w2 = T(0.856), beta1 = T(0.075), Cw1s = T(0.533), beta2 = T(0.0828), Cw2s = T(0.440);
    const T a1 = T(0.31), betas = T(0.09), c009 = T((double)0.09f);
    const T Re = T(p.Re);
    const T *uu = prim_of<T>(b, 1), *vv = prim_of<T>(b, 2), *ww = prim_of<T>(b, 3);
    size_t s = IX(b, i, j, k);
    auto DIFF = [&](const T* a, int di, int dj, int dk) { return a[IX(b, i + di, j + dj, k + dk)] - a[IX(b, i - di, j - dj, k - dk)]; };
    auto DK = [&](int di, int dj, int dk) { return T(kt_of(b, IX(b, i + di, j + dj, k + dk))) - T(kt_of(b, IX(b, i - di, j - dj, k - dk))); };
    auto DW = [&](int di, int dj, int dk) { return T(wt_of(b, IX(b, i + di, j + dj, k + dk))) - T(wt_of(b, IX(b, i - di, j - dj, k - dk))); };
    T ui = DIFF(uu, 1, 0, 0), vi = DIFF(vv, 1, 0, 0), wi = DIFF(ww, 1, 0, 0), kti = DK(1, 0, 0), wti = DW(1, 0, 0);
    T uj = DIFF(uu, 0, 1, 0), vj = DIFF(vv, 0, 1, 0), wj = DIFF(ww, 0, 1, 0), ktj = DK(0, 1, 0), wtj = DW(0, 1, 0);
    T uk = DIFF(uu, 0, 0, 1), vk = DIFF(vv, 0, 0, 1), wk = DIFF(ww, 0, 0, 1), ktk = DK(0, 0, 1), wtk = DW(0, 0, 1);
    auto M0 = [&](int q) { return met_of<T>(b, 0, q)[s]; };
    T ix = M0(0), iy = M0(1), iz = M0(2), jx = M0(3), jy = M0(4), jz = M0(5), kx = M0(6), ky = M0(7), kz = M0(8);
    T vx = (vi * ix + vj * jx) + vk * kx, wx = (wi * ix + wj * jx) + wk * kx;
    T ktx = (kti * ix + ktj * jx) + ktk * kx, wtx = (wti * ix + wtj * jx) + wtk * kx;
    T uy = (ui * iy + uj * jy) + uk * ky, wy = (wi * iy + wj * jy) + wk * ky;
    T kty = (kti * iy + ktj * jy) + ktk * ky, wty = (wti * iy + wtj * jy) + wtk * ky;
    T uz = (ui * iz + uj * jz) + uk * kz, vz = (vi * iz + vj * jz) + vk * kz;
    T ktz = (kti * iz + ktj * jz) + ktk * kz, wtz = (wti * iz + wtj * jz) + wtk * kz;
    T omega = sqrt((sq(wy - vz) + sq(uz - wx)) + sq(vx - uy));
    T d = T(b.prim[0][s]), mu = T(b.mu[s] * p.Re), dw = T(b.dw[s]), kt = T(kt_of(b, s)), wt = T(wt_of(b, s));
    T a2a = T(2) * sqrt(fabs(kt)) / (((c009 * wt) * dw) * Re);
    T a2b = (T(500) * mu) / ((((d * wt) * sq(dw)) * Re) * Re);
    T arg2 = a2a > a2b ? a2a : a2b;
    T f2 = tanh(sq(arg2));
    T den = a1 * wt, alt = (f2 * fabs(omega)) / Re;
    T mutc = ((a1 * d) * kt) / (den > alt ? den : alt);
    T Kws = (((T(2) * ((ktx * wtx + kty * wty) + ktz * wtz)) * d) * sigma_w2) / (wt + T(1.0e-20));
    T CD_kw = Kws > T(1.0e-20) ? Kws : T(1.0e-20);
    T a3a = sqrt(fabs(kt)) / (((c009 * wt) * dw) * Re);
    T arg3 = a3a > a2b ? a3a : a2b;
    T a1b = (((T(4) * d) * sigma_w2) * kt) / (CD_kw * sq(dw));
    T arg1 = arg3 < a1b ? arg3 : a1b;
    T f1 = tanh(((arg1 * arg1) * arg1) * arg1);
    T Pk = (mutc * omega) * omega;
    T Qk = Pk / Re - (((Re * betas) * d) * wt) * kt;
    T Cw = f1 * Cw1s + (T(1) - f1) * Cw2s;
    T beta = f1 * beta1 + (T(1) - f1) * beta2;
    T Qw = ((((Cw * d) * omega) * omega) / Re - ((Re * beta) * d) * sq(wt)) + ((T(1) - f1) * Kws) / Re;
    T vol = T(b.vol[s]);
    b.Res[5][s] = (double)(Qk * vol);
    b.Res[6][s] = (double)(Qw * vol);
    b.mut[s] = (double)mutc;      // Re-scaled value, divided by Re in k_sst_finish
    b.aux[s] = (double)f1;
}

// mu_t ghost copies over the whole -1:n+1 extent (sub_turbulence_SST.f90:170-175): face f in
// order i-, i+, j-, j+, k-, k+ launched one after the other
__global__ void k_sst_ghost(BSet S, int f) {
    const Blk& b = S.blk[blockIdx.z];
    int a = blockIdx.x * blockDim.x + threadIdx.x - 1, c = blockIdx.y - 1;
    int nx = b.nx, ny = b.ny, nz = b.nz;
    if (f < 2) { if (a > ny + 1 || c > nz + 1) return; b.mut[IX(b, f == 0 ? 0 : nx, a, c)] = b.mut[IX(b, f == 0 ? 1 : nx - 1, a, c)]; }
    else if (f < 4) { if (a > nx + 1 || c > nz + 1) return; b.mut[IX(b, a, f == 2 ? 0 : ny, c)] = b.mut[IX(b, a, f == 2 ? 1 : ny - 1, c)]; }
    else { if (a > nx + 1 || c > ny + 1) return; b.mut[IX(b, a, c, f == 4 ? 0 : nz)] = b.mut[IX(b, a, c, f == 4 ? 1 : nz - 1)]; }
}

// k and omega flux through face (i,j,k) of direction DIR, between cells (i,j,k) - e_DIR and
// (i,j,k): convection upwinded on the normal velocity, diffusion with the face gradient of k and
// omega; one thread per face, each face evaluated once for both of its cells; compute type T,
// stored in FP64
template <int DIR, typename T>
__global__ void __launch_bounds__(64) k_sst_face(BSet S) {
    OCFD_LOCATE(SP_FI + DIR)
    const T sigma_k1 = T(0.85), sigma_w1 = T(0.5), sigma_k2 = T(1.0), sigma_w2 = T(0.856);
    const T Re = T(p.Re);
    const T *uu = prim_of<T>(b, 1), *vv = prim_of<T>(b, 2), *ww = prim_of<T>(b, 3);
    const int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
    const int fi = i, fj = j, fk_ = k;
    size_t f = IX(b, fi, fj, fk_), fm = IX(b, fi - di, fj - dj, fk_ - dk);
    T f1 = T(b.aux[f]);
    T skk = f1 * sigma_k1 + (T(1) - f1) * sigma_k2, sww = f1 * sigma_w1 + (T(1) - f1) * sigma_w2;
    T mua = T(b.mu[fm] * p.Re), mub = T(b.mu[f] * p.Re), mta = T(b.mut[fm]), mtb = T(b.mut[f]);
    T muk = (((mua + mub) + skk * (mta + mtb)) * T(0.5)) / Re;
    T muw = (((mua + mub) + sww * (mta + mtb)) * T(0.5)) / Re;
    auto K = [&](int a, int bb, int cc) { return T(kt_of(b, IX(b, a, bb, cc))); };
    auto W = [&](int a, int bb, int cc) { return T(wt_of(b, IX(b, a, bb, cc))); };
    const T q = T(0.25);
    T kti, ktj, ktk, wti, wtj, wtk;
    if constexpr (DIR == 0) {
        kti = K(fi, fj, fk_) - K(fi - 1, fj, fk_); wti = W(fi, fj, fk_) - W(fi - 1, fj, fk_);
        ktj = q * (((K(fi, fj + 1, fk_) - K(fi, fj - 1, fk_)) + K(fi - 1, fj + 1, fk_)) - K(fi - 1, fj - 1, fk_));
        wtj = q * (((W(fi, fj + 1, fk_) - W(fi, fj - 1, fk_)) + W(fi - 1, fj + 1, fk_)) - W(fi - 1, fj - 1, fk_));
        ktk = q * (((K(fi, fj, fk_ + 1) - K(fi, fj, fk_ - 1)) + K(fi - 1, fj, fk_ + 1)) - K(fi - 1, fj, fk_ - 1));
        wtk = q * (((W(fi, fj, fk_ + 1) - W(fi, fj, fk_ - 1)) + W(fi - 1, fj, fk_ + 1)) - W(fi - 1, fj, fk_ - 1));
    } else if constexpr (DIR == 1) {
        kti = q * (((K(fi + 1, fj, fk_) - K(fi - 1, fj, fk_)) + K(fi + 1, fj - 1, fk_)) - K(fi - 1, fj - 1, fk_));
        ktj = K(fi, fj, fk_) - K(fi, fj - 1, fk_);
        ktk = q * (((K(fi, fj, fk_ + 1) - K(fi, fj, fk_ - 1)) + K(fi, fj - 1, fk_ + 1)) - K(fi, fj - 1, fk_ - 1));
        wti = q * (((W(fi + 1, fj, fk_) - W(fi - 1, fj, fk_)) + W(fi + 1, fj - 1, fk_)) - W(fi - 1, fj - 1, fk_));
        wtj = W(fi, fj, fk_) - W(fi, fj - 1, fk_);
        wtk = q * (((W(fi, fj, fk_ + 1) - W(fi, fj, fk_ - 1)) + W(fi, fj - 1, fk_ + 1)) - W(fi, fj - 1, fk_ - 1));
    } else {
        kti = q * (((K(fi + 1, fj, fk_) - K(fi - 1, fj, fk_)) + K(fi + 1, fj, fk_ - 1)) - K(fi - 1, fj, fk_ - 1));
        ktj = q * (((K(fi, fj + 1, fk_) - K(fi, fj - 1, fk_)) + K(fi, fj + 1, fk_ - 1)) - K(fi, fj - 1, fk_ - 1));
        ktk = K(fi, fj, fk_) - K(fi, fj, fk_ - 1);
        wti = q * (((W(fi + 1, fj, fk_) - W(fi - 1, fj, fk_)) + W(fi + 1, fj, fk_ - 1)) - W(fi - 1, fj, fk_ - 1));
        wtj = q * (((W(fi, fj + 1, fk_) - W(fi, fj - 1, fk_)) + W(fi, fj + 1, fk_ - 1)) - W(fi, fj - 1, fk_ - 1));
        wtk = W(fi, fj, fk_) - W(fi, fj, fk_ - 1);
    }
    auto M = [&](int c) { return met_of<T>(b, DIR + 1, c)[f]; };
    T ktx = (kti * M(0) + ktj * M(3)) + ktk * M(6), wtx = (wti * M(0) + wtj * M(3)) + wtk * M(6);
    T kty = (kti * M(1) + ktj * M(4)) + ktk * M(7), wty = (wti * M(1) + wtj * M(4)) + wtk * M(7);
    T ktz = (kti * M(2) + ktj * M(5)) + ktk * M(8), wtz = (wti * M(2) + wtj * M(5)) + wtk * M(8);
    T s1x = nv_of<T>(b, 3 * DIR)[f], s1y = nv_of<T>(b, 3 * DIR + 1)[f], s1z = nv_of<T>(b, 3 * DIR + 2)[f];
    T vn1 = (uu[fm] * s1x + vv[fm] * s1y) + ww[fm] * s1z;
    T vn2 = (uu[f] * s1x + vv[f] * s1y) + ww[f] * s1z;
    T kfi = T(0.5) * ((vn1 + fabs(vn1)) * K(fi - di, fj - dj, fk_ - dk) + (vn2 - fabs(vn2)) * K(fi, fj, fk_));
    T wfi = T(0.5) * ((vn1 + fabs(vn1)) * W(fi - di, fj - dj, fk_ - dk) + (vn2 - fabs(vn2)) * W(fi, fj, fk_));
    T Sf = S_of<T>(b, DIR)[f];
    b.Fsa[DIR][0][f] = (double)((-kfi + muk * ((ktx * s1x + kty * s1y) + ktz * s1z)) * Sf);
    b.Fsa[DIR][1][f] = (double)((-wfi + muw * ((wtx * s1x + wty * s1y) + wtz * s1z)) * Sf);
}

// the face fluxes accumulated into Res(6:7) per direction in i, j, k order
__global__ void __launch_bounds__(64) k_sst_res(BSet S) {
    OCFD_LOCATE(SP_CELL)
    size_t s = IX(b, i, j, k);
    const size_t sp[3] = {s + 1, s + (size_t)b.NI, s + (size_t)b.NI * b.NJ};
    double r6 = b.Res[5][s], r7 = b.Res[6][s];
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        r6 = (r6 + b.Fsa[dir][0][sp[dir]]) - b.Fsa[dir][0][s];
        r7 = (r7 + b.Fsa[dir][1][sp[dir]]) - b.Fsa[dir][1][s];
    }
    b.Res[5][s] = r6;
    b.Res[6][s] = r7;
}

// mu_t/Re and the limit_mut pass of Residual (sub_turbulence_SST.f90:333-340; sub_Residual.f90:1387-1408)
__global__ void __launch_bounds__(64) k_sst_finish(BSet S) {
    OCFD_LOCATE(SP_N0)
    size_t s = IX(b, i, j, k);
    double m2 = b.mut[s] / p.Re;
    double muf = (b.mu[s] * p.Re) / p.Re;
    if (m2 < 0.0) m2 = 0.0;
    if (p.MUT_MAX >= 0.0 && m2 > p.MUT_MAX * muf) m2 = p.MUT_MAX * muf;
    b.mut[s] = m2;
}

// ---------------------------------------------------------------------------------------------
// the 1-D flux of the face frame: FLUX fixes it at compile time (the register budget of a face
// kernel is then that of one solver); FLUX = 0 selects it at run time from iflux
template <int FLUX, typename T>
__device__ __forceinline__ void riemann_t(const T* QL, const T* QR, T* F, T gamma, int iflux) {
    if constexpr (FLUX == 0) riemann(QL, QR, F, gamma, iflux);
    else if constexpr (FLUX == F_SW) flux_sw(QL, QR, F, gamma);
    else if constexpr (FLUX == F_ROE) flux_roe(QL, QR, F, gamma);
    else if constexpr (FLUX == F_VL) flux_vl(QL, QR, F, gamma);
    else if constexpr (FLUX == F_AUSM) flux_ausmpw(QL, QR, F, gamma);
    else flux_hll_hllc(QL, QR, F, gamma, FLUX);
}

// flux_inviscous_i/j/k (sub_Residual.f90:182-705) of face (i,j,k) of direction DIR, stored
template <int DIR, typename T, int RECON, int FLUX>
__device__ __forceinline__ void inv_face(const Blk& b, const Par& p, int i, int j, int k) {
    size_t s = IX(b, i, j, k);
    int pos = DIR == 0 ? i : DIR == 1 ? j : k;
    int nd = DIR == 0 ? b.nx : DIR == 1 ? b.ny : b.nz;
    int scheme = p.scheme;
    uint8_t fl = DIR == 0 ? b.bcf[0][JK(b, j, k)] : DIR == 1 ? b.bcf[1][IK(b, i, k)] : b.bcf[2][IJ(b, i, j)];
    if (fl & 1) { if (pos == 1) scheme = S_CD2; else if (pos < LAP) scheme = p.bscheme; }
    if (fl & 2) { if (pos == nd) scheme = S_CD2; else if (pos > nd - LAP + 1) scheme = p.bscheme; }
    if (b.flags[0] == 1) scheme = S_UD1;
    T s1x = nv_of<T>(b, 3 * DIR)[s], s1y = nv_of<T>(b, 3 * DIR + 1)[s], s1z = nv_of<T>(b, 3 * DIR + 2)[s];
    T t1x, t1y, t1z;
    if constexpr (std::is_same<T, float>::value) { t1x = b.t1f[DIR][0][s]; t1y = b.t1f[DIR][1][s]; t1z = b.t1f[DIR][2][s]; }
    else if (DIR == 0) { size_t a = IX(b, i, j + 1, k), c = IX(b, i, j, k + 1); t1x = T(b.x[a] - b.x[c]); t1y = T(b.y[a] - b.y[c]); t1z = T(b.z[a] - b.z[c]); }
    else if (DIR == 1) { size_t a = IX(b, i + 1, j, k + 1); t1x = T(b.x[a] - b.x[s]); t1y = T(b.y[a] - b.y[s]); t1z = T(b.z[a] - b.z[s]); }
    else { size_t a = IX(b, i + 1, j + 1, k); t1x = T(b.x[a] - b.x[s]); t1y = T(b.y[a] - b.y[s]); t1z = T(b.z[a] - b.z[s]); }
    T t10 = T(1) / sqrt((t1x * t1x + t1y * t1y) + t1z * t1z);
    T s2x = t1x * t10, s2y = t1y * t10, s2z = t1z * t10;
    T s3x = s1y * s2z - s1z * s2y, s3y = s1z * s2x - s1x * s2z, s3z = s1x * s2y - s1y * s2x;
    const long st = DIR == 0 ? 1 : DIR == 1 ? (long)b.NI : (long)b.NI * b.NJ;
    T UL[5], UR[5];
    if constexpr (RECON == 0) {
        for (int m = 0; m < 5; m++) {
            const T* a = prim_of<T>(b, m) + s;
            T u[8];
#pragma unroll
            for (int q = 0; q < 8; q++) u[q] = a[(q - 4) * st];     // u[q] = cell (face - 4 + q): u(q-3)
            UL[m] = scheme_fP(u + 3, scheme);
            UR[m] = scheme_fm(u + 3, scheme);
        }
    } else {
        T U0[5][8];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            const double* a = b.U[m] + s;
#pragma unroll
            for (int q = 0; q < 8; q++) U0[m][q] = T(a[(q - 4) * st]);
        }
        T QL[5], QR[5];
        if constexpr (RECON == 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) { QL[m] = scheme_fP(U0[m] + 3, scheme); QR[m] = scheme_fm(U0[m] + 3, scheme); }
        } else {
            // Reconstuction_Characteristic (sub_Residual.f90:1291-1347)
            T gamma = T(p.gamma), Uh[5];
            for (int m = 0; m < 5; m++) Uh[m] = T(0.5) * (U0[m][3] + U0[m][4]);
            T d1 = Uh[0], u1 = Uh[1] / d1, v1 = Uh[2] / d1, w1 = Uh[3] / d1;
            T p1 = (Uh[4] - ((Uh[1] * u1 + Uh[2] * v1) + Uh[3] * w1) * T(0.5)) * (gamma - T(1));
            T c1 = sqrt(gamma * p1 / d1);
            T V2 = ((u1 * u1 + v1 * v1) + w1 * w1) * T(0.5);
            T tmp1 = (gamma - T(1)) / c1, tmp3 = (gamma - T(1)) / (c1 * c1), tmp5 = T(1) / (T(2) * c1), tmp0 = T(1) / tmp3;
            const T o = T(1), z = T(0), h = T(0.5);
            T S[5][5] = {{V2 - tmp0, -u1, -v1, -w1, o}, {-v1, z, o, z, z}, {-w1, z, z, o, z},
                         {-u1 - V2 * tmp1, o + tmp1 * u1, tmp1 * v1, tmp1 * w1, -tmp1},
                         {-u1 + V2 * tmp1, o - tmp1 * u1, -tmp1 * v1, -tmp1 * w1, tmp1}};
            T S1[5][5] = {{-tmp3, z, z, -tmp5, tmp5}, {-tmp3 * u1, z, z, h - u1 * tmp5, h + u1 * tmp5},
                          {-tmp3 * v1, o, z, -v1 * tmp5, v1 * tmp5}, {-tmp3 * w1, z, o, -w1 * tmp5, w1 * tmp5},
                          {-tmp3 * V2, v1, w1, ((c1 * u1 - V2) - tmp0) * tmp5, ((c1 * u1 + V2) + tmp0) * tmp5}};
            T VL[5], VR[5];
#pragma unroll
            for (int m = 0; m < 5; m++) {
                T V0[8];
                for (int q = 0; q < 8; q++) { T a = T(0); for (int jj = 0; jj < 5; jj++) a = a + S[m][jj] * U0[jj][q]; V0[q] = a; }
                VL[m] = scheme_fP(V0 + 3, scheme); VR[m] = scheme_fm(V0 + 3, scheme);
            }
            for (int m = 0; m < 5; m++) {
                T a = T(0), c = T(0);
                for (int jj = 0; jj < 5; jj++) { a = a + S1[m][jj] * VL[jj]; c = c + S1[m][jj] * VR[jj]; }
                QL[m] = a; QR[m] = c;
            }
        }
        T g = T(p.gamma);
        UL[0] = QL[0]; UL[1] = QL[1] / UL[0]; UL[2] = QL[2] / UL[0]; UL[3] = QL[3] / UL[0];
        UL[4] = (QL[4] - ((UL[1] * QL[1] + UL[2] * QL[2]) + UL[3] * QL[3]) * T(0.5)) * (g - T(1));
        UR[0] = QR[0]; UR[1] = QR[1] / UR[0]; UR[2] = QR[2] / UR[0]; UR[3] = QR[3] / UR[0];
        UR[4] = (QR[4] - ((UR[1] * QR[1] + UR[2] * QR[2]) + UR[3] * QR[3]) * T(0.5)) * (g - T(1));
    }
    if (p.positivity == 1) {
        if (UL[0] <= T(1.0e-20) || UL[4] <= T(1.0e-20)) { size_t t = s - st; for (int m = 0; m < 5; m++) UL[m] = prim_of<T>(b, m)[t]; }
        if (UR[0] <= T(1.0e-20) || UR[4] <= T(1.0e-20)) { for (int m = 0; m < 5; m++) UR[m] = prim_of<T>(b, m)[s]; }
    }
    T QL[5], QR[5], F0[5];
    QL[0] = UL[0]; QL[1] = (UL[1] * s1x + UL[2] * s1y) + UL[3] * s1z; QL[2] = (UL[1] * s2x + UL[2] * s2y) + UL[3] * s2z;
    QL[3] = (UL[1] * s3x + UL[2] * s3y) + UL[3] * s3z; QL[4] = UL[4];
    QR[0] = UR[0]; QR[1] = (UR[1] * s1x + UR[2] * s1y) + UR[3] * s1z; QR[2] = (UR[1] * s2x + UR[2] * s2y) + UR[3] * s2z;
    QR[3] = (UR[1] * s3x + UR[2] * s3y) + UR[3] * s3z; QR[4] = UR[4];
    riemann_t<FLUX, T>(QL, QR, F0, T(p.gamma), p.flux);
    T Sf = S_of<T>(b, DIR)[s];
    put_F(b, DIR, 0, s, -F0[0] * Sf);
    put_F(b, DIR, 1, s, -((F0[1] * s1x + F0[2] * s2x) + F0[3] * s3x) * Sf);
    put_F(b, DIR, 2, s, -((F0[1] * s1y + F0[2] * s2y) + F0[3] * s3y) * Sf);
    put_F(b, DIR, 3, s, -((F0[1] * s1z + F0[2] * s2z) + F0[3] * s3z) * Sf);
    put_F(b, DIR, 4, s, -F0[4] * Sf);
}

// explicit average-state flux on a physical k subface (sub_Residual.f90:667-688)
template <typename T>
__global__ void __launch_bounds__(64) k_inv_kbc(SubSet S) {
    OCFD_SUB_LOCATE
    int i = B.ib + a, j = B.jb + c, k = B.kb;
    if (i > B.ie - 1 || j > B.je - 1) return;
    size_t s = IX(b, i, j, k), t = IX(b, i, j, k - 1);
    double g = p.gamma;
    double d0 = (b.prim[0][s] + b.prim[0][t]) * 0.5, uu0 = (b.prim[1][s] + b.prim[1][t]) * 0.5, v0 = (b.prim[2][s] + b.prim[2][t]) * 0.5;
    double w0 = (b.prim[3][s] + b.prim[3][t]) * 0.5, p0 = (b.prim[4][s] + b.prim[4][t]) * 0.5;
    double E0 = p0 / (g - 1.0) + (d0 * ((uu0 * uu0 + v0 * v0) + w0 * w0)) * 0.5;
    double n1 = b.nv[6][s], n2 = b.nv[7][s], n3 = b.nv[8][s];
    double un = (uu0 * n1 + v0 * n2) + w0 * n3;
    double Sk = b.S[2][s];
    put_F(b, 2, 0, s, T(-(d0 * un) * Sk));
    put_F(b, 2, 1, s, T(-((d0 * un) * uu0 + p0 * n1) * Sk));
    put_F(b, 2, 2, s, T(-((d0 * un) * v0 + p0 * n2) * Sk));
    put_F(b, 2, 3, s, T(-((d0 * un) * w0 + p0 * n3) * Sk));
    put_F(b, 2, 4, s, T(-((E0 + p0) * un) * Sk));
}

// flux_viscous_i/j/k (sub_Residual.f90:712-1247) of face (i,j,k) of direction DIR, stored
template <int DIR, typename T>
__device__ __forceinline__ void visc_face(const Blk& b, const Par& p, int i, int j, int k) {
    constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
    const T tmp1 = T(4.0 / 3.0), tmp2 = T(2.0 / 3.0);
    size_t s = IX(b, i, j, k);
    if (b.flags[0] == 1) {      // over-limit block: no viscous flux, zero wall forces (:734-750)
        int pos0 = DIR == 0 ? i : DIR == 1 ? j : k, nd0 = DIR == 0 ? b.nx : DIR == 1 ? b.ny : b.nz;
        if (pos0 == 1 || pos0 == nd0) {
            int f = DIR + (pos0 == 1 ? 0 : 3);
            size_t m = DIR == 0 ? JK(b, j, k) : DIR == 1 ? IK(b, i, k) : IJ(b, i, j);
            b.surf[f][0][m] = 0.0; b.surf[f][1][m] = 0.0; b.surf[f][2][m] = 0.0;
        }
        return;
    }
    int im = i - di, jm = j - dj, km = k - dk;
    size_t sm = IX(b, im, jm, km);
    const T* Q[4] = {prim_of<T>(b, 1), prim_of<T>(b, 2), prim_of<T>(b, 3), prim_of<T>(b, 5)};
    // tangential directions: i-faces (j,k), j-faces (i,k), k-faces (i,j)
    constexpr int ai = DIR == 0 ? 0 : 1, aj = DIR == 0 ? 1 : 0;          // first tangential unit step
    constexpr int bj = DIR == 2 ? 1 : 0, bk = DIR == 2 ? 0 : 1;          // second tangential unit step
    int pos = DIR == 0 ? i : DIR == 1 ? j : k;
    int nd = DIR == 0 ? b.nx : DIR == 1 ? b.ny : b.nz;
    int posA = DIR == 0 ? j : i, nA = DIR == 0 ? b.ny : b.nx;
    int posB = DIR == 2 ? j : k, nB = DIR == 2 ? b.ny : b.nz;
    bool zeroA = (pos == 1 || pos == nd) && (posA == 1 || posA == nA - 1);
    bool zeroB = (pos == 1 || pos == nd) && (posB == 1 || posB == nB - 1);
    T gi[4], gj[4], gk[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const T* a = Q[q];
        T dn = a[s] - a[sm];
        T ta1 = T(0), ta2 = T(0), tb1 = T(0), tb2 = T(0);
        if (!zeroA) {
            ta1 = T(0.5) * (a[IX(b, i - ai, j - aj, k)] + a[IX(b, im - ai, jm - aj, km)]);
            ta2 = T(0.5) * (a[IX(b, i + ai, j + aj, k)] + a[IX(b, im + ai, jm + aj, km)]);
        }
        if (!zeroB) {
            tb1 = T(0.5) * (a[IX(b, i, j - bj, k - bk)] + a[IX(b, im, jm - bj, km - bk)]);
            tb2 = T(0.5) * (a[IX(b, i, j + bj, k + bk)] + a[IX(b, im, jm + bj, km + bk)]);
        }
        T ga = T(0.5) * (ta2 - ta1), gb = T(0.5) * (tb2 - tb1);
        if (DIR == 0) { gi[q] = dn; gj[q] = ga; gk[q] = gb; }
        else if (DIR == 1) { gi[q] = ga; gj[q] = dn; gk[q] = gb; }
        else { gi[q] = ga; gj[q] = gb; gk[q] = dn; }
    }
    auto M = [&](int q) { return met_of<T>(b, DIR + 1, q)[s]; };
    T ix = M(0), iy = M(1), iz = M(2), jx = M(3), jy = M(4), jz = M(5), kx = M(6), ky = M(7), kz = M(8);
    T ux = (gi[0] * ix + gj[0] * jx) + gk[0] * kx, vx = (gi[1] * ix + gj[1] * jx) + gk[1] * kx;
    T wx = (gi[2] * ix + gj[2] * jx) + gk[2] * kx, Tx = (gi[3] * ix + gj[3] * jx) + gk[3] * kx;
    T uy = (gi[0] * iy + gj[0] * jy) + gk[0] * ky, vy = (gi[1] * iy + gj[1] * jy) + gk[1] * ky;
    T wy = (gi[2] * iy + gj[2] * jy) + gk[2] * ky, Ty = (gi[3] * iy + gj[3] * jy) + gk[3] * ky;
    T uz = (gi[0] * iz + gj[0] * jz) + gk[0] * kz, vz = (gi[1] * iz + gj[1] * jz) + gk[1] * kz;
    T wz = (gi[2] * iz + gj[2] * jz) + gk[2] * kz, Tz = (gi[3] * iz + gj[3] * jz) + gk[3] * kz;
    T mua = T(b.muf[s]), mub = T(b.muf[sm]), mta = T(b.mut[s]), mtb = T(b.mut[sm]);
    T mu0 = T(0.5) * (((mua + mta) + mub) + mtb);
    T PrL = T(p.PrL), PrT = T(p.PrT);
    T k0 = (T(0.5) * T(p.Cp)) * (((mua / PrL + mta / PrT) + mub / PrL) + mtb / PrT);
    T s11 = mu0 * (tmp1 * ux - tmp2 * (vy + wz)), s12 = mu0 * (uy + vx), s13 = mu0 * (uz + wx);
    T s22 = mu0 * (tmp1 * vy - tmp2 * (ux + wz)), s23 = mu0 * (vz + wy), s33 = mu0 * (tmp1 * wz - tmp2 * (ux + vy));
    T u1 = (Q[0][s] + Q[0][sm]) * T(0.5), v1 = (Q[1][s] + Q[1][sm]) * T(0.5), w1 = (Q[2][s] + Q[2][sm]) * T(0.5);
    T E1 = ((u1 * s11 + v1 * s12) + w1 * s13) + k0 * Tx;
    T E2 = ((u1 * s12 + v1 * s22) + w1 * s23) + k0 * Ty;
    T E3 = ((u1 * s13 + v1 * s23) + w1 * s33) + k0 * Tz;
    T s1x = nv_of<T>(b, 3 * DIR)[s], s1y = nv_of<T>(b, 3 * DIR + 1)[s], s1z = nv_of<T>(b, 3 * DIR + 2)[s];
    T Sf = S_of<T>(b, DIR)[s];
    T f1 = ((s11 * s1x + s12 * s1y) + s13 * s1z) * Sf;
    T f2 = ((s12 * s1x + s22 * s1y) + s23 * s1z) * Sf;
    T f3 = ((s13 * s1x + s23 * s1y) + s33 * s1z) * Sf;
    put_Fv(b, DIR, 0, s, f1); put_Fv(b, DIR, 1, s, f2); put_Fv(b, DIR, 2, s, f3);
    put_Fv(b, DIR, 3, s, ((E1 * s1x + E2 * s1y) + E3 * s1z) * Sf);
    // Surf1..6: viscous force on the first and last face of each direction
    if (pos == 1 || pos == nd) {
        int f = DIR + (pos == 1 ? 0 : 3);
        size_t m = DIR == 0 ? JK(b, j, k) : DIR == 1 ? IK(b, i, k) : IJ(b, i, j);
        b.surf[f][0][m] = f1; b.surf[f][1][m] = f2; b.surf[f][2][m] = f3;
    }
}

// Every flux through face (i,j,k) of direction DIR, one thread per face: the SA face terms
// (sub_turbulence_SA.f90:60-174), the inviscid flux and the viscous flux, each stored for the cell
// pass.  The SA terms go first, so their registers are free again for the reconstruction; the
// viscous flux, mostly loads of the face's gradient stencil and metrics, overlaps the inviscid
// flux's arithmetic in the same kernel.
template <int DIR, typename T, int RECON, int FLUX>
__global__ void __launch_bounds__(64) k_face(BSet S) {
    OCFD_LOCATE(SP_FI + DIR)
    if (p.turb == 2) {
        const size_t s = IX(b, i, j, k);
        T fv, f2;
        sa_face<DIR, T>(b, p, i, j, k, fv, f2);
        put_Fsa(b, DIR, 0, s, fv);
        put_Fsa(b, DIR, 1, s, f2);
    }
    inv_face<DIR, T, RECON, FLUX>(b, p, i, j, k);
    if (p.visc == 1) visc_face<DIR, T>(b, p, i, j, k);
}

// Res(1:5): inviscid i, j, k then viscous i, j, k face differences, in the reference order
// (sub_Residual.f90:346-354,507-516,692-701,864-872,1039-1048,1219-1229)
// followed by Res(6) of SA, then comput_Lijk + comput_dt (sub_time_acceleraction.f90:7-112)
// Res(m) goes to res(m, value) and the spectral radii of direction dir to lam(dir, Lc, Lv) as
// soon as they are formed, so a caller's stores free their registers
template <typename T, typename ResSink, typename LamSink>
__device__ __forceinline__ void cell_assemble(const Blk& b, const Par& p, int i, int j, int k, size_t s, ResSink&& res,
                                              LamSink&& lam, double* Lc, double* Lv, double& dt) {
    size_t sp[3] = {s + 1, s + (size_t)b.NI, s + (size_t)b.NI * b.NJ};
    bool visc = p.visc == 1 && b.flags[0] != 1;
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double r = get_F<T>(b, 0, m, sp[0]) - get_F<T>(b, 0, m, s);
        r = (r + get_F<T>(b, 1, m, sp[1])) - get_F<T>(b, 1, m, s);
        r = (r + get_F<T>(b, 2, m, sp[2])) - get_F<T>(b, 2, m, s);
        if (visc && m > 0) {
            r = (r + get_Fv<T>(b, 0, m - 1, sp[0])) - get_Fv<T>(b, 0, m - 1, s);
            r = (r + get_Fv<T>(b, 1, m - 1, sp[1])) - get_Fv<T>(b, 1, m - 1, s);
            r = (r + get_Fv<T>(b, 2, m - 1, sp[2])) - get_Fv<T>(b, 2, m - 1, s);
        }
        res(m, r);
    }
    if (p.turb == 2) res(5, sa_residual<T>(b, p, i, j, k, s));
    const double g43 = p.gamma > (double)(4.0f / 3.0f) ? p.gamma : (double)(4.0f / 3.0f);
    double d = b.prim[0][s], u = b.prim[1][s], v = b.prim[2][s], w = b.prim[3][s], c = b.prim[6][s];
    double D0 = p.visc == 1 ? (g43 / d) * (b.muf[s] / p.PrL + b.mut[s] / p.PrT) : 0.0;
    double vol = b.vol[s], vol1 = 1.0 / vol;
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        double un = 0.5 * ((u * (b.nv[3 * dir][s] + b.nv[3 * dir][sp[dir]]) + v * (b.nv[3 * dir + 1][s] + b.nv[3 * dir + 1][sp[dir]]))
                           + w * (b.nv[3 * dir + 2][s] + b.nv[3 * dir + 2][sp[dir]]));
        double S0 = (b.S[dir][s] + b.S[dir][sp[dir]]) * 0.5;
        Lc[dir] = (fabs(un) + c) * S0;
        Lv[dir] = ((D0 * S0) * S0) * vol1;
        lam(dir, Lc[dir], Lv[dir]);
    }
    if (p.local_dt == 1) {
        double dt_fac = b.flags[0] == 0 ? 1.0 : 0.1;
        dt = ((dt_fac * p.CFL) * vol) / (((Lc[0] + Lc[1]) + Lc[2]) + 1.0 * ((Lv[0] + Lv[1]) + Lv[2]));
        if (p.dtime_mesh == 1) dt = dt * b.dtm[s];
        if (dt > p.dtmax) dt = p.dtmax;
        if (dt < p.dtmin) dt = p.dtmin;
    } else {
        dt = p.dt_global;
    }
}

// the cell pass in box order: Res, Lc, Lv and dt into the block arrays
template <typename T>
__global__ void __launch_bounds__(64) k_cell(BSet S) {
    OCFD_LOCATE(SP_CELL)
    const size_t s = IX(b, i, j, k);
    double Lc[3], Lv[3], dt;
    cell_assemble<T>(
        b, p, i, j, k, s, [&](int m, double r) { b.Res[m][s] = r; },
        [&](int dir, double lc, double lv) { b.Lc[dir][s] = lc; b.Lv[dir][s] = lv; }, Lc, Lv, dt);
    b.dt[s] = dt;
    // du_LU_SGS's diagonal (sub_LU_SGS.f90:27-37 forward, :81-90 backward): of components 1..5 (and
    // 6, 7 of SST), and for SA of the SA component in the forward and in the backward sweep, which
    // sum the viscous radii in different orders
    const double vol = b.vol[s];
    const double sumLc = (Lc[0] + Lc[1]) + Lc[2], sumLv = (Lv[0] + Lv[1]) + Lv[2];
    double alfa = vol / dt + p.w_LU * sumLc;
    if (p.visc == 1) alfa = alfa + 2.0 * sumLv;
    b.al[0][s] = alfa + 0.0 * vol;
    if (p.NV == 6) {
        const double a1 = p.visc == 1 ? alfa + sumLv : alfa, a2 = p.visc == 1 ? ((alfa + Lv[0]) + Lv[1]) + Lv[2] : alfa;
        b.al[1][s] = a1 + 0.0 * vol;
        b.al[2][s] = a2 + 0.0 * vol;
    }
}


// ---------------------------------------------------------------------------------------------
// du_LU_SGS (sub_LU_SGS.f90:4-129): cells of hyperplane i+j+k = plane, one thread per (j,k)
template <int NV, bool FWD>
__global__ void k_lusgs_plane(Blk b, Par p, int plane) {
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
    if (j > b.ny - 1) return;
    int i = plane - k - j;
    if (i < 1 || i > b.nx - 1) return;
    size_t s = IX(b, i, j, k);
    double w = p.w_LU, gamma = p.gamma;
    bool visc = p.visc == 1;
    double sumLc = (b.Lc[0][s] + b.Lc[1][s]) + b.Lc[2][s];
    double sumLv = (b.Lv[0][s] + b.Lv[1][s]) + b.Lv[2][s];
    double alfa0 = b.vol[s] / b.dt[s] + w * sumLc;
    double alfa[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) alfa[m] = alfa0;
    if (visc) {
#pragma unroll
        for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 2.0 * sumLv;
        if (NV == 6) {
            if (FWD) alfa[5] = alfa[5] + sumLv;
            else alfa[5] = ((alfa[5] + b.Lv[0][s]) + b.Lv[1][s]) + b.Lv[2][s];
        }
    }
#pragma unroll
    for (int m = 0; m < NV; m++) alfa[m] = alfa[m] + 0.0 * b.vol[s];
    double acc[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) acc[m] = FWD ? b.Res[m][s] : 0.0;
    const size_t st[3] = {1, (size_t)b.NI, (size_t)b.NI * b.NJ};
    int pos[3] = {i, j, k};
    int lim[3] = {b.nx - 1, b.ny - 1, b.nz - 1};
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
        bool has = FWD ? (pos[dir] != 1) : (pos[dir] != lim[dir]);
        double du[NV];
        if (has) {
            size_t t = FWD ? s - st[dir] : s + st[dir];
            double Un[NV], DUn[NV], DF[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) { Un[m] = b.U[m][t]; DUn[m] = b.DU[m][t]; }
#pragma unroll
            for (int m = 5; m < NV; m++) if (Un[m] < 0.0) Un[m] = 0.0;     // limit_vt
            comput_dfn<NV>(DF, Un, DUn, b.nv[3 * dir][s], b.nv[3 * dir + 1][s], b.nv[3 * dir + 2][s], gamma);
            double wLc = w * b.Lc[dir][t], Lvn = b.Lv[dir][t];
            if (FWD) {
                double Sd = b.S[dir][s];
#pragma unroll
                for (int m = 0; m < NV; m++) du[m] = 0.5 * (DF[m] * Sd + wLc * DUn[m]);
            } else {
                double Sd = b.S[dir][t];
#pragma unroll
                for (int m = 0; m < NV; m++) du[m] = -0.5 * (DF[m] * Sd - wLc * DUn[m]);
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
        if (FWD) b.DU[m][s] = acc[m] / alfa[m];
        else b.DU[m][s] = b.DU[m][s] + acc[m] / alfa[m];
    }
}

// ---------------------------------------------------------------------------------------------
// U = Un + DU (sub_time_advance.f90:92-101), the limit_flow_U snapshot/detection
// (sub_limitflow.f90:46-94) and, for SA, limit_flow_SA's detection (:182-231).  The snapshot
// (rho,u,v,w,p) goes to prim[0..4]; U6 after the update goes to prim[5], U6 below SAmin is clamped
// and U6 above LSAmax is flagged in the block's mask.  limit_flow_U changes only U(1:5) and the SA
// pass only U6, so detecting both before either replaces anything gives the reference's order.
__device__ __forceinline__ long sa_cell_index(const Blk& b, int i, int j, int k) {
    return (long)(i - 1) + (long)(b.nx - 1) * ((long)(j - 1) + (long)(b.ny - 1) * (long)(k - 1));
}
__device__ __forceinline__ bool sa_flagged(const Blk& b, long c) { return (b.samask[c >> 6] >> (c & 63)) & 1ull; }

// limit_flow_U's primitive state of conservative variables U(1:5) (sub_limitflow.f90:46-60)
__device__ __forceinline__ void lim_prims(const double* U, double gamma, double& d1, double& u1, double& v1, double& w1, double& p1) {
    d1 = U[0]; u1 = U[1] / d1; v1 = U[2] / d1; w1 = U[3] / d1;
    p1 = (U[4] - (0.5 * d1) * ((u1 * u1 + v1 * v1) + w1 * w1)) * (gamma - 1.0);
}
__device__ __forceinline__ bool lim_out(const Par& p, double d1, double u1, double v1, double w1, double p1) {
    return d1 < p.Ldmin || p1 < p.Lpmin || d1 > p.Ldmax || p1 > p.Lpmax || fabs(u1) > p.Lumax || fabs(v1) > p.Lumax || fabs(w1) > p.Lumax;
}

__device__ __forceinline__ void update_cell(const Blk& b, const Par& p, int i, int j, int k, size_t s, const double* DU) {
    double U[MAXV];
    for (int m = 0; m < p.NV; m++) { U[m] = b.U[m][s] + DU[m]; b.U[m][s] = U[m]; }
    if (p.limit_flow != 1) return;
    double d1, u1, v1, w1, p1;
    lim_prims(U, p.gamma, d1, u1, v1, w1, p1);
    if (lim_out(p, d1, u1, v1, w1, p1)) atomicAdd(&b.flags[1], 1);
    if (p.NV == 6) {
        const double v = U[5];
        b.prim[5][s] = v;
        if (v > p.LSAmax) {
            atomicAdd(&b.flags[2], 1);
            long c = sa_cell_index(b, i, j, k);
            atomicOr(&b.samask[c >> 6], 1ull << (c & 63));
        } else if (v < 1.0e-8) {
            b.U[5][s] = 1.0e-8;
        }
    }
}

// the update in box order, DU from the block arrays
__global__ void __launch_bounds__(64) k_update_detect(BSet S) {
    if (halted(S)) return;
    OCFD_LOCATE(SP_CELL)
    const size_t s = IX(b, i, j, k);
    double DU[MAXV];
    for (int m = 0; m < p.NV; m++) DU[m] = b.DU[m][s];
    update_cell(b, p, i, j, k, s, DU);
}

// The update of an iteration with the pipelined sweeps: the backward sweep's DU comes from the
// sweep-ordered storage.  A workgroup takes one sweep tile's cells of an i-window (the tiles and
// windows of that storage), the lanes of UP_ROWS consecutive positions of a step's run per pass
// (one 128-byte line of each step).  A pass reads its lanes' DU along the steps and parks it in
// LDS; then its threads update the pass's cells, i fastest.  DU stays in the storage until the
// next iteration's pack; the block arrays get it only when the host asks for it.  Lane (jl, kl)
// holds cell i at step i - 1 + jl + kl.  After a sweep gave up, the update leaves U and the
// storage as they are; the host puts the sentinel back into the whole storage.
constexpr int UP_ROWS = 16, UP_T = UP_ROWS * PK_W;

template <int NV, int TJ>
__global__ void __launch_bounds__(UP_T) k_update_pk(BSet S, const LuTask* __restrict__ tiles, const double* __restrict__ pk,
                                                    const int2* __restrict__ wins) {
    using F = Pk<NV>;
    constexpr int DMAX = TJ + LU_NT / TJ - 2;                 // largest jl + kl of a tile
    constexpr int SLOTS = (PK_W + DMAX) * UP_ROWS;            // (step, position) slots a pass can touch
    __shared__ double buf[NV][UP_ROWS][PK_W + 1];
    if (halted(S)) return;
    const int2 wn = wins[xcd_tile((int)blockIdx.x, (int)gridDim.x)];     // (tile, i-window), windows fastest
    const LuTask T_ = tiles[wn.x];
    const Blk& b = S.blk[T_.blk];
    const Par& p = S.par[T_.blk];
    const int nxm = b.nx - 1;
    const int i0 = wn.y;
    const int W = min(PK_W, nxm - i0 + 1);
    const int t = (int)threadIdx.x, rr = t / PK_W, c = t % PK_W;
#pragma unroll 1
    for (int pos0 = 0; pos0 < LU_NT; pos0 += UP_ROWS) {
        // step i0 - 1 + qo: the lane at position pos0 + l holds cell i0 + qo - (jl + kl) there
        for (int slot = t; slot < SLOTS; slot += UP_T) {
            const int l = slot % UP_ROWS, qo = slot / UP_ROWS;
            const int lane = pos_lane<TJ>(pos0 + l);
            const int jl = lane % TJ, kl = lane / TJ, ci = qo - (jl + kl);
            if (jl < T_.tj && kl < T_.tk && ci >= 0 && ci < W) {
                const double* el = pk + T_.pofs + ((long long)(i0 - 1 + qo) * F::NF + F::DUB) * LU_NT + pos0 + l;
#pragma unroll
                for (int m = 0; m < NV; m++) buf[m][l][ci] = el[m * LU_NT];
            }
        }
        __syncthreads();
        const int lane = pos_lane<TJ>(pos0 + rr);
        const int jl = lane % TJ, kl = lane / TJ;
        if (jl < T_.tj && kl < T_.tk && c < W) {
            const int i = i0 + c, j = T_.j0 + jl, k = T_.k0 + kl;
            const size_t s = IX(b, i, j, k);
            double DU[MAXV];
#pragma unroll
            for (int m = 0; m < NV; m++) DU[m] = buf[m][rr][c];
            update_cell(b, p, i, j, k, s, DU);
        }
        __syncthreads();
    }
}

// the limiter's primitive state of cell s from U, and whether it is out of the limits
__device__ __forceinline__ bool lim_cell(const Blk& b, const Par& p, size_t s, double& d1, double& u1, double& v1, double& w1, double& p1) {
    const double U[5] = {b.U[0][s], b.U[1][s], b.U[2][s], b.U[3][s], b.U[4][s]};
    lim_prims(U, p.gamma, d1, u1, v1, w1, p1);
    return lim_out(p, d1, u1, v1, w1, p1);
}

// limit_flow_U replacement step (sub_limitflow.f90:84-150): bad cells take the mean of the valid
// 3x3x3 neighbours' primitive states, or their own clamped state; sets the IF_OverLimit latch.
// The states come from U after the update: the valid cells keep it through the step, and the
// new U of a bad cell goes to prim[0..4] (free after the residual) until k_limit_commit, once
// every bad cell has read its neighbours.  (One launch of a few thousand workgroups looping over
// the cell tiles: when no block has a bad cell, the usual case, it ends within microseconds.)
__device__ __forceinline__ void limit_apply_cell(const Blk& b, const Par& p, int i, int j, int k) {
    if (i == 1 && j == 1 && k == 1) b.flags[0] = 1;
    size_t s = IX(b, i, j, k);
    double ds, us, vs, ws, ps;
    if (!lim_cell(b, p, s, ds, us, vs, ws, ps)) return;
    double d0 = 0, u0 = 0, v0 = 0, w0 = 0, p0 = 0; int kn = 0;
    for (int ka = -1; ka <= 1; ka++) for (int ja = -1; ja <= 1; ja++) for (int ia = -1; ia <= 1; ia++) {
        int i1 = i + ia, j1 = j + ja, k1 = k + ka;
        if (i1 > 0 && i1 < b.nx && j1 > 0 && j1 < b.ny && k1 > 0 && k1 < b.nz) {
            double dt, ut, vt, wt, pt;
            if (!lim_cell(b, p, IX(b, i1, j1, k1), dt, ut, vt, wt, pt)) { d0 = d0 + dt; u0 = u0 + ut; v0 = v0 + vt; w0 = w0 + wt; p0 = p0 + pt; kn++; }
        }
    }
    if (kn == 0) {
        d0 = ds; u0 = us; v0 = vs; w0 = ws; p0 = ps;
        if (d0 < p.Ldmin) d0 = p.Ldmin;
        if (d0 > p.Ldmax) d0 = p.Ldmax;
        if (p0 < p.Lpmin) p0 = p.Lpmin;
        if (p0 > p.Lpmax) p0 = p.Lpmax;
        if (fabs(u0) > p.Lumax) u0 = copysign(p.Lumax, u0);
        if (fabs(v0) > p.Lumax) v0 = copysign(p.Lumax, v0);
        if (fabs(w0) > p.Lumax) w0 = copysign(p.Lumax, w0);
    } else {
        double sn1 = 1.0 / (double)kn;
        d0 = d0 * sn1; u0 = u0 * sn1; v0 = v0 * sn1; w0 = w0 * sn1; p0 = p0 * sn1;
    }
    b.prim[0][s] = d0; b.prim[1][s] = d0 * u0; b.prim[2][s] = d0 * v0; b.prim[3][s] = d0 * w0;
    b.prim[4][s] = p0 / (p.gamma - 1.0) + (0.5 * d0) * ((u0 * u0 + v0 * v0) + w0 * w0);
}

// the cell tiles of the blocks whose flag word `which` is nonzero, spread over the grid: every
// workgroup walks the block list (uniform, scalar loads) and skips an unflagged block whole
template <typename F>
__device__ __forceinline__ void flagged_tiles(const BSet& S, int nb, int ntiles, int which, F&& f) {
    for (int bi = 0; bi < nb; bi++) {
        const Blk& b = S.blk[bi];
        if (b.flags[which] == 0) continue;
        const int t1 = bi + 1 < nb ? S.tile0[bi + 1] : ntiles;
        for (int t = S.tile0[bi] + (int)blockIdx.x; t < t1; t += (int)gridDim.x) {
            int bj, i, j, k;
            if (locate_tile<SP_CELL>(S, t, bj, i, j, k)) f(b, S.par[bi], i, j, k);
        }
    }
}

__global__ void __launch_bounds__(64) k_limit_apply(BSet S, int nb, int ntiles) {
    if (halted(S)) return;
    flagged_tiles(S, nb, ntiles, 1, [](const Blk& b, const Par& p, int i, int j, int k) { limit_apply_cell(b, p, i, j, k); });
}

// the new U of the bad cells, once every bad cell has read its neighbours
__global__ void __launch_bounds__(64) k_limit_commit(BSet S, int nb, int ntiles) {
    if (halted(S)) return;
    flagged_tiles(S, nb, ntiles, 1, [](const Blk& b, const Par& p, int i, int j, int k) {
        const size_t s = IX(b, i, j, k);
        double d1, u1, v1, w1, p1;
        if (lim_cell(b, p, s, d1, u1, v1, w1, p1))
            for (int m = 0; m < 5; m++) b.U[m][s] = b.prim[m][s];
    });
}

// limit_flow_SA (sub_limitflow.f90:182-231).  The reference clamps and averages in place in
// lexicographic (k, j, i) order: U6 below SAmin is clamped, U6 above LSAmax is replaced by the
// mean of the admissible neighbours, the earlier neighbours with their values after this pass
// (clamped or already replaced), the later ones with their values before it.  The update kernel
// clamps in parallel (the original value is kept in prim[5]) and marks the flagged cells.  A
// flagged cell none of whose 26 neighbours is flagged reads only values that the pass does not
// change, so those cells are done in parallel first; the rest follow in order of
// w = i + 2j + 4k.  Every earlier neighbour in the reference's loop has a smaller w (dk = -1:
// di + 2dj - 4 < 0; dk = 0, dj = -1: di - 2 < 0; dk = dj = 0: di = -1) and two cells with the same
// w are never neighbours, so the cells of one w are done in parallel after all smaller w, and
// each reads exactly what the sequential loop gives it.
__device__ __forceinline__ void sa_replace(const Blk& b, const Par& p, int i, int j, int k) {
    const double SAmin = 1.0e-8;
    double* U6 = b.U[5];
    const double* orig = b.prim[5];
    double s0 = 0.0; int kn = 0;
    for (int ka = -1; ka <= 1; ka++) for (int ja = -1; ja <= 1; ja++) for (int ia = -1; ia <= 1; ia++) {
        int i1 = i + ia, j1 = j + ja, k1 = k + ka;
        if (i1 > 0 && i1 < b.nx && j1 > 0 && j1 < b.ny && k1 > 0 && k1 < b.nz) {
            size_t t = IX(b, i1, j1, k1);
            bool later = (k1 > k) || (k1 == k && (j1 > j || (j1 == j && i1 > i)));
            double v = later ? orig[t] : U6[t];
            if (v >= SAmin && v <= p.LSAmax) { s0 = s0 + v; kn++; }
        }
    }
    U6[IX(b, i, j, k)] = kn == 0 ? p.LSAmax : s0 / (double)kn;
}

__device__ __forceinline__ void limit_sa_isolated_cell(const Blk& b, const Par& p, int i, int j, int k) {
    long c = sa_cell_index(b, i, j, k);
    if (!sa_flagged(b, c)) return;
    for (int ka = -1; ka <= 1; ka++) for (int ja = -1; ja <= 1; ja++) for (int ia = -1; ia <= 1; ia++) {
        int i1 = i + ia, j1 = j + ja, k1 = k + ka;
        if ((ia | ja | ka) != 0 && i1 > 0 && i1 < b.nx && j1 > 0 && j1 < b.ny && k1 > 0 && k1 < b.nz &&
            sa_flagged(b, sa_cell_index(b, i1, j1, k1)))
            return;                                   // clustered: left to the ordered pass
    }
    sa_replace(b, p, i, j, k);
    // no flagged cell reads this cell's bit (it has no flagged neighbour), so it can go now
    atomicAnd(&b.samask[c >> 6], ~(1ull << (c & 63)));
}

__global__ void __launch_bounds__(64) k_limit_sa_isolated(BSet S, int nb, int ntiles) {
    if (halted(S)) return;
    flagged_tiles(S, nb, ntiles, 2, [](const Blk& b, const Par& p, int i, int j, int k) { limit_sa_isolated_cell(b, p, i, j, k); });
}

// one workgroup per block: the flagged cells left by the isolated pass, by increasing w over the
// box that bounds them; clears the mask
constexpr int SA_ORD_T = 256;
__global__ void __launch_bounds__(SA_ORD_T) k_limit_sa_ordered(BSet S) {
    if (halted(S)) return;
    const Blk& b = S.blk[blockIdx.x];
    const Par& p = S.par[blockIdx.x];
    if (b.flags[2] == 0) return;
    const int nxm = b.nx - 1, nym = b.ny - 1;
    const long ncell = (long)nxm * nym * (b.nz - 1);
    const long nword = (ncell + 63) / 64;
    __shared__ int lo[3], hi[3];
    if (threadIdx.x < 3) { lo[threadIdx.x] = 1 << 30; hi[threadIdx.x] = -1; }
    __syncthreads();
    int l0 = 1 << 30, l1 = 1 << 30, l2 = 1 << 30, h0 = -1, h1 = -1, h2 = -1;
    for (long wi = threadIdx.x; wi < nword; wi += SA_ORD_T) {
        unsigned long long m = b.samask[wi];
        while (m) {
            const int bit = __builtin_ctzll(m);
            m &= m - 1;
            const long cc = wi * 64 + bit;
            const int i = (int)(cc % nxm) + 1, j = (int)((cc / nxm) % nym) + 1, k = (int)(cc / ((long)nxm * nym)) + 1;
            l0 = min(l0, i); l1 = min(l1, j); l2 = min(l2, k);
            h0 = max(h0, i); h1 = max(h1, j); h2 = max(h2, k);
        }
    }
    if (h0 >= 0) {
        atomicMin(&lo[0], l0); atomicMin(&lo[1], l1); atomicMin(&lo[2], l2);
        atomicMax(&hi[0], h0); atomicMax(&hi[1], h1); atomicMax(&hi[2], h2);
    }
    __syncthreads();
    if (hi[0] < 0) return;                      // every flagged cell was isolated
    const int nj = hi[1] - lo[1] + 1, njk = nj * (hi[2] - lo[2] + 1);
    const int w0 = lo[0] + 2 * lo[1] + 4 * lo[2], w1 = hi[0] + 2 * hi[1] + 4 * hi[2];
    for (int w = w0; w <= w1; w++) {
        for (int a = (int)threadIdx.x; a < njk; a += SA_ORD_T) {
            const int j = lo[1] + a % nj, k = lo[2] + a / nj, i = w - 2 * j - 4 * k;
            if (i >= lo[0] && i <= hi[0] && sa_flagged(b, sa_cell_index(b, i, j, k))) sa_replace(b, p, i, j, k);
        }
        // the replaced values of this w before the next reads them
        __threadfence_block();
        __syncthreads();
    }
    for (long wi = threadIdx.x; wi < nword; wi += SA_ORD_T)
        if (b.samask[wi]) b.samask[wi] = 0ull;
}

__global__ void k_reset_flags(BSet S) {
    const Blk& b = S.blk[blockIdx.x];
    if (threadIdx.x == 0) { b.flags[1] = 0; b.flags[2] = 0; }
}

// force_vt_kw (sub_time_advance.f90:27-61)
__global__ void __launch_bounds__(64) k_force_vt_kw(BSet S) {
    if (halted(S)) return;
    OCFD_LOCATE(SP_CELL)
    size_t s = IX(b, i, j, k);
    for (int m = 5; m < p.NV; m++) if (b.U[m][s] < 0) b.U[m][s] = 1.0e-10;
}

// ---------------------------------------------------------------------------------------------
// Boundary conditions on one subface (sub_boundary.f90:6-504)
__device__ __forceinline__ void d_bc_wall(const Blk& b, const Par& p, const Sub& B, int a, int c) {   // boundary_wall + wall_bound (:87-147, :405-454)
    size_t s1, sg;
    switch (B.face) {
    case 1: case 4: { int j = B.jb + a, k = B.kb + c; if (j > B.je - 1 || k > B.ke - 1) return;
        s1 = IX(b, B.face == 1 ? 1 : B.ie - 1, j, k); sg = IX(b, B.face == 1 ? 0 : B.ie, j, k); break; }
    case 2: case 5: { int i = B.ib + a, k = B.kb + c; if (i > B.ie - 1 || k > B.ke - 1) return;
        s1 = IX(b, i, B.face == 2 ? 1 : B.je - 1, k); sg = IX(b, i, B.face == 2 ? 0 : B.je, k); break; }
    default: { int i = B.ib + a, j = B.jb + c; if (i > B.ie - 1 || j > B.je - 1) return;
        s1 = IX(b, i, j, B.face == 3 ? 1 : B.ke - 1); sg = IX(b, i, j, B.face == 3 ? 0 : B.ke); break; }
    }
    double g = p.gamma, Ma = p.Ma, U1[MAXV], Ug[MAXV];
    for (int m = 0; m < p.NV; m++) U1[m] = b.U[m][s1];
    if (p.Twall > 0.0) {
        double d1 = U1[0], uu1 = U1[1] / d1, v1 = U1[2] / d1, w1 = U1[3] / d1;
        double p1 = (U1[4] - (0.5 * d1) * ((uu1 * uu1 + v1 * v1) + w1 * w1)) * (g - 1.0);
        double T1 = ((g * Ma) * Ma) * p1 / d1;
        double p2 = p1, T2 = 2.0 * p.Twall - T1;
        if (T2 < 0.5 * T1) T2 = 0.5 * T1;
        double uu2 = -uu1, v2 = -v1, w2 = -w1;
        double d2 = ((g * Ma) * Ma) * p2 / T2;
        Ug[0] = d2; Ug[1] = d2 * uu2; Ug[2] = d2 * v2; Ug[3] = d2 * w2;
        Ug[4] = p2 / (g - 1.0) + (0.5 * d2) * ((uu2 * uu2 + v2 * v2) + w2 * w2);
    } else {
        Ug[0] = U1[0]; Ug[1] = -U1[1]; Ug[2] = -U1[2]; Ug[3] = -U1[3]; Ug[4] = U1[4];
    }
    if (p.NV == 6) Ug[5] = 0.0;
    else if (p.NV == 7) {
        const double beta1 = 0.075;
        double dw = b.dw[s1];
        double wt = (60.0 * b.muf[s1]) / (((U1[0] * beta1) * (dw * dw)) * p.Re);
        Ug[5] = -U1[5];
        Ug[6] = (U1[0] + Ug[0]) * wt - U1[6];
    }
    for (int m = 0; m < p.NV; m++) b.U[m][sg] = Ug[m];
}

// boundary_Farfield (:152-283); flag 0 far field, 1 forced outlet, 2 forced inlet.  The node
// range of the subface is inclusive, as in the reference loop.
__device__ __forceinline__ void d_bc_far(const Blk& b, const Par& p, const Sub& B, int flag, int a, int c) {
    int i, j, k;
    switch (B.face) {
    case 1: case 4: j = B.jb + a; k = B.kb + c; i = B.ib; if (j > B.je || k > B.ke) return; break;
    case 2: case 5: i = B.ib + a; k = B.kb + c; j = B.jb; if (i > B.ie || k > B.ke) return; break;
    default: i = B.ib + a; j = B.jb + c; k = B.kb; if (i > B.ie || j > B.je) return; break;
    }
    int i1 = i, j1 = j, k1 = k, i2 = i, j2 = j, k2 = k;
    size_t s = IX(b, i, j, k);
    double n1, n2, n3;
    switch (B.face) {
    case 1: i2 = i - 1; n1 = -b.nv[0][s]; n2 = -b.nv[1][s]; n3 = -b.nv[2][s]; break;
    case 2: j2 = j - 1; n1 = -b.nv[3][s]; n2 = -b.nv[4][s]; n3 = -b.nv[5][s]; break;
    case 3: k2 = k - 1; n1 = -b.nv[6][s]; n2 = -b.nv[7][s]; n3 = -b.nv[8][s]; break;
    case 4: i1 = i - 1; n1 = b.nv[0][s]; n2 = b.nv[1][s]; n3 = b.nv[2][s]; break;
    case 5: j1 = j - 1; n1 = b.nv[3][s]; n2 = b.nv[4][s]; n3 = b.nv[5][s]; break;
    default: k1 = k - 1; n1 = b.nv[6][s]; n2 = b.nv[7][s]; n3 = b.nv[8][s]; break;
    }
    size_t s1 = IX(b, i1, j1, k1), s2 = IX(b, i2, j2, k2);
    double g = p.gamma;
    double d_inf = 1.0, u_inf = cos(p.A_alfa) * cos(p.A_beta), v_inf = sin(p.A_alfa) * cos(p.A_beta), w_inf = sin(p.A_beta);
    double p_inf = 1.0 / (g * p.Ma * p.Ma);
    double d1 = b.U[0][s1], u1 = b.U[1][s1] / d1, v1 = b.U[2][s1] / d1, w1 = b.U[3][s1] / d1;
    double p1 = (b.U[4][s1] - (0.5 * d1) * ((u1 * u1 + v1 * v1) + w1 * w1)) * (g - 1.0);
    double c1 = sqrt(g * p1 / d1);
    double d2, u2, v2, w2, p2, Ma_n = 0.0;
    if (flag == 1) { d2 = d1; u2 = u1; v2 = v1; w2 = w1; p2 = p1; }
    else if (flag == 2) { d2 = d_inf; u2 = u_inf; v2 = v_inf; w2 = w_inf; p2 = p_inf; }
    else {
        if (p.P_outlet <= -1.0) Ma_n = ((u_inf * n1 + v_inf * n2) + w_inf * n3) * p.Ma;
        else Ma_n = ((u1 * n1 + v1 * n2) + w1 * n3) / c1;
        if (Ma_n > 1.0) { d2 = d1; u2 = u1; v2 = v1; w2 = w1; p2 = p1; }
        else if (Ma_n > -1.0e-6) {
            if (p.P_outlet > 0.0) {
                double pb = p.P_outlet;
                double db = d1 + (pb - p1) / (c1 * c1);
                double ub = u1 + ((p1 - pb) / (d1 * c1)) * n1, vb = v1 + ((p1 - pb) / (d1 * c1)) * n2, wb = w1 + ((p1 - pb) / (d1 * c1)) * n3;
                p2 = 2.0 * pb - p1; d2 = 2.0 * db - d1; u2 = 2.0 * ub - u1; v2 = 2.0 * vb - v1; w2 = 2.0 * wb - w1;
            } else { d2 = d1; u2 = u1; v2 = v1; w2 = w1; p2 = p1; }
        } else if (Ma_n > -1.0) {
            double pb = 0.5 * ((p1 + p_inf) - (d1 * c1) * (((u_inf - u1) * n1 + (v_inf - v1) * n2) + (w_inf - w1) * n3));
            double db = d_inf + (pb - p_inf) / (c1 * c1);
            double ub = u_inf - ((p_inf - pb) / (d1 * c1)) * n1, vb = v_inf - ((p_inf - pb) / (d1 * c1)) * n2, wb = w_inf - ((p_inf - pb) / (d1 * c1)) * n3;
            p2 = 2.0 * pb - p1; d2 = 2.0 * db - d1; u2 = 2.0 * ub - u1; v2 = 2.0 * vb - v1; w2 = 2.0 * wb - w1;
        } else { d2 = d_inf; u2 = u_inf; v2 = v_inf; w2 = w_inf; p2 = p_inf; }
    }
    double nv6 = 0, nv7 = 0;
    bool copy = Ma_n > 0.0 || flag == 1;
    if (p.NV >= 6) { nv6 = copy ? b.U[5][s1] : (p.NV == 6 ? 5.0 : d_inf * p.Kt_inf); }
    if (p.NV == 7) { nv7 = copy ? b.U[6][s1] : d_inf * p.Wt_inf; }
    b.U[0][s2] = d2; b.U[1][s2] = d2 * u2; b.U[2][s2] = d2 * v2; b.U[3][s2] = d2 * w2;
    b.U[4][s2] = p2 / (g - 1.0) + (0.5 * d2) * ((u2 * u2 + v2 * v2) + w2 * w2);
    if (p.NV >= 6) b.U[5][s2] = nv6;
    if (p.NV == 7) b.U[6][s2] = nv7;
}

__device__ __forceinline__ void d_bc_sym(const Blk& b, const Par& p, const Sub& B, int a, int c) {   // boundary_Symmetry_or_SlideWall (:291-373)
    size_t s1, s2, sn;
    int nb;
    switch (B.face) {
    case 1: case 4: { int j = B.jb + a, k = B.kb + c; if (j > B.je - 1 || k > B.ke - 1) return;
        int i = B.face == 1 ? B.ib : B.ie, i1 = B.face == 1 ? B.ib : B.ie - 1, i2 = B.face == 1 ? B.ib - 1 : B.ie;
        s1 = IX(b, i1, j, k); s2 = IX(b, i2, j, k); sn = IX(b, i, j, k); nb = 0; break; }
    case 2: case 5: { int i = B.ib + a, k = B.kb + c; if (i > B.ie - 1 || k > B.ke - 1) return;
        int j = B.face == 2 ? B.jb : B.je, j1 = B.face == 2 ? B.jb : B.je - 1, j2 = B.face == 2 ? B.jb - 1 : B.je;
        s1 = IX(b, i, j1, k); s2 = IX(b, i, j2, k); sn = IX(b, i, j, k); nb = 3; break; }
    default: { int i = B.ib + a, j = B.jb + c; if (i > B.ie - 1 || j > B.je - 1) return;
        int k = B.face == 3 ? B.kb : B.ke, k1 = B.face == 3 ? B.kb : B.ke - 1, k2 = B.face == 3 ? B.kb - 1 : B.ke;
        s1 = IX(b, i, j, k1); s2 = IX(b, i, j, k2); sn = IX(b, i, j, k); nb = 6; break; }
    }
    for (int m = 0; m < p.NV; m++) b.U[m][s2] = b.U[m][s1];
    double n1 = b.nv[nb][sn], n2 = b.nv[nb + 1][sn], n3 = b.nv[nb + 2][sn];
    double Vn = (b.U[1][s1] * n1 + b.U[2][s1] * n2) + b.U[3][s1] * n3;
    b.U[1][s2] = b.U[1][s2] - (2.0 * Vn) * n1;
    b.U[2][s2] = b.U[2][s2] - (2.0 * Vn) * n2;
    b.U[3][s2] = b.U[3][s2] - (2.0 * Vn) * n3;
}

__device__ __forceinline__ void d_bc_ext(const Blk& b, const Par& p, const Sub& B, int a, int c) {   // boundary_Extrapolate (:458-504)
    int i, j, k;
    switch (B.face) {
    case 1: case 4: j = B.jb + a; k = B.kb + c; i = B.ib; if (j > B.je || k > B.ke) return; break;
    case 2: case 5: i = B.ib + a; k = B.kb + c; j = B.jb; if (i > B.ie || k > B.ke) return; break;
    default: i = B.ib + a; j = B.jb + c; k = B.kb; if (i > B.ie || j > B.je) return; break;
    }
    int i1 = i, j1 = j, k1 = k, i2 = i, j2 = j, k2 = k;
    switch (B.face) {
    case 1: i2 = i - 1; break; case 2: j2 = j - 1; break; case 3: k2 = k - 1; break;
    case 4: i1 = i - 1; break; case 5: j1 = j - 1; break; default: k1 = k - 1; break;
    }
    for (int m = 0; m < p.NV; m++) b.U[m][IX(b, i2, j2, k2)] = b.U[m][IX(b, i1, j1, k1)];
}

// every physical subface of the set in one launch; the ghost cells written by different
// subfaces are disjoint and only interior cells are read
__global__ void __launch_bounds__(64) k_bc(SubSet S) {
    OCFD_SUB_LOCATE
    switch (T_.kind) {
    case SUB_WALL: d_bc_wall(b, p, B, a, c); break;
    case SUB_FAR: d_bc_far(b, p, B, T_.arg, a, c); break;
    case SUB_SYM: d_bc_sym(b, p, B, a, c); break;
    default: d_bc_ext(b, p, B, a, c); break;
    }
}

// Amut_boundary on every wall subface of the set
__global__ void __launch_bounds__(64) k_amut(SubSet S) {
    OCFD_SUB_LOCATE
    (void)p;
    d_amut(b, B, a, c);
}

// Umessage_send_mpi local copy (sub_update_buffer_mpi.f90:42-141): LAP ghost layers of the
// neighbour block from this block's interior, with the L1..L3 axis permutation and order.
__device__ __forceinline__ void d_halo(const Blk& src, const Blk& dst, const HaloMap& h, int NV, int t) {
    int n0 = h.ke1[0] - h.kb1[0] + 1, n1 = h.ke1[1] - h.kb1[1] + 1, n2 = h.ke1[2] - h.kb1[2] + 1;
    if (t >= n0 * n1 * n2) return;
    int ka[3] = {t % n0, (t / n0) % n1, t / (n0 * n1)};
    int i = h.kb1[0] + ka[0], j = h.kb1[1] + ka[1], k = h.kb1[2] + ka[2];
    int i1 = h.ks[0] + ka[h.L[0] - 1] * h.P[0], j1 = h.ks[1] + ka[h.L[1] - 1] * h.P[1], k1 = h.ks[2] + ka[h.L[2] - 1] * h.P[2];
    size_t d = IX(dst, i, j, k), s = IX(src, i1, j1, k1);
    for (int m = 0; m < NV; m++) dst.U[m][d] = src.U[m][s];
}

// single-pair launch, used while the geometry is set up
__global__ void k_halo(Blk src, Blk dst, HaloMap h, int NV) {
    d_halo(src, dst, h, NV, blockIdx.x * blockDim.x + threadIdx.x);
}

// every inner subface of the set in one launch (the ghost regions written are disjoint)
__global__ void __launch_bounds__(256) k_halo_set(HaloSet S, int NV) {
    const HaloTask& T = S.t[S.tile_task[blockIdx.x]];
    d_halo(S.blk[T.src], S.blk[T.dst], T.h, NV, (blockIdx.x - T.tile0) * 256 + threadIdx.x);
}

// get_U_conner (sub_update_buffer_mpi.f90:515-549): 12 edges, then 8 vertices
__device__ __forceinline__ void d_corner_edges(const Blk& b, int NV, int t, int e, int m) {
    int nx = b.nx, ny = b.ny, nz = b.nz;
    double* U = b.U[m];
    auto V = [&](int i, int j, int k) -> double& { return U[IX(b, i, j, k)]; };
    int k = t + 1, j = t + 1, i = t + 1;
    switch (e) {
    case 0: if (k <= nz - 1) V(0, 0, k) = (V(1, 0, k) + V(0, 1, k)) * 0.5; break;
    case 1: if (k <= nz - 1) V(nx, 0, k) = (V(nx - 1, 0, k) + V(nx, 1, k)) * 0.5; break;
    case 2: if (k <= nz - 1) V(0, ny, k) = (V(1, ny, k) + V(0, ny - 1, k)) * 0.5; break;
    case 3: if (k <= nz - 1) V(nx, ny, k) = (V(nx - 1, ny, k) + V(nx, ny - 1, k)) * 0.5; break;
    case 4: if (j <= ny - 1) V(0, j, 0) = (V(1, j, 0) + V(0, j, 1)) * 0.5; break;
    case 5: if (j <= ny - 1) V(nx, j, 0) = (V(nx - 1, j, 0) + V(nx, j, 1)) * 0.5; break;
    case 6: if (j <= ny - 1) V(0, j, nz) = (V(1, j, nz) + V(0, j, nz - 1)) * 0.5; break;
    case 7: if (j <= ny - 1) V(nx, j, nz) = (V(nx - 1, j, nz) + V(nx, j, nz - 1)) * 0.5; break;
    case 8: if (i <= nx - 1) V(i, 0, 0) = (V(i, 1, 0) + V(i, 0, 1)) * 0.5; break;
    case 9: if (i <= nx - 1) V(i, ny, 0) = (V(i, ny - 1, 0) + V(i, ny, 1)) * 0.5; break;
    case 10: if (i <= nx - 1) V(i, 0, nz) = (V(i, 1, nz) + V(i, 0, nz - 1)) * 0.5; break;
    default: if (i <= nx - 1) V(i, ny, nz) = (V(i, ny - 1, nz) + V(i, ny, nz - 1)) * 0.5; break;
    }
    (void)NV;
}

__device__ __forceinline__ void d_corner_vertices(const Blk& b, int c, int m) {
    if (c >= 8) return;
    int nx = b.nx, ny = b.ny, nz = b.nz;
    double* U = b.U[m];
    auto V = [&](int i, int j, int k) { return U[IX(b, i, j, k)]; };
    int I = (c & 1) ? nx : 0, J = (c & 2) ? ny : 0, K = (c & 4) ? nz : 0;
    int Ii = (c & 1) ? nx - 1 : 1, Jj = (c & 2) ? ny - 1 : 1, Kk = (c & 4) ? nz - 1 : 1;
    double r = (2.0 * ((V(Ii, J, K) + V(I, Jj, K)) + V(I, J, Kk)) - ((V(Ii, Jj, K) + V(Ii, J, Kk)) + V(I, Jj, Kk))) / 3.0;
    U[IX(b, I, J, K)] = r;
}

// edges then vertices of every block of the set (blockIdx.y: edge, blockIdx.z: block x variable);
// a workgroup walks its whole edge
__global__ void k_corner_edges(BSet S, int NV) {
    const Blk& b = S.blk[blockIdx.z / NV];
    const int len = max(b.nx, max(b.ny, b.nz));
    for (int t = (int)threadIdx.x; t < len; t += (int)blockDim.x) d_corner_edges(b, NV, t, blockIdx.y, blockIdx.z % NV);
}
__global__ void k_corner_vertices(BSet S, int NV) {
    d_corner_vertices(S.blk[blockIdx.x / NV], threadIdx.x, blockIdx.x % NV);
}
// single-block forms, used while the geometry is set up
__global__ void k_corner_edges1(Blk b, int NV) {
    d_corner_edges(b, NV, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, blockIdx.z);
}
__global__ void k_corner_vertices1(Blk b) { d_corner_vertices(b, threadIdx.x, blockIdx.x); }

// ---------------------------------------------------------------------------------------------
// comput_max_Res_onemesh (sub_time_advance.f90:330-380): per block, component and part (a fixed
// slice of the block's cells) the max |Res|, the sum of squares and a NaN flag, each part reduced
// in a fixed tree order (deterministic, independent of co-scheduled work)
__global__ void k_norms(const NormBlk* tab, int NV, double* part) {
    const NormBlk& b = tab[blockIdx.z];
    __shared__ double smax[256], ssum[256];
    __shared__ int snan[256];
    int m = blockIdx.y;
    long ncell = (long)(b.nx - 1) * (b.ny - 1) * (b.nz - 1);
    double mx = 0.0, sm = 0.0; int nan = 0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
        int i = (int)(c % (b.nx - 1)) + 1, j = (int)((c / (b.nx - 1)) % (b.ny - 1)) + 1, k = (int)(c / ((long)(b.nx - 1) * (b.ny - 1))) + 1;
        double r = b.res[m][(size_t)(i + 3) + (size_t)b.NI * ((size_t)(j + 3) + (size_t)b.NJ * (size_t)(k + 3))];
        if (r != r) nan = 1;
        mx = fmax(mx, fabs(r));
        sm = sm + r * r;
    }
    smax[threadIdx.x] = mx; ssum[threadIdx.x] = sm; snan[threadIdx.x] = nan;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + w];
            snan[threadIdx.x] |= snan[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = part + 3 * (((size_t)blockIdx.z * NV + m) * gridDim.x + blockIdx.x);
        o[0] = smax[0]; o[1] = ssum[0]; o[2] = snan[0];
    }
    (void)NV;
}

// the parts of every block combined in the order of the reference's block loop: per component,
// blocks in order and each block's parts in order; out[v] = max |Res_v|, out[NV + v] = sum of
// squares, out[2 NV + v] = 1 when a NaN was seen
__global__ void k_norms_final(int nb, int NV, int P, const double* part, double* out) {
    const int v = (int)threadIdx.x;
    if (v >= NV) return;
    double mx = 0.0, sm = 0.0, nan = 0.0;
    for (int b = 0; b < nb; b++)
        for (int q = 0; q < P; q++) {
            const double* o = part + 3 * (((size_t)b * NV + v) * P + q);
            mx = mx < o[0] ? o[0] : mx;
            sm = sm + o[1];
            if (o[2] != 0.0) nan = 1.0;
        }
    out[v] = mx; out[NV + v] = sm; out[2 * NV + v] = nan;
}

// comput_force (sub_Post.f90:4-265): the pressure force, the viscous force (Surf1..6) and their
// moments about the centroid, one thread per wall face.  Every workgroup of 64 faces sums its
// faces in a fixed tree order into one part; k_force_final adds the parts in tile order (blocks,
// subfaces, face rows): deterministic, though not the reference's serial order.
__global__ void __launch_bounds__(64) k_force_parts(SubSet S, double cx, double cy, double cz, double* part) {
    OCFD_SUB_LOCATE
    const int f = B.face, d = (f - 1) % 3;
    int lo1, hi1, lo2, hi2;
    if (d == 0) { lo1 = B.kb; hi1 = B.ke - 1; lo2 = B.jb; hi2 = B.je - 1; }
    else if (d == 1) { lo1 = B.kb; hi1 = B.ke - 1; lo2 = B.ib; hi2 = B.ie - 1; }
    else { lo1 = B.jb; hi1 = B.je - 1; lo2 = B.ib; hi2 = B.ie - 1; }
    const int o = lo1 + c, in = lo2 + a;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (o <= hi1 && in <= hi2) {
        size_t s1, s0, sf, sm;
        if (d == 0) { int j = in, k = o; s1 = IX(b, f == 1 ? 1 : b.nx - 1, j, k); s0 = IX(b, f == 1 ? 0 : b.nx, j, k); sf = IX(b, f == 1 ? 1 : b.nx, j, k); sm = JK(b, j, k); }
        else if (d == 1) { int i = in, k = o; s1 = IX(b, i, f == 2 ? 1 : b.ny - 1, k); s0 = IX(b, i, f == 2 ? 0 : b.ny, k); sf = IX(b, i, f == 2 ? 1 : b.ny, k); sm = IK(b, i, k); }
        else { int i = in, j = o; s1 = IX(b, i, j, f == 3 ? 1 : b.nz - 1); s0 = IX(b, i, j, f == 3 ? 0 : b.nz); sf = IX(b, i, j, f == 3 ? 1 : b.nz); sm = IJ(b, i, j); }
        const double g = p.gamma, p_inf = 1.0 / (g * p.Ma * p.Ma);
        double p1 = (g - 1.0) * (b.U[4][s1] - ((b.U[1][s1] * b.U[1][s1] + b.U[2][s1] * b.U[2][s1]) + b.U[3][s1] * b.U[3][s1]) / b.U[0][s1]);
        double p2 = (g - 1.0) * (b.U[4][s0] - ((b.U[1][s0] * b.U[1][s0] + b.U[2][s0] * b.U[2][s0]) + b.U[3][s0] * b.U[3][s0]) / b.U[0][s0]);
        double Pw = (f <= 3) ? -(0.5 * (p1 + p2) - p_inf) * b.S[d][sf] : (0.5 * (p1 + p2) - p_inf) * b.S[d][sf];
        double n1 = b.nv[3 * d][sf], n2 = b.nv[3 * d + 1][sf], n3 = b.nv[3 * d + 2][sf];
        double sg = f <= 3 ? 1.0 : -1.0;
        double sv0 = sg * b.surf[f - 1][0][sm], sv1 = sg * b.surf[f - 1][1][sm], sv2 = sg * b.surf[f - 1][2][sm];
        v[0] = Pw * n1; v[1] = Pw * n2; v[2] = Pw * n3;
        v[3] = sv0; v[4] = sv1; v[5] = sv2;
        double fx0 = Pw * n1 + sv0, fy0 = Pw * n2 + sv1, fz0 = Pw * n3 + sv2;
        double xc = (b.xc[s1] + b.xc[s0]) * 0.5 - cx, yc = (b.yc[s1] + b.yc[s0]) * 0.5 - cy, zc = (b.zc[s1] + b.zc[s0]) * 0.5 - cz;
        v[6] = yc * fz0 - zc * fy0;
        v[7] = zc * fx0 - xc * fz0;
        v[8] = xc * fy0 - yc * fx0;
    }
    __shared__ double sh[9][64];
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int q = 0; q < 9; q++) sh[q][t] = v[q];
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int q = 0; q < 9; q++) sh[q][t] = sh[q][t] + sh[q][t + w];
        }
        __syncthreads();
    }
    if (t < 9) part[(size_t)blockIdx.x * 9 + t] = sh[t][0];
}

__global__ void k_force_final(int nparts, const double* part, double* out9) {
    const int q = (int)threadIdx.x;
    if (q >= 9) return;
    double acc = 0.0;
    for (int w = 0; w < nparts; w++) acc = acc + part[(size_t)w * 9 + q];
    out9[q] = acc;
}

// ---------------------------------------------------------------------------------------------
// host launchers.  Every launch covers all blocks of the set; the per-block conditions of the
// reference (turbulence model, limiter, NVAR) are uniform over a set.
static inline dim3 g3(int n0, int n1, int n2, int bs) { return dim3((unsigned)((n0 + bs - 1) / bs), (unsigned)n1, (unsigned)n2); }
static inline BSet bset(const SetPlan& P, int sp) { return BSet{P.blk, P.par, P.tile_blk[sp], P.tile0[sp], P.halt}; }

template <int SP>
static void launch_space(void (*kern)(BSet), const SetPlan& P, hipStream_t st) {
    if (P.ntiles[SP] > 0) hipLaunchKernelGGL(kern, dim3((unsigned)P.ntiles[SP]), dim3(64), 0, st, bset(P, SP));
}

static void launch_subs(void (*kern)(SubSet), const SetPlan& P, const SubList& L, hipStream_t st) {
    if (L.ntiles > 0) hipLaunchKernelGGL(kern, dim3((unsigned)L.ntiles), dim3(64), 0, st, SubSet{P.blk, P.par, L.t, L.tile_task});
}

void set_pre(const SetPlan& P, hipStream_t st) {
    const Par& p = P.p0;
    hipLaunchKernelGGL(k_reset_flags, dim3((unsigned)P.nb), dim3(64), 0, st, bset(P, SP_CELL));
    launch_space<SP_BOX>(k_prims, P, st);      // with get_viscous on the interior cells
    if (p.visc != 1) return;
    launch_subs(k_visc_ghost, P, P.vghost, st);
    if (p.turb == 2) {
        launch_space<SP_N0>(k_sa_mut, P, st);
        launch_subs(k_amut, P, P.amut, st);
    } else if (p.turb == 3) {
        if (p.mixed) launch_space<SP_CELL>(k_sst_cell<float>, P, st);
        else launch_space<SP_CELL>(k_sst_cell<double>, P, st);
        int m3 = P.mx + 3;
        for (int f = 0; f < 6; f++)
            hipLaunchKernelGGL(k_sst_ghost, dim3((unsigned)((m3 + 63) / 64), (unsigned)m3, (unsigned)P.nb), dim3(64), 0, st, bset(P, SP_CELL), f);
        launch_subs(k_amut, P, P.amut, st);
        if (p.mixed) {
            launch_space<SP_FI>(k_sst_face<0, float>, P, st);
            launch_space<SP_FJ>(k_sst_face<1, float>, P, st);
            launch_space<SP_FK>(k_sst_face<2, float>, P, st);
        } else {
            launch_space<SP_FI>(k_sst_face<0, double>, P, st);
            launch_space<SP_FJ>(k_sst_face<1, double>, P, st);
            launch_space<SP_FK>(k_sst_face<2, double>, P, st);
        }
        launch_space<SP_CELL>(k_sst_res, P, st);
        launch_space<SP_N0>(k_sst_finish, P, st);
        launch_subs(k_amut, P, P.amut, st);
    } else {
        launch_subs(k_amut, P, P.amut, st);
    }
}

template <typename T, int R, int FL>
static void faces_rf(const SetPlan& P, hipStream_t st) {
    launch_space<SP_FI>(k_face<0, T, R, FL>, P, st);
    launch_space<SP_FJ>(k_face<1, T, R, FL>, P, st);
    launch_space<SP_FK>(k_face<2, T, R, FL>, P, st);
}
// the reconstruction variable set (IFlag_Reconstruction) and, for the primitive-variable
// reconstruction, the 1-D flux as template arguments: each kernel holds only its own
// reconstruction and Riemann solver
template <typename T>
static void faces_t(const SetPlan& P, hipStream_t st) {
    if (P.p0.recon == 0) {
        switch (P.p0.flux) {
        case F_SW: faces_rf<T, 0, F_SW>(P, st); break;
        case F_HLL: faces_rf<T, 0, F_HLL>(P, st); break;
        case F_HLLC: faces_rf<T, 0, F_HLLC>(P, st); break;
        case F_ROE: faces_rf<T, 0, F_ROE>(P, st); break;
        case F_VL: faces_rf<T, 0, F_VL>(P, st); break;
        default: faces_rf<T, 0, F_AUSM>(P, st); break;
        }
    } else if (P.p0.recon == 1) {
        faces_rf<T, 1, 0>(P, st);
    } else {
        faces_rf<T, 2, 0>(P, st);
    }
    launch_subs(k_inv_kbc<T>, P, P.kbc, st);
}
void set_inviscid(const SetPlan& P, hipStream_t st) {
    if (P.p0.mixed) faces_t<float>(P, st);
    else faces_t<double>(P, st);
}

template <typename T>
static void viscous_dt_t(const SetPlan& P, hipStream_t st) {
    launch_space<SP_CELL>(k_cell<T>, P, st);
}
void set_viscous_dt(const SetPlan& P, hipStream_t st) {
    if (P.p0.mixed) viscous_dt_t<float>(P, st);
    else viscous_dt_t<double>(P, st);
}


// limit_flow_U's replacement and limit_flow_SA after the update
static void limit_after_update(const SetPlan& P, hipStream_t st) {
    if (P.p0.limit_flow != 1) return;
    const int n = P.ntiles[SP_CELL];
    if (n <= 0) return;
    const dim3 grid((unsigned)std::min(n, 4096));
    hipLaunchKernelGGL(k_limit_apply, grid, dim3(64), 0, st, bset(P, SP_CELL), P.nb, n);
    hipLaunchKernelGGL(k_limit_commit, grid, dim3(64), 0, st, bset(P, SP_CELL), P.nb, n);
    if (P.p0.NV == 6) {
        hipLaunchKernelGGL(k_limit_sa_isolated, grid, dim3(64), 0, st, bset(P, SP_CELL), P.nb, n);
        hipLaunchKernelGGL(k_limit_sa_ordered, dim3((unsigned)P.nb), dim3(SA_ORD_T), 0, st, bset(P, SP_CELL));
    }
}

void set_update_limit(const SetPlan& P, hipStream_t st) {
    launch_space<SP_CELL>(k_update_detect, P, st);
    limit_after_update(P, st);
}

template <int NV, int TJ>
static void update_pk(const SetPlan& P, const LuPlan& L, hipStream_t st) {
    if (L.nwins > 0)
        hipLaunchKernelGGL((k_update_pk<NV, TJ>), dim3((unsigned)L.nwins), dim3(UP_T), 0, st, bset(P, SP_CELL), L.fwd,
                           L.pk, L.win);
}
void set_update_limit_sweep_order(const SetPlan& P, const LuPlan& L, hipStream_t st) {
    switch (P.p0.NV * 100 + L.tj) {
    case 504: update_pk<5, 4>(P, L, st); break;
    case 508: update_pk<5, 8>(P, L, st); break;
    case 516: update_pk<5, 16>(P, L, st); break;
    case 604: update_pk<6, 4>(P, L, st); break;
    case 608: update_pk<6, 8>(P, L, st); break;
    case 616: update_pk<6, 16>(P, L, st); break;
    case 704: update_pk<7, 4>(P, L, st); break;
    case 708: update_pk<7, 8>(P, L, st); break;
    default: update_pk<7, 16>(P, L, st); break;
    }
    limit_after_update(P, st);
}

void set_bc(const SetPlan& P, hipStream_t st) { launch_subs(k_bc, P, P.bc, st); }

void set_halo(const SetPlan& P, hipStream_t st) {
    if (P.halo.ntiles > 0)
        hipLaunchKernelGGL(k_halo_set, dim3((unsigned)P.halo.ntiles), dim3(256), 0, st, HaloSet{P.blk, P.halo.t, P.halo.tile_task}, P.p0.NV);
    hipLaunchKernelGGL(k_corner_edges, dim3(1, 12, (unsigned)(P.p0.NV * P.nb)), dim3(64), 0, st, bset(P, SP_CELL), P.p0.NV);
    hipLaunchKernelGGL(k_corner_vertices, dim3((unsigned)(P.p0.NV * P.nb)), dim3(64), 0, st, bset(P, SP_CELL), P.p0.NV);
}

// force_vt_kw (sub_time_advance.f90:27-61): U6 (and U7) below 0 set to 1e-10 in the cells.  With SA
// and the limiter on, the update has already raised every cell's U6 below 1e-8 to 1e-8 and the SA
// replacement writes positive values only, so the pass cannot change anything and is not launched.
void set_force_vt_kw(const SetPlan& P, hipStream_t st) {
    if (P.p0.NV == 6 && P.p0.limit_flow == 1) return;
    if (P.p0.NV >= 6) launch_space<SP_CELL>(k_force_vt_kw, P, st);
}

void launch_norms(const NormBlk* tab, int nb, int NV, double* part, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_norms, dim3((unsigned)norms_parts(), (unsigned)NV, (unsigned)nb), dim3(256), 0, st, tab, NV, part);
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(64), 0, st, nb, NV, norms_parts(), (const double*)part, out);
}

// ---- per-block launches: the hyperplane LU-SGS schedule, the setup-time halo, forces ----
template <int NV>
static void lusgs_nv(const Blk& b, const Par& p, bool fwd, hipStream_t st) {
    int np = b.nx + b.ny + b.nz - 3;
    dim3 grid((unsigned)((b.ny - 1 + 63) / 64), (unsigned)(b.nz - 1)), blk(64);
    if (fwd)
        for (int plane = 3; plane <= np; plane++) hipLaunchKernelGGL((k_lusgs_plane<NV, true>), grid, blk, 0, st, b, p, plane);
    else
        for (int plane = np; plane >= 3; plane--) hipLaunchKernelGGL((k_lusgs_plane<NV, false>), grid, blk, 0, st, b, p, plane);
}

void launch_lusgs_sweep(const Blk& b, const Par& p, bool forward, hipStream_t st) {
    if (p.NV == 5) lusgs_nv<5>(b, p, forward, st);
    else if (p.NV == 6) lusgs_nv<6>(b, p, forward, st);
    else lusgs_nv<7>(b, p, forward, st);
}

HaloMap halo_map(const Sub& B) {
    HaloMap h;
    int kb[3] = {B.ib, B.jb, B.kb}, ke[3] = {B.ie - 1, B.je - 1, B.ke - 1};
    int k = (B.face - 1) % 3;
    if (B.face > 3) kb[k] -= LAP;
    ke[k] = kb[k] + LAP - 1;
    int kb1[3] = {B.ib1, B.jb1, B.kb1}, ke1[3] = {B.ie1 - 1, B.je1 - 1, B.ke1 - 1};
    k = (B.face1 - 1) % 3;
    if (B.face1 <= 3) kb1[k] -= LAP;
    ke1[k] = kb1[k] + LAP - 1;
    int Ls[3] = {B.L1, B.L2, B.L3};
    for (int q = 0; q < 3; q++) {
        h.kb[q] = kb[q]; h.kb1[q] = kb1[q]; h.ke1[q] = ke1[q];
        h.L[q] = std::abs(Ls[q]); h.P[q] = Ls[q] >= 0 ? 1 : -1;
        h.ks[q] = h.P[q] > 0 ? kb[q] : ke[q];
    }
    return h;
}

int halo_points(const HaloMap& h) {
    return (h.ke1[0] - h.kb1[0] + 1) * (h.ke1[1] - h.kb1[1] + 1) * (h.ke1[2] - h.kb1[2] + 1);
}

void launch_halo(const Blk& src, const Blk& dst, const Sub& B, int NV, hipStream_t st) {
    HaloMap h = halo_map(B);
    int n = halo_points(h);
    hipLaunchKernelGGL(k_halo, GRID1(n, 256), 0, st, src, dst, h, NV);
}

void launch_forces(const SetPlan& P, const double cen[3], double* part, double* out9, hipStream_t st) {
    if (P.force.ntiles > 0)
        hipLaunchKernelGGL(k_force_parts, dim3((unsigned)P.force.ntiles), dim3(64), 0, st,
                           SubSet{P.blk, P.par, P.force.t, P.force.tile_task}, cen[0], cen[1], cen[2], part);
    hipLaunchKernelGGL(k_force_final, dim3(1), dim3(64), 0, st, P.force.ntiles, (const double*)part, out9);
}

// ---------------------------------------------------------------------------------------------
// self-check of the shared-reciprocal division: n operand pairs from a counter-based hash, the
// numerator's exponent within +-2^emax, the divisor's within +-2^40 (signs and exact zeros
// included); counts the quotients whose bits differ from the compiler's division
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}
__device__ __forceinline__ double hashed(unsigned long long h, int emax) {
    const double m = 1.0 + (double)(h >> 11) * (1.0 / 9007199254740992.0);    // [1, 2)
    const int e = (int)((h >> 3) % (unsigned long long)(2 * emax + 1)) - emax;
    return (h & 1) ? -ldexp(m, e) : ldexp(m, e);
}
__global__ void k_check_division(long long n, unsigned long long seed, int emax, unsigned long long* bad) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const unsigned long long h1 = mix64(seed ^ (2 * (unsigned long long)i)), h2 = mix64(seed ^ (2 * (unsigned long long)i + 1));
        const double a = (h1 % 97 == 0) ? 0.0 : hashed(h1, emax), b = hashed(h2, 40);
        const double q1 = div_nr(a, b, recip_nr(b)), q2 = a / b;
        // the two zeros compare equal: a zero quotient may differ in sign only
        if (q1 != q2 && __double_as_longlong(q1) != __double_as_longlong(q2)) atomicAdd(bad, 1ull);
    }
}

int check_division(long long n, unsigned long long seed, int emax, long long* mismatches) {
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof(unsigned long long)) != hipSuccess) return -1;
    unsigned long long h = 0;
    hipError_t e = hipMemset(d, 0, sizeof(h));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_division, dim3(2048), dim3(256), 0, 0, n, seed, emax, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return -1;
    *mismatches = (long long)h;
    return 0;
}

}  // namespace ocfd
