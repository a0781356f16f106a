// Device physics of the OpenCFD-EC hot path: face reconstruction, 1-D flux splitting and the
// LU-SGS flux difference.  Every expression keeps the reference's evaluation order (Fortran
// left-to-right, parentheses below) so the FP64 path reproduces the CPU build bit for bit
// when compiled without contraction.  Citations are file:line in the OpenCFD-EC tree.
#pragma once
#include <hip/hip_runtime.h>

namespace ocfd {

enum Scheme { S_UD1 = 0, S_NND2 = 1, S_UD3 = 2, S_MUSCL2U = 3, S_MUSCL2C = 4, S_MUSCL3 = 5,
              S_OMUSCL2 = 6, S_WENO5 = 7, S_UD5 = 8, S_WENO7 = 9, S_CD2 = 20 };
enum FluxId { F_SW = 1, F_HLL = 2, F_HLLC = 3, F_ROE = 4, F_VL = 5, F_AUSM = 6 };

template <typename T> __device__ __forceinline__ T sq(T a) { return a * a; }

// minmod (sub_scheme.f90:281-292)
template <typename T> __device__ __forceinline__ T minmod(T a, T b) {
    if (a * b <= T(0)) return T(0);
    if (fabs(a) <= fabs(b)) return a;
    return b;
}

// u points at u(0) of a stencil u(-3..4) (face between u(0) and u(1)).
// WENO5-JS left/right values (sub_scheme.f90:101-140)
// Double division a / b with the divisor's reciprocal shared: v_rcp_f64 refined by two Newton
// steps, then q = a r corrected once by the residual a - b q.  These are the steps of the
// compiler's IEEE division without its operand pre-scaling (v_div_scale) and special-value
// fix-up (v_div_fixup), which change nothing when a, b and a / b lie well inside the normal range
// (|a| above about 1e-290 and |b| between about 1e-290 and 1e290), so the bits are those of a / b
// there; divisions by one divisor then cost one reciprocal between them.
__device__ __forceinline__ double recip_nr(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double div_nr(double a, double b, double r) {
    const double q = a * r;
    const double e = __builtin_fma(-b, q, a);
    return __builtin_fma(e, r, q);
}

// a / b in the compute type: the shared-reciprocal sequence above for double, the compiler's
// division for float; rcp_of / dvr share one reciprocal between divisions by b
template <typename T> __device__ __forceinline__ T dvd(T a, T b) {
    if constexpr (sizeof(T) == 8) return div_nr(a, b, recip_nr(b));
    else return a / b;
}
template <typename T> __device__ __forceinline__ T rcp_of(T b) {
    if constexpr (sizeof(T) == 8) return recip_nr(b);
    else return T(0);
}
template <typename T> __device__ __forceinline__ T dvr(T a, T b, T r) {
    if constexpr (sizeof(T) == 8) return div_nr(a, b, r);
    else return a / b;
}

template <typename T> __device__ __forceinline__ T weno5_L(const T* u) {
    const T ep = T(1.0e-6), C03 = T(3.0 / 10.0), C13 = T(3.0 / 5.0), C23 = T(1.0 / 10.0);
    const T c1312 = T(13.0 / 12.0), c14 = T(0.25);
    T S0 = c1312 * sq((u[0] - T(2) * u[1]) + u[2]) + c14 * sq((T(3) * u[0] - T(4) * u[1]) + u[2]);
    T S1 = c1312 * sq((u[-1] - T(2) * u[0]) + u[1]) + c14 * sq(u[-1] - u[1]);
    T S2 = c1312 * sq((u[-2] - T(2) * u[-1]) + u[0]) + c14 * sq((u[-2] - T(4) * u[-1]) + T(3) * u[0]);
    T a0 = dvd(C03, sq(ep + S0)), a1 = dvd(C13, sq(ep + S1)), a2 = dvd(C23, sq(ep + S2));
    T am = (a0 + a1) + a2;
    T q03 = (T(1.0 / 3.0) * u[0] + T(5.0 / 6.0) * u[1]) - T(1.0 / 6.0) * u[2];
    T q13 = (-T(1.0 / 6.0) * u[-1] + T(5.0 / 6.0) * u[0]) + T(1.0 / 3.0) * u[1];
    T q23 = (T(1.0 / 3.0) * u[-2] - T(7.0 / 6.0) * u[-1]) + T(11.0 / 6.0) * u[0];
    return dvd((a0 * q03 + a1 * q13) + a2 * q23, am);
}
template <typename T> __device__ __forceinline__ T weno5_R(const T* u) {
    const T ep = T(1.0e-6), C03 = T(3.0 / 10.0), C13 = T(3.0 / 5.0), C23 = T(1.0 / 10.0);
    const T c1312 = T(13.0 / 12.0), c14 = T(0.25);
    T S0 = c1312 * sq((u[1] - T(2) * u[0]) + u[-1]) + c14 * sq((T(3) * u[1] - T(4) * u[0]) + u[-1]);
    T S1 = c1312 * sq((u[2] - T(2) * u[1]) + u[0]) + c14 * sq(u[2] - u[0]);
    T S2 = c1312 * sq((u[3] - T(2) * u[2]) + u[1]) + c14 * sq((u[3] - T(4) * u[2]) + T(3) * u[1]);
    T a0 = dvd(C03, sq(ep + S0)), a1 = dvd(C13, sq(ep + S1)), a2 = dvd(C23, sq(ep + S2));
    T am = (a0 + a1) + a2;
    T q03 = (T(1.0 / 3.0) * u[1] + T(5.0 / 6.0) * u[0]) - T(1.0 / 6.0) * u[-1];
    T q13 = (-T(1.0 / 6.0) * u[2] + T(5.0 / 6.0) * u[1]) + T(1.0 / 3.0) * u[0];
    T q23 = (T(1.0 / 3.0) * u[3] - T(7.0 / 6.0) * u[2]) + T(11.0 / 6.0) * u[1];
    return dvd((a0 * q03 + a1 * q13) + a2 * q23, am);
}

// scheme_fP / scheme_fm (sub_scheme.f90:6-96); U points at u(0)
template <typename T> __device__ __forceinline__ T scheme_fP(const T* U, int scheme) {
    const T k2c = T(1), k2u = T(-1), b2 = T(2), k3 = T(1.0 / 3.0), ep = T(1.0e-6);
    const T eta = T((double)0.05f);
    switch (scheme) {
    case S_CD2: return T(0.5) * (U[0] + U[1]);
    case S_UD1: return U[0];
    case S_UD3: return ((-U[-1] + T(5) * U[0]) + T(2) * U[1]) / T(6);
    case S_NND2: return U[0] + T(0.5) * minmod(U[0] - U[-1], U[1] - U[0]);
    case S_MUSCL2U:
        return U[0] + T(0.25) * ((T(1) - k2u) * minmod(U[0] - U[-1], b2 * (U[1] - U[0])) + (T(1) + k2u) * minmod(U[1] - U[0], b2 * (U[0] - U[-1])));
    case S_MUSCL2C:
        return U[0] + T(0.25) * ((T(1) - k2c) * minmod(U[0] - U[-1], b2 * (U[1] - U[0])) + (T(1) + k2c) * minmod(U[1] - U[0], b2 * (U[0] - U[-1])));
    case S_MUSCL3: {
        T up = U[1] - U[0], um = U[0] - U[-1];
        T s = (T(2) * up * um + ep) / ((up * up + um * um) + ep);
        return U[0] + (T(0.25) * s) * ((T(1) - k3 * s) * um + (T(1) + k3 * s) * up);
    }
    case S_OMUSCL2: {
        T r1 = ((U[0] - U[-1]) + ep) / ((U[1] - U[0]) + ep), r2 = ((U[1] - U[0]) + ep) / ((U[2] - U[1]) + ep);
        T f1 = ((T(1) - eta) + ((eta - T(0.55)) * T(0.5)) / r2) + ((eta + T(0.55)) * T(0.5)) * r1;
        T mn = T(2); if (f1 < mn) mn = f1; if (T(2) * r1 < mn) mn = T(2) * r1;
        T f = mn > T(0) ? mn : T(0);
        return U[0] + (T(0.5) * f) * (U[1] - U[0]);
    }
    case S_WENO5: return weno5_L(U);
    case S_UD5: return ((((T(2) * U[-2] - T(13) * U[-1]) + T(47) * U[0]) + T(27) * U[1]) - T(3) * U[2]) / T(60);
    default: return U[0];
    }
}
template <typename T> __device__ __forceinline__ T scheme_fm(const T* U, int scheme) {
    const T k2c = T(1), k2u = T(-1), b2 = T(2), k3 = T(1.0 / 3.0), ep = T(1.0e-6);
    const T eta = T((double)0.05f);
    switch (scheme) {
    case S_CD2: return T(0.5) * (U[0] + U[1]);
    case S_UD1: return U[1];
    case S_UD3: return ((T(2) * U[0] + T(5) * U[1]) - U[2]) / T(6);
    case S_NND2: return U[1] - T(0.5) * minmod(U[1] - U[0], U[2] - U[1]);
    case S_MUSCL2U:
        return U[1] - T(0.25) * ((T(1) - k2u) * minmod(U[2] - U[1], b2 * (U[1] - U[0])) + (T(1) + k2u) * minmod(U[1] - U[0], b2 * (U[2] - U[1])));
    case S_MUSCL2C:
        return U[1] - T(0.25) * ((T(1) - k2c) * minmod(U[2] - U[1], b2 * (U[1] - U[0])) + (T(1) + k2c) * minmod(U[1] - U[0], b2 * (U[2] - U[1])));
    case S_MUSCL3: {
        T up = U[2] - U[1], um = U[1] - U[0];
        T s = (T(2) * up * um + ep) / ((up * up + um * um) + ep);
        return U[1] - (T(0.25) * s) * ((T(1) - k3 * s) * up + (T(1) + k3 * s) * um);
    }
    case S_OMUSCL2: {
        T r1 = ((U[2] - U[1]) + ep) / ((U[1] - U[0]) + ep), r2 = ((U[1] - U[0]) + ep) / ((U[0] - U[-1]) + ep);
        T f1 = ((T(1) - eta) + ((eta - T(0.55)) * T(0.5)) / r2) + ((eta + T(0.55)) * T(0.5)) * r1;
        T mn = T(2); if (f1 < mn) mn = f1; if (T(2) * r1 < mn) mn = T(2) * r1;
        T f = mn > T(0) ? mn : T(0);
        return U[1] - (T(0.5) * f) * (U[1] - U[0]);
    }
    case S_WENO5: return weno5_R(U);
    case S_UD5: return ((((T(-3) * U[-1] + T(27) * U[0]) + T(47) * U[1]) - T(13) * U[2]) + T(2) * U[3]) / T(60);
    default: return U[1];
    }
}

// ---------------------------------------------------------------------------------------------
// 1-D fluxes in the rotated face frame (sub_flux_split.f90).  Q = (rho, u_n, u_t1, u_t2, p).

template <typename T>
__device__ __forceinline__ void flux_hll_hllc(const T* QL, const T* QR, T* F, T gamma, int iflux) {  // :13-95
    T dl = QL[0], uul = QL[1], vvl = QL[2], wwl = QL[3], pl = QL[4];
    T dr = QR[0], uur = QR[1], vvr = QR[2], wwr = QR[3], pr = QR[4];
    T UL[5], UR[5], Fl[5], Fr[5];
    UL[0] = dl; UL[1] = dl * uul; UL[2] = dl * vvl; UL[3] = dl * wwl;
    UL[4] = pl / (gamma - T(1)) + (dl * ((uul * uul + vvl * vvl) + wwl * wwl)) * T(0.5);
    UR[0] = dr; UR[1] = dr * uur; UR[2] = dr * vvr; UR[3] = dr * wwr;
    UR[4] = pr / (gamma - T(1)) + (dr * ((uur * uur + vvr * vvr) + wwr * wwr)) * T(0.5);
    T al = sqrt(gamma * pl / dl), ar = sqrt(gamma * pr / dr);
    T p_pvrs = T(0.5) * (pl + pr) - (((uur - uul) * (dl + dr)) * (al + ar)) * T(0.125);
    T p_star = p_pvrs > T(0) ? p_pvrs : T(0);
    T qql = (p_star <= pl) ? T(1) : sqrt(T(1) + ((gamma + T(1)) / (T(2) * gamma)) * (p_star / pl - T(1)));
    T qqr = (p_star <= pr) ? T(1) : sqrt(T(1) + ((gamma + T(1)) / (T(2) * gamma)) * (p_star / pr - T(1)));
    T Sl = uul - al * qql, Sr = uur + ar * qqr;
    Fl[0] = UL[1]; Fl[1] = UL[1] * uul + pl; Fl[2] = UL[2] * uul; Fl[3] = UL[3] * uul; Fl[4] = uul * (UL[4] + pl);
    Fr[0] = UR[1]; Fr[1] = UR[1] * uur + pr; Fr[2] = UR[2] * uur; Fr[3] = UR[3] * uur; Fr[4] = uur * (UR[4] + pr);
    if (iflux == F_HLL) {
        if (Sl >= T(0)) { for (int m = 0; m < 5; m++) F[m] = Fl[m]; }
        else if (Sr <= T(0)) { for (int m = 0; m < 5; m++) F[m] = Fr[m]; }
        else { for (int m = 0; m < 5; m++) F[m] = ((Sr * Fl[m] - Sl * Fr[m]) + (Sl * Sr) * (UR[m] - UL[m])) / (Sr - Sl); }
        return;
    }
    (void)wwr;
    if (Sl >= T(0)) { for (int m = 0; m < 5; m++) F[m] = Fl[m]; return; }
    if (Sr <= T(0)) { for (int m = 0; m < 5; m++) F[m] = Fr[m]; return; }
    T S_star = ((pr - pl + (dl * uul) * (Sl - uul)) - (dr * uur) * (Sr - uur)) / (dl * (Sl - uul) - dr * (Sr - uur));
    if (S_star >= T(0)) {
        T tmpl = dl * (Sl - uul) / (Sl - S_star);
        F[0] = Fl[0] + Sl * (tmpl - UL[0]);
        F[1] = Fl[1] + Sl * (tmpl * S_star - UL[1]);
        F[2] = Fl[2] + Sl * (tmpl * vvl - UL[2]);
        F[3] = Fl[3] + Sl * (tmpl * wwl - UL[3]);
        T tmp = UL[4] / dl + (S_star - uul) * (S_star + pl / (dl * (Sl - uul)));
        F[4] = Fl[4] + Sl * (tmpl * tmp - UL[4]);
    } else {
        T tmpr = dr * (Sr - uur) / (Sr - S_star);
        F[0] = Fr[0] + Sr * (tmpr - UR[0]);
        F[1] = Fr[1] + Sr * (tmpr * S_star - UR[1]);
        F[2] = Fr[2] + Sr * (tmpr * vvr - UR[2]);
        F[3] = Fr[3] + Sr * (tmpr * vvr - UR[3]);   // the reference uses vvr here (sub_flux_split.f90:67)
        T tmp = UR[4] / dr + (S_star - uur) * (S_star + pr / (dr * (Sr - uur)));
        F[4] = Fr[4] + Sr * (tmpr * tmp - UR[4]);
    }
}

template <typename T>
__device__ __forceinline__ void flux_sw(const T* QL, const T* QR, T* F, T gamma) {  // :102-141
    T dl = QL[0], uul = QL[1], vvl = QL[2], wwl = QL[3], pl = QL[4];
    T dr = QR[0], uur = QR[1], vvr = QR[2], wwr = QR[3], pr = QR[4];
    T al = sqrt(gamma * pl / dl), ar = sqrt(gamma * pr / dr);
    T tmp1 = T(2) * (gamma - T(1)), tmp3 = (T(3) - gamma) / (T(2) * (gamma - T(1)));
    T E1P = (uul + fabs(uul)) * T(0.5), E2P = ((uul - al) + fabs(uul - al)) * T(0.5), E3P = ((uul + al) + fabs(uul + al)) * T(0.5);
    T tmp0 = dl / (T(2) * gamma);
    T fp0 = tmp0 * ((tmp1 * E1P + E2P) + E3P);
    T fp1 = tmp0 * (((tmp1 * E1P) * uul + E2P * (uul - al)) + E3P * (uul + al));
    T fp2 = tmp0 * (((tmp1 * E1P) * vvl + E2P * vvl) + E3P * vvl);
    T fp3 = tmp0 * (((tmp1 * E1P) * wwl + E2P * wwl) + E3P * wwl);
    T fp4 = tmp0 * ((((E1P * (gamma - T(1))) * ((uul * uul + vvl * vvl) + wwl * wwl) + (E2P * ((sq(uul - al) + vvl * vvl) + wwl * wwl)) * T(0.5))
                     + (E3P * ((sq(uul + al) + vvl * vvl) + wwl * wwl)) * T(0.5)) + ((tmp3 * al) * al) * (E2P + E3P));
    T E1M = (uur - fabs(uur)) * T(0.5), E2M = ((uur - ar) - fabs(uur - ar)) * T(0.5), E3M = ((uur + ar) - fabs(uur + ar)) * T(0.5);
    tmp0 = dr / (T(2) * gamma);
    T fm0 = tmp0 * ((tmp1 * E1M + E2M) + E3M);
    T fm1 = tmp0 * (((tmp1 * E1M) * uur + E2M * (uur - ar)) + E3M * (uur + ar));
    T fm2 = tmp0 * (((tmp1 * E1M) * vvr + E2M * vvr) + E3M * vvr);
    T fm3 = tmp0 * (((tmp1 * E1M) * wwr + E2M * wwr) + E3M * wwr);
    T fm4 = tmp0 * ((((E1M * (gamma - T(1))) * ((uur * uur + vvr * vvr) + wwr * wwr) + (E2M * ((sq(uur - ar) + vvr * vvr) + wwr * wwr)) * T(0.5))
                     + (E3M * ((sq(uur + ar) + vvr * vvr) + wwr * wwr)) * T(0.5)) + ((tmp3 * ar) * ar) * (E2M + E3M));
    F[0] = fp0 + fm0; F[1] = fp1 + fm1; F[2] = fp2 + fm2; F[3] = fp3 + fm3; F[4] = fp4 + fm4;
}

template <typename T>
__device__ __forceinline__ void flux_roe(const T* QL, const T* QR, T* F, T gamma) {  // :148-210
    const T delt = T(0.1);
    const T eps8 = T((double)1.e-8f);   // single-precision literal 1.e-8 (:195,197)
    T dl = QL[0], uul = QL[1], vvl = QL[2], wwl = QL[3], pl = QL[4];
    T dr = QR[0], uur = QR[1], vvr = QR[2], wwr = QR[3], pr = QR[4];
    T hl = gamma * pl / ((gamma - T(1)) * dl) + T(0.5) * ((uul * uul + vvl * vvl) + wwl * wwl);
    T hr = gamma * pr / ((gamma - T(1)) * dr) + T(0.5) * ((uur * uur + vvr * vvr) + wwr * wwr);
    T UL1 = dl * uul, UL2 = dl * vvl, UL3 = dl * wwl;
    T UL4 = pl / (gamma - T(1)) + (dl * ((uul * uul + vvl * vvl) + wwl * wwl)) * T(0.5);
    T UR1 = dr * uur, UR2 = dr * vvr, UR3 = dr * wwr;
    T UR4 = pr / (gamma - T(1)) + (dr * ((uur * uur + vvr * vvr) + wwr * wwr)) * T(0.5);
    T Fl0 = UL1, Fl1 = UL1 * uul + pl, Fl2 = UL2 * uul, Fl3 = UL3 * uul, Fl4 = uul * (UL4 + pl);
    T Fr0 = UR1, Fr1 = UR1 * uur + pr, Fr2 = UR2 * uur, Fr3 = UR3 * uur, Fr4 = uur * (UR4 + pr);
    T tmpp = sqrt(dr / dl);
    const T den = T(1) + tmpp, rden = rcp_of(den);
    T avu = dvr(uul + tmpp * uur, den, rden);
    T avv = dvr(vvl + tmpp * vvr, den, rden);
    T avw = dvr(wwl + tmpp * wwr, den, rden);
    T avh = dvr(hl + tmpp * hr, den, rden);
    T ava = sqrt((gamma - T(1)) * (avh - T(0.5) * ((avu * avu + avv * avv) + avw * avw)));
    T l0 = fabs(avu - ava), l1 = fabs(avu), l2 = fabs(avu + ava);
    if (l0 < delt) l0 = (sq(l0) + delt * delt) / (T(2) * delt);
    if (l1 < delt) l1 = (sq(l1) + delt * delt) / (T(2) * delt);
    if (l2 < delt) l2 = (sq(l2) + delt * delt) / (T(2) * delt);
    T ddr = dr - dl;
    T a1 = (UR2 - UL2) / (avv + eps8) - ddr;
    T a2 = (UR3 - UL3) / (avw + eps8) - ddr;
    T a3 = (gamma - T(1)) * (((((ddr * (((avh - avu * avu) - avv * avv) - avw * avw) + avu * (UR1 - UL1)) + avv * (UR2 - UL2))
                             + avw * (UR3 - UL3)) - (UR4 - UL4))) / (ava * ava);
    T a0 = ((((avu + ava) * ddr - UR1) + UL1) - ava * a3) / (T(2) * ava);
    T a4 = ddr - (a0 + a3);
    F[0] = T(0.5) * (Fl0 + Fr0) - T(0.5) * ((l0 * a0 + l1 * a3) + l2 * a4);
    F[1] = T(0.5) * (Fl1 + Fr1) - T(0.5) * ((((l0 * a0) * (avu - ava) + (l1 * a3) * avu) + (l2 * a4) * (avu + ava)));
    F[2] = T(0.5) * (Fl2 + Fr2) - T(0.5) * (((((l0 * a0) * avv + (l1 * a1) * avv) + (l1 * a3) * avv) + (l2 * a4) * avv));
    F[3] = T(0.5) * (Fl3 + Fr3) - T(0.5) * (((((l0 * a0) * avw + (l1 * a2) * avw) + (l1 * a3) * avw) + (l2 * a4) * avw));
    F[4] = T(0.5) * (Fl4 + Fr4) - T(0.5) * ((((((l0 * a0) * (avh - avu * ava) + ((l1 * a1) * avv) * avv) + ((l1 * a2) * avw) * avw)
                                               + ((l1 * a3) * T(0.5)) * ((avu * avu + avv * avv) + avw * avw)) + (l2 * a4) * (avh + avu * ava)));
}

template <typename T>
__device__ __forceinline__ void flux_vl(const T* QL, const T* QR, T* F, T gamma) {  // :217-275
    T dl = QL[0], uul = QL[1], vvl = QL[2], wwl = QL[3], pl = QL[4];
    T dr = QR[0], uur = QR[1], vvr = QR[2], wwr = QR[3], pr = QR[4];
    T al = sqrt(gamma * pl / dl), ar = sqrt(gamma * pr / dr);
    T Ml = uul / al, Mr = uur / ar;
    T fp[5] = {T(0), T(0), T(0), T(0), T(0)}, fm[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (Ml >= T(1)) {
        fp[0] = dl * uul; fp[1] = (dl * uul) * uul + pl; fp[2] = (dl * uul) * vvl; fp[3] = (dl * uul) * wwl;
        fp[4] = uul * (gamma * pl / (gamma - T(1)) + (T(0.5) * dl) * ((uul * uul + vvl * vvl) + wwl * wwl));
    } else if (fabs(Ml) < T(1)) {
        T Mp = (T(0.25) * (T(1) + Ml)) * (T(1) + Ml), tmp0 = (dl * al) * Mp;
        fp[0] = tmp0;
        fp[1] = tmp0 * ((gamma - T(1)) * uul + T(2) * al) / gamma;
        fp[2] = tmp0 * vvl; fp[3] = tmp0 * wwl;
        fp[4] = tmp0 * ((((gamma - T(1)) * uul + T(2) * al) * ((gamma - T(1)) * uul + T(2) * al)) * T(0.5) / (gamma * gamma - T(1)) + T(0.5) * (vvl * vvl + wwl * wwl));
    }
    if (Mr >= T(1)) {
    } else if (fabs(Mr) < T(1)) {
        T Mm = (T(-0.25) * (Mr - T(1))) * (Mr - T(1)), tmp0 = (dr * ar) * Mm;
        fm[0] = tmp0;
        fm[1] = tmp0 * ((gamma - T(1)) * uur - T(2) * ar) / gamma;
        fm[2] = tmp0 * vvr; fm[3] = tmp0 * wwr;
        fm[4] = tmp0 * ((((gamma - T(1)) * uur - T(2) * ar) * ((gamma - T(1)) * uur - T(2) * ar)) * T(0.5) / (gamma * gamma - T(1)) + T(0.5) * (vvr * vvr + wwr * wwr));
    } else if (Mr <= T(-1)) {
        fm[0] = dr * uur; fm[1] = (dr * uur) * uur + pr; fm[2] = (dr * uur) * vvr; fm[3] = (dr * uur) * wwr;
        fm[4] = uur * (gamma * pr / (gamma - T(1)) + (T(0.5) * dr) * ((uur * uur + vvr * vvr) + wwr * wwr));
    }
    for (int m = 0; m < 5; m++) F[m] = fp[m] + fm[m];
}

template <typename T> __device__ __forceinline__ T fmin2(T a, T b) { return a < b ? a : b; }

template <typename T>
__device__ __forceinline__ void flux_ausmpw(const T* QL, const T* QR, T* F, T gamma) {  // :356-442 (flux id 6)
    T dl = QL[0], uul = QL[1], vvl = QL[2], wwl = QL[3], pl = QL[4];
    T dr = QR[0], uur = QR[1], vvr = QR[2], wwr = QR[3], pr = QR[4];
    T hl = gamma * pl / ((gamma - T(1)) * dl) + T(0.5) * ((uul * uul + vvl * vvl) + wwl * wwl);
    T hr = gamma * pr / ((gamma - T(1)) * dr) + T(0.5) * ((uur * uur + vvr * vvr) + wwr * wwr);
    T al = sqrt(gamma * pl / dl), ar = sqrt(gamma * pr / dr);
    T a = T(0.5) * (al + ar), Ml = uul / a, Mr = uur / a;
    T mlr = fmin2(pl / pr, pr / pl);
    T wpl = T(1) - (mlr * mlr) * mlr;            // x**3.d0 evaluates as (x*x)*x
    T plpl = (T(0.75) <= mlr && T(1) > mlr) ? T(4) * mlr - T(3) : T(0);
    T mrl = fmin2(pr / pl, pl / pr);
    T plpr = (T(0.75) <= mrl && T(1) > mrl) ? T(4) * mrl - T(3) : T(0);
    T Pp5, Pm5, Mp4, Mm4, fL, fR;
    if (fabs(Ml) >= T(1)) Pp5 = T(0.5) * (Ml + fabs(Ml)) / Ml;
    else Pp5 = ((T(0.25) * (Ml + T(1))) * (Ml + T(1))) * (T(2) - Ml) + (((T(3) * Ml) * (Ml * Ml - T(1))) * (Ml * Ml - T(1))) / T(16);
    if (fabs(Mr) >= T(1)) Pm5 = T(0.5) * (Mr - fabs(Mr)) / Mr;
    else Pm5 = ((T(0.25) * (Mr - T(1))) * (Mr - T(1))) * (T(2) + Mr) - (((T(3) * Mr) * (Mr * Mr - T(1))) * (Mr * Mr - T(1))) / T(16);
    T p = pl * Pp5 + pr * Pm5;
    if (fabs(Ml) >= T(1)) { Mp4 = T(0.5) * (Ml + fabs(Ml)); fL = T(0); }
    else {
        Mp4 = (T(0.25) * (Ml + T(1))) * (Ml + T(1)) + (T(0.125) * (Ml * Ml - T(1))) * (Ml * Ml - T(1));
        fL = (((pl / p - T(1)) * plpl) * fabs((T(0.25) * (Ml + T(1))) * (Ml + T(1)))) * fmin2(T(1), pow(sqrt((uul * uul + vvl * vvl) + wwl * wwl) / a, T(0.25)));
    }
    if (fabs(Mr) >= T(1)) { Mm4 = T(0.5) * (Mr - fabs(Mr)); fR = T(0); }
    else {
        Mm4 = (T(-0.25) * (Mr - T(1))) * (Mr - T(1)) - (T(0.125) * (Mr * Mr - T(1))) * (Mr * Mr - T(1));
        fR = (((pr / p - T(1)) * plpr) * fabs((T(-0.25) * (Mr - T(1))) * (Mr - T(1)))) * fmin2(T(1), pow(sqrt((uur * uur + vvr * vvr) + wwr * wwr) / a, T(0.25)));
    }
    T M = Mp4 + Mm4, MLp, MRm;
    if (M >= T(0)) MLp = ((Mp4 + Mm4) - (Mm4 * wpl) * (T(1) + fR)) + (fL * Mp4 + fR * Mm4);
    else MLp = (Mp4 * wpl) * (T(1) + fL);
    if (M < T(0)) MRm = ((Mp4 + Mm4) - (Mp4 * wpl) * (T(1) + fL)) + (fL * Mp4 + fR * Mm4);
    else MRm = (Mm4 * wpl) * (T(1) + fR);
    T fpv[5] = {dl, dl * uul, dl * vvl, dl * wwl, dl * hl}, fmv[5] = {dr, dr * uur, dr * vvr, dr * wwr, dr * hr};
    for (int m = 0; m < 5; m++) F[m] = (MLp * a) * fpv[m] + (MRm * a) * fmv[m];
    F[1] = F[1] + p;
}

template <typename T>
__device__ __forceinline__ void riemann(const T* QL, const T* QR, T* F, T gamma, int iflux) {
    if (iflux == F_SW) flux_sw(QL, QR, F, gamma);
    else if (iflux == F_ROE) flux_roe(QL, QR, F, gamma);
    else if (iflux == F_VL) flux_vl(QL, QR, F, gamma);
    else if (iflux == F_AUSM) flux_ausmpw(QL, QR, F, gamma);   // id 6 calls AUSMPW (sub_Residual.f90:327)
    else flux_hll_hllc(QL, QR, F, gamma, iflux);
}

// comput_DFn (sub_LU_SGS.f90:135-158): DF = F(U+DU) - F(U), convective part, face normal n
template <int NV>
__device__ __forceinline__ void comput_dfn(double* DF, const double* U, const double* DU, double n1, double n2, double n3, double gamma) {
    double U2[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) U2[m] = U[m] + DU[m];
    double un1 = ((U[1] * n1 + U[2] * n2) + U[3] * n3) / U[0];
    double p1 = (gamma - 1.0) * (U[4] - (0.5 * ((U[1] * U[1] + U[2] * U[2]) + U[3] * U[3])) / U[0]);
    double un2 = ((U2[1] * n1 + U2[2] * n2) + U2[3] * n3) / U2[0];
    double p2 = (gamma - 1.0) * (U2[4] - (0.5 * ((U2[1] * U2[1] + U2[2] * U2[2]) + U2[3] * U2[3])) / U2[0]);
    DF[0] = U2[0] * un2 - U[0] * un1;
    DF[1] = (U2[1] * un2 + p2 * n1) - (U[1] * un1 + p1 * n1);
    DF[2] = (U2[2] * un2 + p2 * n2) - (U[2] * un1 + p1 * n2);
    DF[3] = (U2[3] * un2 + p2 * n3) - (U[3] * un1 + p1 * n3);
    DF[4] = (U2[4] + p2) * un2 - (U[4] + p1) * un1;
#pragma unroll
    for (int m = 5; m < NV; m++) DF[m] = U2[m] * un2 - U[m] * un1;
}

// the same expression with the two divisions by each density sharing one reciprocal
template <int NV>
__device__ __forceinline__ void comput_dfn_nr(double* DF, const double* U, const double* DU, double n1, double n2, double n3, double gamma) {
    double U2[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) U2[m] = U[m] + DU[m];
    const double r1 = recip_nr(U[0]), r2 = recip_nr(U2[0]);
    double un1 = div_nr((U[1] * n1 + U[2] * n2) + U[3] * n3, U[0], r1);
    double p1 = (gamma - 1.0) * (U[4] - div_nr(0.5 * ((U[1] * U[1] + U[2] * U[2]) + U[3] * U[3]), U[0], r1));
    double un2 = div_nr((U2[1] * n1 + U2[2] * n2) + U2[3] * n3, U2[0], r2);
    double p2 = (gamma - 1.0) * (U2[4] - div_nr(0.5 * ((U2[1] * U2[1] + U2[2] * U2[2]) + U2[3] * U2[3]), U2[0], r2));
    DF[0] = U2[0] * un2 - U[0] * un1;
    DF[1] = (U2[1] * un2 + p2 * n1) - (U[1] * un1 + p1 * n1);
    DF[2] = (U2[2] * un2 + p2 * n2) - (U[2] * un1 + p1 * n2);
    DF[3] = (U2[3] * un2 + p2 * n3) - (U[3] * un1 + p1 * n3);
    DF[4] = (U2[4] + p2) * un2 - (U[4] + p1) * un1;
#pragma unroll
    for (int m = 5; m < NV; m++) DF[m] = U2[m] * un2 - U[m] * un1;
}

}  // namespace ocfd
