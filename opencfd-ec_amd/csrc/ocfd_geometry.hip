// Device geometry precompute: Comput_Goemetric_var (sub_geometry.f90:6-346), the dtime_mesh
// factor of check_mesh_quality_onemesh (sub_geometry.f90:361-515) and the wall distance of
// comput_wall_dist (sub_comput_dw.f90:43-100).  Inputs: node coordinates with their ghost layer
// already filled by the host (Update_coordinate_buffer_onemesh).  Same evaluation order as
// the reference; acos/exp come from the device math library.
#include "ocfd_dev.h"
#include "ocfd_launch.h"
#include "ocfd_phys.h"

namespace ocfd {

// face areas, unit normals and the per-face volume contributions Vi/Vj/Vk (kept in F[0][0..2])
__global__ void k_geo_faces(Blk b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx) return;
    const double *x = b.x, *y = b.y, *z = b.z;
    auto X = [&](int a, int bb, int c) { return x[IX(b, a, bb, c)]; };
    auto Y = [&](int a, int bb, int c) { return y[IX(b, a, bb, c)]; };
    auto Z = [&](int a, int bb, int c) { return z[IX(b, a, bb, c)]; };
    size_t s = IX(b, i, j, k);
    double t1x, t1y, t1z, t2x, t2y, t2z, s1x, s1y, s1z, ss, xa, ya, za;
    // i face
    t1x = X(i, j + 1, k) - X(i, j, k + 1); t1y = Y(i, j + 1, k) - Y(i, j, k + 1); t1z = Z(i, j + 1, k) - Z(i, j, k + 1);
    t2x = X(i, j + 1, k + 1) - X(i, j, k); t2y = Y(i, j + 1, k + 1) - Y(i, j, k); t2z = Z(i, j + 1, k + 1) - Z(i, j, k);
    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
    b.S[0][s] = ss * 0.5;
    if (ss >= 1.0e-20) { b.nv[0][s] = s1x / ss; b.nv[1][s] = s1y / ss; b.nv[2][s] = s1z / ss; }
    else { b.nv[0][s] = 1.0; b.nv[1][s] = 0.0; b.nv[2][s] = 0.0; }
    xa = (((X(i, j, k) + X(i, j + 1, k)) + X(i, j, k + 1)) + X(i, j + 1, k + 1)) * 0.25;
    ya = (((Y(i, j, k) + Y(i, j + 1, k)) + Y(i, j, k + 1)) + Y(i, j + 1, k + 1)) * 0.25;
    za = (((Z(i, j, k) + Z(i, j + 1, k)) + Z(i, j, k + 1)) + Z(i, j + 1, k + 1)) * 0.25;
    b.F[0][0][s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;
    // j face
    t1x = X(i + 1, j, k + 1) - X(i, j, k); t1y = Y(i + 1, j, k + 1) - Y(i, j, k); t1z = Z(i + 1, j, k + 1) - Z(i, j, k);
    t2x = X(i + 1, j, k) - X(i, j, k + 1); t2y = Y(i + 1, j, k) - Y(i, j, k + 1); t2z = Z(i + 1, j, k) - Z(i, j, k + 1);
    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
    b.S[1][s] = ss * 0.5;
    if (ss >= 1.0e-20) { b.nv[3][s] = s1x / ss; b.nv[4][s] = s1y / ss; b.nv[5][s] = s1z / ss; }
    else { b.nv[3][s] = 1.0; b.nv[4][s] = 0.0; b.nv[5][s] = 0.0; }
    xa = (((X(i, j, k) + X(i + 1, j, k)) + X(i, j, k + 1)) + X(i + 1, j, k + 1)) * 0.25;
    ya = (((Y(i, j, k) + Y(i + 1, j, k)) + Y(i, j, k + 1)) + Y(i + 1, j, k + 1)) * 0.25;
    za = (((Z(i, j, k) + Z(i + 1, j, k)) + Z(i, j, k + 1)) + Z(i + 1, j, k + 1)) * 0.25;
    b.F[0][1][s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;
    // k face
    t1x = X(i + 1, j + 1, k) - X(i, j, k); t1y = Y(i + 1, j + 1, k) - Y(i, j, k); t1z = Z(i + 1, j + 1, k) - Z(i, j, k);
    t2x = X(i, j + 1, k) - X(i + 1, j, k); t2y = Y(i, j + 1, k) - Y(i + 1, j, k); t2z = Z(i, j + 1, k) - Z(i + 1, j, k);
    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
    b.S[2][s] = ss * 0.5;
    if (ss >= 1.0e-20) { b.nv[6][s] = s1x / ss; b.nv[7][s] = s1y / ss; b.nv[8][s] = s1z / ss; }
    else { b.nv[6][s] = 1.0; b.nv[7][s] = 0.0; b.nv[8][s] = 0.0; }
    xa = (((X(i, j, k) + X(i + 1, j, k)) + X(i, j + 1, k)) + X(i + 1, j + 1, k)) * 0.25;
    ya = (((Y(i, j, k) + Y(i + 1, j, k)) + Y(i, j + 1, k)) + Y(i + 1, j + 1, k)) * 0.25;
    za = (((Z(i, j, k) + Z(i + 1, j, k)) + Z(i, j + 1, k)) + Z(i + 1, j + 1, k)) * 0.25;
    b.F[0][2][s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;
}

__global__ void k_geo_vol(Blk b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx - 1) return;
    const double *Vi = b.F[0][0], *Vj = b.F[0][1], *Vk = b.F[0][2];
    size_t s = IX(b, i, j, k);
    b.vol[s] = (((((Vi[IX(b, i + 1, j, k)] - Vi[s]) + Vj[IX(b, i, j + 1, k)]) - Vj[s]) + Vk[IX(b, i, j, k + 1)]) - Vk[s]) / 3.0;
}

__global__ void k_geo_centres(Blk b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i > b.nx) return;
    const double* A[3] = {b.x, b.y, b.z};
    double* C[3] = {b.xc, b.yc, b.zc};
    for (int q = 0; q < 3; q++) {
        const double* x = A[q];
        auto X = [&](int a, int bb, int c) { return x[IX(b, a, bb, c)]; };
        C[q][IX(b, i, j, k)] = (((((((X(i, j, k) + X(i, j + 1, k)) + X(i, j, k + 1)) + X(i, j + 1, k + 1)) + X(i + 1, j, k)) + X(i + 1, j + 1, k))
                                 + X(i + 1, j, k + 1)) + X(i + 1, j + 1, k + 1)) * 0.125;
    }
}

__device__ __forceinline__ void metric_store(double* const* mt, size_t s, double xi, double yi, double zi, double xj, double yj,
                                             double zj, double xk, double yk, double zk) {
    double Jac1 = ((((xi * yj * zk + yi * zj * xk) + zi * xj * yk) - xi * zj * yk) - yi * xj * zk) - zi * yj * xk;
    if (fabs(Jac1) < 1.0e-20) Jac1 = 1.0e-20;
    double Jac = 1.0 / Jac1;
    mt[0][s] = Jac * (yj * zk - zj * yk); mt[1][s] = Jac * (zj * xk - xj * zk); mt[2][s] = Jac * (xj * yk - yj * xk);
    mt[3][s] = Jac * (yk * zi - zk * yi); mt[4][s] = Jac * (zk * xi - xk * zi); mt[5][s] = Jac * (xk * yi - yk * xi);
    mt[6][s] = Jac * (yi * zj - zi * yj); mt[7][s] = Jac * (zi * xj - xi * zj); mt[8][s] = Jac * (xi * yj - yi * xj);
}

// metric set DIR+1 on (I-1/2,J,K) / (I,J-1/2,K) / (I,J,K-1/2) faces (sub_geometry.f90:110-310)
template <int DIR>
__global__ void k_geo_metric(Blk b) {
    constexpr int di = DIR == 0, dj = DIR == 1, dk = DIR == 2;
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx - 1 + di || j > b.ny - 1 + dj) return;
    int im = i - di, jm = j - dj, km = k - dk;
    const double* C[3] = {b.xc, b.yc, b.zc};
    constexpr int ai = DIR == 0 ? 0 : 1, aj = DIR == 0 ? 1 : 0;
    constexpr int bj = DIR == 2 ? 1 : 0, bk = DIR == 2 ? 0 : 1;
    int pos = DIR == 0 ? i : DIR == 1 ? j : k, nd = DIR == 0 ? b.nx : DIR == 1 ? b.ny : b.nz;
    int posA = DIR == 0 ? j : i, nA = DIR == 0 ? b.ny : b.nx;
    int posB = DIR == 2 ? j : k, nB = DIR == 2 ? b.ny : b.nz;
    bool cA = (pos == 1 || pos == nd) && (posA == 1 || posA == nA - 1);
    bool cB = (pos == 1 || pos == nd) && (posB == 1 || posB == nB - 1);
    double dn[3], ga[3], gb[3];
    for (int q = 0; q < 3; q++) {
        const double* c = C[q];
        dn[q] = c[IX(b, i, j, k)] - c[IX(b, im, jm, km)];
        double a1, a2, b1, b2;
        if (cA) { a1 = c[IX(b, i - ai, j - aj, k)]; a2 = c[IX(b, i + ai, j + aj, k)]; }
        else { a1 = 0.5 * (c[IX(b, i - ai, j - aj, k)] + c[IX(b, im - ai, jm - aj, km)]);
               a2 = 0.5 * (c[IX(b, i + ai, j + aj, k)] + c[IX(b, im + ai, jm + aj, km)]); }
        if (cB) { b1 = c[IX(b, i, j - bj, k - bk)]; b2 = c[IX(b, i, j + bj, k + bk)]; }
        else { b1 = 0.5 * (c[IX(b, i, j - bj, k - bk)] + c[IX(b, im, jm - bj, km - bk)]);
               b2 = 0.5 * (c[IX(b, i, j + bj, k + bk)] + c[IX(b, im, jm + bj, km + bk)]); }
        ga[q] = 0.5 * (a2 - a1); gb[q] = 0.5 * (b2 - b1);
    }
    size_t s = IX(b, i, j, k);
    if (DIR == 0) metric_store(b.met[1], s, dn[0], dn[1], dn[2], ga[0], ga[1], ga[2], gb[0], gb[1], gb[2]);
    else if (DIR == 1) metric_store(b.met[2], s, ga[0], ga[1], ga[2], dn[0], dn[1], dn[2], gb[0], gb[1], gb[2]);
    else metric_store(b.met[3], s, ga[0], ga[1], ga[2], gb[0], gb[1], gb[2], dn[0], dn[1], dn[2]);
}

// metric set 0 at cell centres (sub_geometry.f90:312-341)
__global__ void k_geo_metric0(Blk b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx - 1 || b.met[0][0] == nullptr) return;
    const double *xc = b.xc, *yc = b.yc, *zc = b.zc;
    auto D = [&](const double* c, int di, int dj, int dk) { return c[IX(b, i + di, j + dj, k + dk)] - c[IX(b, i - di, j - dj, k - dk)]; };
    metric_store(b.met[0], IX(b, i, j, k), D(xc, 1, 0, 0), D(yc, 1, 0, 0), D(zc, 1, 0, 0), D(xc, 0, 1, 0), D(yc, 0, 1, 0), D(zc, 0, 1, 0),
                 D(xc, 0, 0, 1), D(yc, 0, 0, 1), D(zc, 0, 0, 1));
}

// dtime_mesh (sub_geometry.f90:410-444)
__global__ void k_mesh_quality(Blk b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx - 1) return;
    const double *xc = b.xc, *yc = b.yc, *zc = b.zc;
    auto XC = [&](int a, int bb, int c) { return xc[IX(b, a, bb, c)]; };
    auto YC = [&](int a, int bb, int c) { return yc[IX(b, a, bb, c)]; };
    auto ZC = [&](int a, int bb, int c) { return zc[IX(b, a, bb, c)]; };
    auto L2 = [&](int a, int bb, int c) { return (sq(XC(i, j, k) - XC(a, bb, c)) + sq(YC(i, j, k) - YC(a, bb, c))) + sq(ZC(i, j, k) - ZC(a, bb, c)); };
    double fl1 = sqrt(L2(i - 1, j, k)) / sqrt(L2(i + 1, j, k));
    double fl2 = sqrt(L2(i, j - 1, k)) / sqrt(L2(i, j + 1, k));
    double fl3 = sqrt(L2(i, j, k - 1)) / sqrt(L2(i, j, k + 1));
    if (fl1 < 1.0) fl1 = 1.0 / fl1;
    if (fl2 < 1.0) fl2 = 1.0 / fl2;
    if (fl3 < 1.0) fl3 = 1.0 / fl3;
    double fl = fl1; if (fl2 > fl) fl = fl2; if (fl3 > fl) fl = fl3;
    double ft[3];
    for (int dd = 0; dd < 3; dd++) {
        int di = dd == 0, dj = dd == 1, dk = dd == 2;
        double x1 = XC(i, j, k) - XC(i - di, j - dj, k - dk), x2 = XC(i + di, j + dj, k + dk) - XC(i, j, k);
        double y1 = YC(i, j, k) - YC(i - di, j - dj, k - dk), y2 = YC(i + di, j + dj, k + dk) - YC(i, j, k);
        double z1 = ZC(i, j, k) - ZC(i - di, j - dj, k - dk), z2 = ZC(i + di, j + dj, k + dk) - ZC(i, j, k);
        ft[dd] = ((x1 * x2 + y1 * y2) + z1 * z2) / sqrt(((x1 * x1 + y1 * y1) + z1 * z1) * ((x2 * x2 + y2 * y2) + z2 * z2));
    }
    double mn = ft[0]; if (ft[1] < mn) mn = ft[1]; if (ft[2] < mn) mn = ft[2];
    double f = 1.0 * mn; if (f > 1.0) f = 1.0;
    f = acos(f);
    double Af = 0.9 * exp(-4.0 * sq(fl - 1.0)) + 0.1;
    double a = (4.0 / 3.1415926535) * f;
    double Aa = 0.9 * exp(-(a * a)) + 0.1;
    b.dtm[IX(b, i, j, k)] = Af < Aa ? Af : Aa;
}

// brute-force minimum distance from every node to the wall nodes; the minimum of the distances
// is the square root of the minimum squared distance (sqrt is monotone, correctly rounded).
// Wall nodes are staged through LDS in tiles.
__global__ void k_wall_node_dist(Blk b, const double* __restrict__ W, long npw, double* dis) {
    __shared__ double sw[3 * 256];
    long nnode = (long)b.nx * b.ny * b.nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i = 1, j = 1, k = 1;
    double px = 0, py = 0, pz = 0;
    if (t < nnode) {
        i = (int)(t % b.nx) + 1; j = (int)((t / b.nx) % b.ny) + 1; k = (int)(t / ((long)b.nx * b.ny)) + 1;
        size_t s = IX(b, i, j, k);
        px = b.x[s]; py = b.y[s]; pz = b.z[s];
    }
    double best = 1.0e16;
    for (long base = 0; base < npw; base += 256) {
        long q = base + threadIdx.x;
        __syncthreads();
        if (q < npw) { sw[3 * threadIdx.x] = W[3 * q]; sw[3 * threadIdx.x + 1] = W[3 * q + 1]; sw[3 * threadIdx.x + 2] = W[3 * q + 2]; }
        __syncthreads();
        int nq = (int)((npw - base) < 256 ? (npw - base) : 256);
        for (int r = 0; r < nq; r++) {
            double dx = px - sw[3 * r], dy = py - sw[3 * r + 1], dz = pz - sw[3 * r + 2];
            double d2 = (dx * dx + dy * dy) + dz * dz;
            best = d2 < best ? d2 : best;
        }
    }
    if (t < nnode) dis[IX(b, i, j, k)] = npw > 0 ? sqrt(best) : 1.0e8;
}

__global__ void k_wall_cell_dist(Blk b, const double* dis) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.nx - 1) return;
    auto D = [&](int a, int bb, int c) { return dis[IX(b, a, bb, c)]; };
    b.dw[IX(b, i, j, k)] = (((((((D(i, j, k) + D(i, j + 1, k)) + D(i, j, k + 1)) + D(i, j + 1, k + 1)) + D(i + 1, j, k)) + D(i + 1, j + 1, k))
                             + D(i + 1, j, k + 1)) + D(i + 1, j + 1, k + 1)) * 0.125;
}

static inline dim3 g3(int n0, int n1, int n2, int bs) { return dim3((unsigned)((n0 + bs - 1) / bs), (unsigned)n1, (unsigned)n2); }

void launch_geometry(const Blk& b, const Sub* subs, int nsub, hipStream_t st) {
    (void)subs; (void)nsub;
    hipLaunchKernelGGL(k_geo_faces, g3(b.nx, b.ny, b.nz, 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_geo_vol, g3(b.nx - 1, b.ny - 1, b.nz - 1, 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_geo_centres, g3(b.nx + 1, b.ny + 1, b.nz + 1, 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_geo_metric<0>, g3(b.nx, b.ny - 1, b.nz - 1, 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_geo_metric<1>, g3(b.nx - 1, b.ny, b.nz - 1, 64), dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_geo_metric<2>, g3(b.nx - 1, b.ny - 1, b.nz, 64), dim3(64), 0, st, b);
    if (b.met[0][0] != nullptr) hipLaunchKernelGGL(k_geo_metric0, g3(b.nx - 1, b.ny - 1, b.nz - 1, 64), dim3(64), 0, st, b);
}

// FP32 copies of the geometry the mixed-precision kernels read, once per block: face normals,
// areas and metrics converted from the FP64 arrays, and per direction the node difference t1 of
// the inviscid flux's face frame (sub_Residual.f90:199-205, 366-372, 533-539), the FP64 difference
// converted as the FP64 kernels form it.  One thread per point of the padded box; t1 where both
// of its nodes lie in the box.
__global__ void k_geo_f32(Blk b) {
    const long n = (long)b.NI * b.NJ * b.NK;
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int q = 0; q < 9; q++) b.nvf[q][s] = (float)b.nv[q][s];
    for (int d = 0; d < 3; d++) b.Sf[d][s] = (float)b.S[d][s];
    for (int m = 0; m < 4; m++)
        if (b.metf[m][0] != nullptr)
            for (int q = 0; q < 9; q++) b.metf[m][q][s] = (float)b.met[m][q][s];
    const int i = (int)(s % b.NI) - 3, j = (int)((s / b.NI) % b.NJ) - 3, k = (int)(s / ((long)b.NI * b.NJ)) - 3;
    const bool ip = i + 1 <= b.nx + 3, jp = j + 1 <= b.ny + 3, kp = k + 1 <= b.nz + 3;
    const double* X[3] = {b.x, b.y, b.z};
    for (int c = 0; c < 3; c++) {
        b.t1f[0][c][s] = jp && kp ? (float)(X[c][IX(b, i, j + 1, k)] - X[c][IX(b, i, j, k + 1)]) : 0.0f;
        b.t1f[1][c][s] = ip && kp ? (float)(X[c][IX(b, i + 1, j, k + 1)] - X[c][s]) : 0.0f;
        b.t1f[2][c][s] = ip && jp ? (float)(X[c][IX(b, i + 1, j + 1, k)] - X[c][s]) : 0.0f;
    }
}

void launch_geometry_f32(const Blk& b, hipStream_t st) {
    const long n = (long)b.NI * b.NJ * b.NK;
    hipLaunchKernelGGL(k_geo_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b);
}

void launch_mesh_quality(const Blk& b, hipStream_t st) {
    hipLaunchKernelGGL(k_mesh_quality, g3(b.nx - 1, b.ny - 1, b.nz - 1, 64), dim3(64), 0, st, b);
}

void launch_wall_distance(const Blk& b, const double* wall_xyz, long npw, hipStream_t st) {
    long nnode = (long)b.nx * b.ny * b.nz;
    double* dis = b.F[1][0];      // face-flux storage is free during setup
    hipLaunchKernelGGL(k_wall_node_dist, dim3((unsigned)((nnode + 255) / 256)), dim3(256), 0, st, b, wall_xyz, npw, dis);
    hipLaunchKernelGGL(k_wall_cell_dist, g3(b.nx - 1, b.ny - 1, b.nz - 1, 64), dim3(64), 0, st, b, dis);
}

}  // namespace ocfd
