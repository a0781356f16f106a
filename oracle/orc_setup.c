/* Oracle setup: parameters, Gridgen subface conversion, coordinate ghosts, geometry, mesh
 * quality, wall distance, boundary indicators, freestream initialisation.
 * TEST INFRASTRUCTURE ONLY (see orc_internal.h). */
#include "orc_internal.h"
#include "ocfd_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double *zalloc(size_t n) {
    double *p = (double *)calloc(n, sizeof(double));
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static int *ialloc(size_t n) {
    int *p = (int *)calloc(n, sizeof(int));
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* set_default_parameter + read_parameter_ec + set_const_para (sub_read_parameter.f90:30-106,
 * 371-429).  rp/ip carry the namelist values in the order documented in ocfd_oracle.h. */
orc_ctx *orc_create(const double *rp, const int *ip) {
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->Ma = rp[0]; c->Re = rp[1]; c->gamma = rp[2]; c->AoA = rp[3]; c->AoS = rp[4];
    c->P_outlet = rp[5]; c->PrL = rp[6]; c->PrT = rp[7]; c->T_inf = rp[8]; c->Twall = rp[9];
    c->CFL = rp[10]; c->dtmax = rp[11]; c->dtmin = rp[12]; c->dt_global = rp[13]; c->w_LU = rp[14];
    c->MUT_MAX = rp[15]; c->Kt_inf = rp[16]; c->Wt_inf = rp[17]; c->Ldmin = rp[18]; c->Ldmax = rp[19];
    c->Lpmin = rp[20]; c->Lpmax = rp[21]; c->Lumax = rp[22]; c->LSAmax = rp[23]; c->Ref_S = rp[24];
    c->Ref_L = rp[25]; c->Centroid[0] = rp[26]; c->Centroid[1] = rp[27]; c->Centroid[2] = rp[28];
    c->Periodic_dX = rp[29]; c->Periodic_dY = rp[30]; c->Periodic_dZ = rp[31];
    c->If_viscous = ip[0]; c->turb = ip[1]; c->Scheme = ip[2]; c->Flux = ip[3]; c->Recon = ip[4];
    c->Bound_Scheme = ip[5]; c->Iflag_local_dt = ip[6]; c->If_dtime_mesh = ip[7];
    c->IF_Scheme_Positivity = ip[8]; c->IFLAG_LIMIT_FLOW = ip[9]; c->Cood_Y_UP = ip[10];
    c->Time_Method = ip[11];
    /* set_const_para */
    c->Twall = c->Twall / c->T_inf;
    c->Lpmin = c->Lpmin / (c->gamma * c->Ma * c->Ma);
    c->Lpmax = c->Lpmax / (c->gamma * c->Ma * c->Ma);
    if (c->Bound_Scheme == -1) c->Bound_Scheme = c->Scheme;
    if (c->If_viscous == 0) c->turb = 0;
    c->NVAR = (c->turb == 2 || c->turb == 21) ? 6 : (c->turb == 3 ? 7 : 5);
    c->AoA = c->AoA * PI_EC / 180.0;
    c->AoS = c->AoS * PI_EC / 180.0;
    if (c->Cood_Y_UP == 1) { c->A_alfa = c->AoA; c->A_beta = c->AoS; }
    else { c->A_alfa = c->AoS; c->A_beta = c->AoA; }
    c->Cv = 1.0 / (c->gamma * (c->gamma - 1.0) * c->Ma * c->Ma);
    c->Cp = c->Cv * c->gamma;
    return c;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        double **all[] = {&b->x, &b->y, &b->z, &b->xc, &b->yc, &b->zc, &b->vol, &b->Si, &b->Sj, &b->Sk,
                          &b->dw, &b->dtime_mesh, &b->dt, &b->mu, &b->mut, &b->d, &b->uu, &b->v, &b->w,
                          &b->T, &b->p, &b->cc, &b->Lci, &b->Lcj, &b->Lck, &b->Lvi, &b->Lvj, &b->Lvk};
        for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); q++) free(*all[q]);
        for (int q = 0; q < 9; q++) { free(b->nv[q]); for (int s = 0; s < 4; s++) free(b->met[s][q]); }
        for (int q = 0; q < ORC_MAXV; q++) { free(b->U[q]); free(b->Un[q]); free(b->Res[q]); free(b->DU[q]); free(b->Flux[q]); }
        for (int q = 0; q < 2; q++) { free(b->BcI[q]); free(b->BcJ[q]); free(b->BcK[q]); }
        for (int f = 0; f < 6; f++) for (int q = 0; q < 3; q++) free(b->Surf[f][q]);
        free(b->bc);
    }
    free(c->blk);
    free(c);
}

/* allocate_mem_Blocks (sub_init.f90:444-541): storage and its initial values. */
int orc_add_block(orc_ctx *c, int nx, int ny, int nz, const double *x, const double *y, const double *z) {
    if (c->nblk == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 4;
        c->blk = (orc_block *)realloc(c->blk, (size_t)c->cap * sizeof(orc_block));
    }
    orc_block *b = &c->blk[c->nblk];
    memset(b, 0, sizeof(*b));
    b->nx = nx; b->ny = ny; b->nz = nz;
    b->NI = nx + 7; b->NJ = ny + 7; b->NK = nz + 7;
    size_t n = (size_t)b->NI * b->NJ * b->NK;
    b->n = n;
    double **arr[] = {&b->x, &b->y, &b->z, &b->xc, &b->yc, &b->zc, &b->vol, &b->Si, &b->Sj, &b->Sk,
                      &b->dw, &b->dtime_mesh, &b->dt, &b->mu, &b->mut, &b->d, &b->uu, &b->v, &b->w,
                      &b->T, &b->p, &b->cc, &b->Lci, &b->Lcj, &b->Lck, &b->Lvi, &b->Lvj, &b->Lvk};
    for (size_t q = 0; q < sizeof(arr) / sizeof(arr[0]); q++) *arr[q] = zalloc(n);
    for (int q = 0; q < 9; q++) { b->nv[q] = zalloc(n); for (int s = 0; s < 4; s++) b->met[s][q] = zalloc(n); }
    for (int q = 0; q < c->NVAR; q++) {
        b->U[q] = zalloc(n); b->Un[q] = zalloc(n); b->Res[q] = zalloc(n); b->DU[q] = zalloc(n);
        b->Flux[q] = zalloc(n);
    }
    size_t n2 = (size_t)(b->NI > b->NJ ? b->NI : b->NJ) * (size_t)(b->NJ > b->NK ? b->NJ : b->NK) * 2;
    for (int q = 0; q < 2; q++) { b->BcI[q] = ialloc(n2); b->BcJ[q] = ialloc(n2); b->BcK[q] = ialloc(n2); }
    for (int f = 0; f < 6; f++) for (int q = 0; q < 3; q++) b->Surf[f][q] = zalloc(n2);
    /* node coordinates (read_main_mesh, sub_IO.f90:48-56) */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                size_t s = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
                b->x[IX(b, i, j, k)] = x[s]; b->y[IX(b, i, j, k)] = y[s]; b->z[IX(b, i, j, k)] = z[s];
            }
    /* initial values (sub_init.f90:496,507-526) */
    for (size_t s = 0; s < n; s++) {
        b->dtime_mesh[s] = 1.0;
        b->mu[s] = 1.0 / c->Re;
        b->U[0][s] = 1.0; b->U[4][s] = 1.0;
        if (c->NVAR == 6) b->U[5][s] = 1.0;
        if (c->NVAR == 7) { b->U[5][s] = 0.0; b->U[6][s] = 1.0; }
    }
    c->Num_Cell += (long)(nx - 1) * (ny - 1) * (nz - 1);
    c->nblk++;
    return c->nblk;
}

/* Convert_bc (sub_convert_inp.f90:131-247) for one Gridgen record of block `blk` (1-based). */
int orc_add_subface(orc_ctx *c, int blk, const int *rng, int bc, const int *link) {
    orc_block *b = &c->blk[blk - 1];
    b->bc = (orc_bcmsg *)realloc(b->bc, (size_t)(b->nsub + 1) * sizeof(orc_bcmsg));
    orc_bcmsg *B = &b->bc[b->nsub];
    memset(B, 0, sizeof(*B));
    B->bc = bc;
    B->f_no = b->nsub + 1;
    int kb[3], ke[3], kb1[3] = {0, 0, 0}, ke1[3] = {0, 0, 0}, s[3], p[3], Lp[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++) { kb[k] = rng[2 * k]; ke[k] = rng[2 * k + 1]; }
    if (bc < 0) { for (int k = 0; k < 3; k++) { kb1[k] = link[2 * k]; ke1[k] = link[2 * k + 1]; } B->nb1 = link[6]; }
    for (int k = 0; k < 3; k++) s[k] = (kb[k] == ke[k]) ? 0 : (kb[k] > 0 ? 1 : -1);
    B->ib = abs(kb[0]) < abs(ke[0]) ? abs(kb[0]) : abs(ke[0]); B->ie = abs(kb[0]) > abs(ke[0]) ? abs(kb[0]) : abs(ke[0]);
    B->jb = abs(kb[1]) < abs(ke[1]) ? abs(kb[1]) : abs(ke[1]); B->je = abs(kb[1]) > abs(ke[1]) ? abs(kb[1]) : abs(ke[1]);
    B->kb = abs(kb[2]) < abs(ke[2]) ? abs(kb[2]) : abs(ke[2]); B->ke = abs(kb[2]) > abs(ke[2]) ? abs(kb[2]) : abs(ke[2]);
    if (s[0] == 0) B->face = (B->ib == 1) ? 1 : 4;
    else if (s[1] == 0) B->face = (B->jb == 1) ? 2 : 5;
    else B->face = (B->kb == 1) ? 3 : 6;
    if (bc < 0) {
        for (int k = 0; k < 3; k++) p[k] = (kb1[k] == ke1[k]) ? 0 : (kb1[k] > 0 ? 1 : -1);
        B->ib1 = abs(kb1[0]) < abs(ke1[0]) ? abs(kb1[0]) : abs(ke1[0]); B->ie1 = abs(kb1[0]) > abs(ke1[0]) ? abs(kb1[0]) : abs(ke1[0]);
        B->jb1 = abs(kb1[1]) < abs(ke1[1]) ? abs(kb1[1]) : abs(ke1[1]); B->je1 = abs(kb1[1]) > abs(ke1[1]) ? abs(kb1[1]) : abs(ke1[1]);
        B->kb1 = abs(kb1[2]) < abs(ke1[2]) ? abs(kb1[2]) : abs(ke1[2]); B->ke1 = abs(kb1[2]) > abs(ke1[2]) ? abs(kb1[2]) : abs(ke1[2]);
        if (p[0] == 0) B->face1 = (B->ib1 == 1) ? 1 : 4;
        else if (p[1] == 0) B->face1 = (B->jb1 == 1) ? 2 : 5;
        else B->face1 = (B->kb1 == 1) ? 3 : 6;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                if (s[k] == p[j]) Lp[k] = j + 1;
        for (int k = 0; k < 3; k++) {
            if (s[k] != 0) {
                int k1 = Lp[k] - 1;
                if ((ke[k] - kb[k]) * (ke1[k1] - kb1[k1]) < 0) Lp[k] = -Lp[k];
            } else {
                if ((B->face - 1) / 3 == (B->face1 - 1) / 3) Lp[k] = -Lp[k];
            }
        }
        B->L1 = Lp[0]; B->L2 = Lp[1]; B->L3 = Lp[2];
    }
    b->nsub++;
    return b->nsub;
}

/* f_no1 search of convert_inp_inc (sub_convert_inp.f90:77-101). */
static int link_subfaces(orc_ctx *c) {
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc >= 0) continue;
            B->f_no1 = 0;
            orc_block *b1 = &c->blk[B->nb1 - 1];
            for (int q1 = 0; q1 < b1->nsub; q1++) {
                orc_bcmsg *C1 = &b1->bc[q1];
                if (B->ib1 == C1->ib && B->ie1 == C1->ie && B->jb1 == C1->jb && B->je1 == C1->je &&
                    B->kb1 == C1->kb && B->ke1 == C1->ke && C1->nb1 == m + 1) { B->f_no1 = q1 + 1; break; }
            }
            if (B->f_no1 == 0) { fprintf(stderr, "oracle: linked subface not found (block %d sub %d)\n", m + 1, q + 1); return -1; }
        }
    }
    return 0;
}

int orc_get_inc(orc_ctx *c, int blk, int *out21) {
    orc_block *b = &c->blk[blk - 1];
    for (int q = 0; q < b->nsub; q++) {
        orc_bcmsg *B = &b->bc[q];
        int v[21] = {B->ib, B->ie, B->jb, B->je, B->kb, B->ke, B->bc, B->face, B->f_no, B->ib1, B->ie1, B->jb1,
                     B->je1, B->kb1, B->ke1, B->nb1, B->face1, B->f_no1, B->L1, B->L2, B->L3};
        memcpy(out21 + 21 * q, v, sizeof(v));
    }
    return b->nsub;
}

/* -------------------------------------------------------------------------------------------
 * Coordinate ghost nodes: Coordinate_send_mpi (sub_update_buffer_mpi.f90:250-352),
 * Coordinate_Periodic (:609-677), coordinate_boundary_and_corner (:413-468), get_xyz_conner (:476-512).
 */
static void xyz_corner(orc_block *b, double *x) {
    int nx = b->nx, ny = b->ny, nz = b->nz;
#define X(i, j, k) x[IX(b, i, j, k)]
    for (int k = 1; k <= nz; k++) {
        X(0, 0, k) = (X(1, 0, k) + X(0, 1, k)) * 0.5;
        X(nx + 1, 0, k) = (X(nx, 0, k) + X(nx + 1, 1, k)) * 0.5;
        X(0, ny + 1, k) = (X(1, ny + 1, k) + X(0, ny, k)) * 0.5;
        X(nx + 1, ny + 1, k) = (X(nx, ny + 1, k) + X(nx + 1, ny, k)) * 0.5;
    }
    for (int j = 1; j <= ny; j++) {
        X(0, j, 0) = (X(1, j, 0) + X(0, j, 1)) * 0.5;
        X(nx + 1, j, 0) = (X(nx, j, 0) + X(nx + 1, j, 1)) * 0.5;
        X(0, j, nz + 1) = (X(1, j, nz + 1) + X(0, j, nz)) * 0.5;
        X(nx + 1, j, nz + 1) = (X(nx, j, nz + 1) + X(nx + 1, j, nz)) * 0.5;
    }
    for (int i = 1; i <= nx; i++) {
        X(i, 0, 0) = (X(i, 1, 0) + X(i, 0, 1)) * 0.5;
        X(i, ny + 1, 0) = (X(i, ny, 0) + X(i, ny + 1, 1)) * 0.5;
        X(i, 0, nz + 1) = (X(i, 1, nz + 1) + X(i, 0, nz)) * 0.5;
        X(i, ny + 1, nz + 1) = (X(i, ny, nz + 1) + X(i, ny + 1, nz)) * 0.5;
    }
    X(0, 0, 0) = (2.0 * ((X(1, 0, 0) + X(0, 1, 0)) + X(0, 0, 1)) - ((X(1, 1, 0) + X(1, 0, 1)) + X(0, 1, 1))) / 3.0;
    X(nx + 1, 0, 0) = (2.0 * ((X(nx, 0, 0) + X(nx + 1, 1, 0)) + X(nx + 1, 0, 1)) - ((X(nx, 1, 0) + X(nx, 0, 1)) + X(nx + 1, 1, 1))) / 3.0;
    X(0, ny + 1, 0) = (2.0 * ((X(1, ny + 1, 0) + X(0, ny, 0)) + X(0, ny + 1, 1)) - ((X(1, ny, 0) + X(1, ny + 1, 1)) + X(0, ny, 1))) / 3.0;
    X(nx + 1, ny + 1, 0) = (2.0 * ((X(nx, ny + 1, 0) + X(nx + 1, ny, 0)) + X(nx + 1, ny + 1, 1)) - ((X(nx, ny, 0) + X(nx, ny + 1, 1)) + X(nx + 1, ny, 1))) / 3.0;
    X(0, 0, nz + 1) = (2.0 * ((X(1, 0, nz + 1) + X(0, 1, nz + 1)) + X(0, 0, nz)) - ((X(1, 1, nz + 1) + X(1, 0, nz)) + X(0, 1, nz))) / 3.0;
    X(nx + 1, 0, nz + 1) = (2.0 * ((X(nx, 0, nz + 1) + X(nx + 1, 1, nz + 1)) + X(nx + 1, 0, nz)) - ((X(nx, 1, nz + 1) + X(nx, 0, nz)) + X(nx + 1, 1, nz))) / 3.0;
    X(0, ny + 1, nz + 1) = (2.0 * ((X(1, ny + 1, nz + 1) + X(0, ny, nz + 1)) + X(0, ny + 1, nz)) - ((X(1, ny, nz + 1) + X(1, ny + 1, nz)) + X(0, ny, nz))) / 3.0;
    X(nx + 1, ny + 1, nz + 1) = (2.0 * ((X(nx, ny + 1, nz + 1) + X(nx + 1, ny, nz + 1)) + X(nx + 1, ny + 1, nz)) - ((X(nx, ny, nz + 1) + X(nx, ny + 1, nz)) + X(nx + 1, ny, nz))) / 3.0;
#undef X
}

static int isign1(int v) { return v >= 0 ? 1 : -1; }

void orc_coordinate_ghosts(orc_ctx *c) {
    /* Coordinate_send_mpi: one ghost node layer from the neighbour's second node layer */
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc >= 0) continue;
            int kb[3] = {B->ib, B->jb, B->kb}, ke[3] = {B->ie, B->je, B->ke};
            int kd = (B->face - 1) % 3;
            if (B->face <= 3) kb[kd] += 1; else kb[kd] -= 1;
            ke[kd] = kb[kd];
            int kb1[3] = {B->ib1, B->jb1, B->kb1}, ke1[3] = {B->ie1, B->je1, B->ke1};
            int kd1 = (B->face1 - 1) % 3;
            if (B->face1 <= 3) kb1[kd1] -= 1; else kb1[kd1] += 1;
            ke1[kd1] = kb1[kd1];
            int L[3] = {abs(B->L1), abs(B->L2), abs(B->L3)}, P[3] = {isign1(B->L1), isign1(B->L2), isign1(B->L3)}, ks[3];
            for (int k = 0; k < 3; k++) ks[k] = P[k] > 0 ? kb[k] : ke[k];
            orc_block *b1 = &c->blk[B->nb1 - 1];
            for (int k = kb1[2]; k <= ke1[2]; k++)
                for (int j = kb1[1]; j <= ke1[1]; j++)
                    for (int i = kb1[0]; i <= ke1[0]; i++) {
                        int ka[3] = {i - kb1[0], j - kb1[1], k - kb1[2]};
                        int i1 = ks[0] + ka[L[0] - 1] * P[0], j1 = ks[1] + ka[L[1] - 1] * P[1], k1 = ks[2] + ka[L[2] - 1] * P[2];
                        b1->x[IX(b1, i, j, k)] = b->x[IX(b, i1, j1, k1)];
                        b1->y[IX(b1, i, j, k)] = b->y[IX(b, i1, j1, k1)];
                        b1->z[IX(b1, i, j, k)] = b->z[IX(b, i1, j1, k1)];
                    }
        }
    }
    /* Coordinate_Periodic (non-turbomachinery branch) */
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc != -2 && B->bc != -3) continue;
            double Xp, Yp, Zp;
            if (B->bc == -2) { Xp = -c->Periodic_dX; Yp = -c->Periodic_dY; Zp = -c->Periodic_dZ; }
            else { Xp = c->Periodic_dX; Yp = c->Periodic_dY; Zp = c->Periodic_dZ; }
            int kb[3] = {B->ib, B->jb, B->kb}, ke[3] = {B->ie, B->je, B->ke};
            int kd = (B->face - 1) % 3;
            if (B->face <= 3) kb[kd] -= 1; else kb[kd] += 1;
            ke[kd] = kb[kd];
            for (int k = kb[2]; k <= ke[2]; k++)
                for (int j = kb[1]; j <= ke[1]; j++)
                    for (int i = kb[0]; i <= ke[0]; i++) {
                        b->x[IX(b, i, j, k)] = b->x[IX(b, i, j, k)] + Xp;
                        b->y[IX(b, i, j, k)] = b->y[IX(b, i, j, k)] + Yp;
                        b->z[IX(b, i, j, k)] = b->z[IX(b, i, j, k)] + Zp;
                    }
        }
    }
    /* coordinate_boundary_and_corner: x0 = 2 x1 - x2 on physical faces */
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc < 0) continue;
            int i1 = B->ib, i2 = B->ie, j1 = B->jb, j2 = B->je, k1 = B->kb, k2 = B->ke;
            double *A[3] = {b->x, b->y, b->z};
            for (int a = 0; a < 3; a++) {
                double *x = A[a];
                for (int k = k1; k <= k2; k++)
                    for (int j = j1; j <= j2; j++)
                        for (int i = i1; i <= i2; i++) {
                            switch (B->face) {
                            case 1: if (i == i1) x[IX(b, i1 - 1, j, k)] = 2.0 * x[IX(b, i1, j, k)] - x[IX(b, i1 + 1, j, k)]; break;
                            case 2: if (j == j1) x[IX(b, i, j1 - 1, k)] = 2.0 * x[IX(b, i, j1, k)] - x[IX(b, i, j1 + 1, k)]; break;
                            case 3: if (k == k1) x[IX(b, i, j, k1 - 1)] = 2.0 * x[IX(b, i, j, k1)] - x[IX(b, i, j, k1 + 1)]; break;
                            case 4: if (i == i2) x[IX(b, i2 + 1, j, k)] = 2.0 * x[IX(b, i2, j, k)] - x[IX(b, i2 - 1, j, k)]; break;
                            case 5: if (j == j2) x[IX(b, i, j2 + 1, k)] = 2.0 * x[IX(b, i, j2, k)] - x[IX(b, i, j2 - 1, k)]; break;
                            case 6: if (k == k2) x[IX(b, i, j, k2 + 1)] = 2.0 * x[IX(b, i, j, k2)] - x[IX(b, i, j, k2 - 1)]; break;
                            }
                        }
            }
        }
    }
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        xyz_corner(b, b->x); xyz_corner(b, b->y); xyz_corner(b, b->z);
    }
}

/* -------------------------------------------------------------------------------------------
 * Comput_Goemetric_var (sub_geometry.f90:6-346)
 */
static void metric_face(orc_block *b, double *const *mt, int i, int j, int k,
                        double xi, double yi, double zi, double xj, double yj, double zj,
                        double xk, double yk, double zk) {
    double Jac1 = ((((xi * yj * zk + yi * zj * xk) + zi * xj * yk) - xi * zj * yk) - yi * xj * zk) - zi * yj * xk;
    if (fabs(Jac1) < LIM_ZERO) Jac1 = LIM_ZERO;
    double Jac = 1.0 / Jac1;
    size_t s = IX(b, i, j, k);
    mt[0][s] = Jac * (yj * zk - zj * yk);
    mt[1][s] = Jac * (zj * xk - xj * zk);
    mt[2][s] = Jac * (xj * yk - yj * xk);
    mt[3][s] = Jac * (yk * zi - zk * yi);
    mt[4][s] = Jac * (zk * xi - xk * zi);
    mt[5][s] = Jac * (xk * yi - yk * xi);
    mt[6][s] = Jac * (yi * zj - zi * yj);
    mt[7][s] = Jac * (zi * xj - xi * zj);
    mt[8][s] = Jac * (xi * yj - yi * xj);
}

void orc_geometry(orc_ctx *c) {
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        int nx = b->nx, ny = b->ny, nz = b->nz;
        double *x = b->x, *y = b->y, *z = b->z;
        double *Vi = zalloc(b->n), *Vj = zalloc(b->n), *Vk = zalloc(b->n);
#define XX(i, j, k) x[IX(b, i, j, k)]
#define YY(i, j, k) y[IX(b, i, j, k)]
#define ZZ(i, j, k) z[IX(b, i, j, k)]
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t s = IX(b, i, j, k);
                    double t1x, t1y, t1z, t2x, t2y, t2z, s1x, s1y, s1z, ss, xa, ya, za;
                    t1x = XX(i, j + 1, k) - XX(i, j, k + 1); t1y = YY(i, j + 1, k) - YY(i, j, k + 1); t1z = ZZ(i, j + 1, k) - ZZ(i, j, k + 1);
                    t2x = XX(i, j + 1, k + 1) - XX(i, j, k); t2y = YY(i, j + 1, k + 1) - YY(i, j, k); t2z = ZZ(i, j + 1, k + 1) - ZZ(i, j, k);
                    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
                    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
                    b->Si[s] = ss * 0.5;
                    if (ss >= LIM_ZERO) { b->nv[0][s] = s1x / ss; b->nv[1][s] = s1y / ss; b->nv[2][s] = s1z / ss; }
                    else { b->nv[0][s] = 1.0; b->nv[1][s] = 0.0; b->nv[2][s] = 0.0; }
                    xa = (((XX(i, j, k) + XX(i, j + 1, k)) + XX(i, j, k + 1)) + XX(i, j + 1, k + 1)) * 0.25;
                    ya = (((YY(i, j, k) + YY(i, j + 1, k)) + YY(i, j, k + 1)) + YY(i, j + 1, k + 1)) * 0.25;
                    za = (((ZZ(i, j, k) + ZZ(i, j + 1, k)) + ZZ(i, j, k + 1)) + ZZ(i, j + 1, k + 1)) * 0.25;
                    Vi[s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;

                    t1x = XX(i + 1, j, k + 1) - XX(i, j, k); t1y = YY(i + 1, j, k + 1) - YY(i, j, k); t1z = ZZ(i + 1, j, k + 1) - ZZ(i, j, k);
                    t2x = XX(i + 1, j, k) - XX(i, j, k + 1); t2y = YY(i + 1, j, k) - YY(i, j, k + 1); t2z = ZZ(i + 1, j, k) - ZZ(i, j, k + 1);
                    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
                    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
                    b->Sj[s] = ss * 0.5;
                    if (ss >= LIM_ZERO) { b->nv[3][s] = s1x / ss; b->nv[4][s] = s1y / ss; b->nv[5][s] = s1z / ss; }
                    else { b->nv[3][s] = 1.0; b->nv[4][s] = 0.0; b->nv[5][s] = 0.0; }
                    xa = (((XX(i, j, k) + XX(i + 1, j, k)) + XX(i, j, k + 1)) + XX(i + 1, j, k + 1)) * 0.25;
                    ya = (((YY(i, j, k) + YY(i + 1, j, k)) + YY(i, j, k + 1)) + YY(i + 1, j, k + 1)) * 0.25;
                    za = (((ZZ(i, j, k) + ZZ(i + 1, j, k)) + ZZ(i, j, k + 1)) + ZZ(i + 1, j, k + 1)) * 0.25;
                    Vj[s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;

                    t1x = XX(i + 1, j + 1, k) - XX(i, j, k); t1y = YY(i + 1, j + 1, k) - YY(i, j, k); t1z = ZZ(i + 1, j + 1, k) - ZZ(i, j, k);
                    t2x = XX(i, j + 1, k) - XX(i + 1, j, k); t2y = YY(i, j + 1, k) - YY(i + 1, j, k); t2z = ZZ(i, j + 1, k) - ZZ(i + 1, j, k);
                    s1x = t1y * t2z - t1z * t2y; s1y = t1z * t2x - t1x * t2z; s1z = t1x * t2y - t1y * t2x;
                    ss = sqrt((s1x * s1x + s1y * s1y) + s1z * s1z);
                    b->Sk[s] = ss * 0.5;
                    if (ss >= LIM_ZERO) { b->nv[6][s] = s1x / ss; b->nv[7][s] = s1y / ss; b->nv[8][s] = s1z / ss; }
                    else { b->nv[6][s] = 1.0; b->nv[7][s] = 0.0; b->nv[8][s] = 0.0; }
                    xa = (((XX(i, j, k) + XX(i + 1, j, k)) + XX(i, j + 1, k)) + XX(i + 1, j + 1, k)) * 0.25;
                    ya = (((YY(i, j, k) + YY(i + 1, j, k)) + YY(i, j + 1, k)) + YY(i + 1, j + 1, k)) * 0.25;
                    za = (((ZZ(i, j, k) + ZZ(i + 1, j, k)) + ZZ(i, j + 1, k)) + ZZ(i + 1, j + 1, k)) * 0.25;
                    Vk[s] = ((s1x * xa + s1y * ya) + s1z * za) * 0.5;
                }
        for (int k = 1; k <= nz - 1; k++)
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx - 1; i++)
                    b->vol[IX(b, i, j, k)] = (((((Vi[IX(b, i + 1, j, k)] - Vi[IX(b, i, j, k)]) + Vj[IX(b, i, j + 1, k)]) - Vj[IX(b, i, j, k)])
                                              + Vk[IX(b, i, j, k + 1)]) - Vk[IX(b, i, j, k)]) / 3.0;
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx; i++) {
                    size_t s = IX(b, i, j, k);
                    b->xc[s] = (((((((XX(i, j, k) + XX(i, j + 1, k)) + XX(i, j, k + 1)) + XX(i, j + 1, k + 1)) + XX(i + 1, j, k)) + XX(i + 1, j + 1, k)) + XX(i + 1, j, k + 1)) + XX(i + 1, j + 1, k + 1)) * 0.125;
                    b->yc[s] = (((((((YY(i, j, k) + YY(i, j + 1, k)) + YY(i, j, k + 1)) + YY(i, j + 1, k + 1)) + YY(i + 1, j, k)) + YY(i + 1, j + 1, k)) + YY(i + 1, j, k + 1)) + YY(i + 1, j + 1, k + 1)) * 0.125;
                    b->zc[s] = (((((((ZZ(i, j, k) + ZZ(i, j + 1, k)) + ZZ(i, j, k + 1)) + ZZ(i, j + 1, k + 1)) + ZZ(i + 1, j, k)) + ZZ(i + 1, j + 1, k)) + ZZ(i + 1, j, k + 1)) + ZZ(i + 1, j + 1, k + 1)) * 0.125;
                }
        free(Vi); free(Vj); free(Vk);
#undef XX
#undef YY
#undef ZZ
        double *xc = b->xc, *yc = b->yc, *zc = b->zc;
#define XC(i, j, k) xc[IX(b, i, j, k)]
#define YC(i, j, k) yc[IX(b, i, j, k)]
#define ZC(i, j, k) zc[IX(b, i, j, k)]
        /* set 1: (I-1/2,J,K) faces */
        for (int k = 1; k <= nz - 1; k++)
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx; i++) {
                    double xi = XC(i, j, k) - XC(i - 1, j, k), yi = YC(i, j, k) - YC(i - 1, j, k), zi = ZC(i, j, k) - ZC(i - 1, j, k);
                    double xj1, yj1, zj1, xj2, yj2, zj2, xk1, yk1, zk1, xk2, yk2, zk2;
                    if ((i == 1 || i == nx) && (j == 1 || j == ny - 1)) {
                        xj1 = XC(i, j - 1, k); yj1 = YC(i, j - 1, k); zj1 = ZC(i, j - 1, k);
                        xj2 = XC(i, j + 1, k); yj2 = YC(i, j + 1, k); zj2 = ZC(i, j + 1, k);
                    } else {
                        xj1 = 0.5 * (XC(i, j - 1, k) + XC(i - 1, j - 1, k)); yj1 = 0.5 * (YC(i, j - 1, k) + YC(i - 1, j - 1, k)); zj1 = 0.5 * (ZC(i, j - 1, k) + ZC(i - 1, j - 1, k));
                        xj2 = 0.5 * (XC(i, j + 1, k) + XC(i - 1, j + 1, k)); yj2 = 0.5 * (YC(i, j + 1, k) + YC(i - 1, j + 1, k)); zj2 = 0.5 * (ZC(i, j + 1, k) + ZC(i - 1, j + 1, k));
                    }
                    if ((i == 1 || i == nx) && (k == 1 || k == nz - 1)) {
                        xk1 = XC(i, j, k - 1); yk1 = YC(i, j, k - 1); zk1 = ZC(i, j, k - 1);
                        xk2 = XC(i, j, k + 1); yk2 = YC(i, j, k + 1); zk2 = ZC(i, j, k + 1);
                    } else {
                        xk1 = 0.5 * (XC(i, j, k - 1) + XC(i - 1, j, k - 1)); yk1 = 0.5 * (YC(i, j, k - 1) + YC(i - 1, j, k - 1)); zk1 = 0.5 * (ZC(i, j, k - 1) + ZC(i - 1, j, k - 1));
                        xk2 = 0.5 * (XC(i, j, k + 1) + XC(i - 1, j, k + 1)); yk2 = 0.5 * (YC(i, j, k + 1) + YC(i - 1, j, k + 1)); zk2 = 0.5 * (ZC(i, j, k + 1) + ZC(i - 1, j, k + 1));
                    }
                    metric_face(b, b->met[1], i, j, k, xi, yi, zi, 0.5 * (xj2 - xj1), 0.5 * (yj2 - yj1), 0.5 * (zj2 - zj1),
                                0.5 * (xk2 - xk1), 0.5 * (yk2 - yk1), 0.5 * (zk2 - zk1));
                }
        /* set 2: (I,J-1/2,K) faces */
        for (int k = 1; k <= nz - 1; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx - 1; i++) {
                    double xj = XC(i, j, k) - XC(i, j - 1, k), yj = YC(i, j, k) - YC(i, j - 1, k), zj = ZC(i, j, k) - ZC(i, j - 1, k);
                    double xi1, yi1, zi1, xi2, yi2, zi2, xk1, yk1, zk1, xk2, yk2, zk2;
                    if ((j == 1 || j == ny) && (i == 1 || i == nx - 1)) {
                        xi1 = XC(i - 1, j, k); yi1 = YC(i - 1, j, k); zi1 = ZC(i - 1, j, k);
                        xi2 = XC(i + 1, j, k); yi2 = YC(i + 1, j, k); zi2 = ZC(i + 1, j, k);
                    } else {
                        xi1 = 0.5 * (XC(i - 1, j, k) + XC(i - 1, j - 1, k)); yi1 = 0.5 * (YC(i - 1, j, k) + YC(i - 1, j - 1, k)); zi1 = 0.5 * (ZC(i - 1, j, k) + ZC(i - 1, j - 1, k));
                        xi2 = 0.5 * (XC(i + 1, j, k) + XC(i + 1, j - 1, k)); yi2 = 0.5 * (YC(i + 1, j, k) + YC(i + 1, j - 1, k)); zi2 = 0.5 * (ZC(i + 1, j, k) + ZC(i + 1, j - 1, k));
                    }
                    if ((j == 1 || j == ny) && (k == 1 || k == nz - 1)) {
                        xk1 = XC(i, j, k - 1); yk1 = YC(i, j, k - 1); zk1 = ZC(i, j, k - 1);
                        xk2 = XC(i, j, k + 1); yk2 = YC(i, j, k + 1); zk2 = ZC(i, j, k + 1);
                    } else {
                        xk1 = 0.5 * (XC(i, j, k - 1) + XC(i, j - 1, k - 1)); yk1 = 0.5 * (YC(i, j, k - 1) + YC(i, j - 1, k - 1)); zk1 = 0.5 * (ZC(i, j, k - 1) + ZC(i, j - 1, k - 1));
                        xk2 = 0.5 * (XC(i, j, k + 1) + XC(i, j - 1, k + 1)); yk2 = 0.5 * (YC(i, j, k + 1) + YC(i, j - 1, k + 1)); zk2 = 0.5 * (ZC(i, j, k + 1) + ZC(i, j - 1, k + 1));
                    }
                    metric_face(b, b->met[2], i, j, k, 0.5 * (xi2 - xi1), 0.5 * (yi2 - yi1), 0.5 * (zi2 - zi1), xj, yj, zj,
                                0.5 * (xk2 - xk1), 0.5 * (yk2 - yk1), 0.5 * (zk2 - zk1));
                }
        /* set 3: (I,J,K-1/2) faces */
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx - 1; i++) {
                    double xk = XC(i, j, k) - XC(i, j, k - 1), yk = YC(i, j, k) - YC(i, j, k - 1), zk = ZC(i, j, k) - ZC(i, j, k - 1);
                    double xi1, yi1, zi1, xi2, yi2, zi2, xj1, yj1, zj1, xj2, yj2, zj2;
                    if ((k == 1 || k == nz) && (i == 1 || i == nx - 1)) {
                        xi1 = XC(i - 1, j, k); yi1 = YC(i - 1, j, k); zi1 = ZC(i - 1, j, k);
                        xi2 = XC(i + 1, j, k); yi2 = YC(i + 1, j, k); zi2 = ZC(i + 1, j, k);
                    } else {
                        xi1 = 0.5 * (XC(i - 1, j, k) + XC(i - 1, j, k - 1)); yi1 = 0.5 * (YC(i - 1, j, k) + YC(i - 1, j, k - 1)); zi1 = 0.5 * (ZC(i - 1, j, k) + ZC(i - 1, j, k - 1));
                        xi2 = 0.5 * (XC(i + 1, j, k) + XC(i + 1, j, k - 1)); yi2 = 0.5 * (YC(i + 1, j, k) + YC(i + 1, j, k - 1)); zi2 = 0.5 * (ZC(i + 1, j, k) + ZC(i + 1, j, k - 1));
                    }
                    if ((k == 1 || k == nz) && (j == 1 || j == ny - 1)) {
                        xj1 = XC(i, j - 1, k); yj1 = YC(i, j - 1, k); zj1 = ZC(i, j - 1, k);
                        xj2 = XC(i, j + 1, k); yj2 = YC(i, j + 1, k); zj2 = ZC(i, j + 1, k);
                    } else {
                        xj1 = 0.5 * (XC(i, j - 1, k) + XC(i, j - 1, k - 1)); yj1 = 0.5 * (YC(i, j - 1, k) + YC(i, j - 1, k - 1)); zj1 = 0.5 * (ZC(i, j - 1, k) + ZC(i, j - 1, k - 1));
                        xj2 = 0.5 * (XC(i, j + 1, k) + XC(i, j + 1, k - 1)); yj2 = 0.5 * (YC(i, j + 1, k) + YC(i, j + 1, k - 1)); zj2 = 0.5 * (ZC(i, j + 1, k) + ZC(i, j + 1, k - 1));
                    }
                    metric_face(b, b->met[3], i, j, k, 0.5 * (xi2 - xi1), 0.5 * (yi2 - yi1), 0.5 * (zi2 - zi1),
                                0.5 * (xj2 - xj1), 0.5 * (yj2 - yj1), 0.5 * (zj2 - zj1), xk, yk, zk);
                }
        /* set 0: cell centres, finest mesh only */
        for (int k = 1; k <= nz - 1; k++)
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx - 1; i++)
                    metric_face(b, b->met[0], i, j, k,
                                XC(i + 1, j, k) - XC(i - 1, j, k), YC(i + 1, j, k) - YC(i - 1, j, k), ZC(i + 1, j, k) - ZC(i - 1, j, k),
                                XC(i, j + 1, k) - XC(i, j - 1, k), YC(i, j + 1, k) - YC(i, j - 1, k), ZC(i, j + 1, k) - ZC(i, j - 1, k),
                                XC(i, j, k + 1) - XC(i, j, k - 1), YC(i, j, k + 1) - YC(i, j, k - 1), ZC(i, j, k + 1) - ZC(i, j, k - 1));
#undef XC
#undef YC
#undef ZC
    }
}

/* Halo index map shared by the cell-centre and conserved-variable exchanges
 * (sub_update_buffer_mpi.f90:70-94; sub_udate_Meshlink_mpi.f90:58-82). */
void orc_halo_map(const orc_bcmsg *B, int layers, int *kb, int *ke, int *kb1, int *ke1, int *L, int *P, int *ks) {
    (void)layers;
    kb[0] = B->ib; ke[0] = B->ie - 1; kb[1] = B->jb; ke[1] = B->je - 1; kb[2] = B->kb; ke[2] = B->ke - 1;
    int k = (B->face - 1) % 3;
    if (B->face > 3) kb[k] = kb[k] - LAP;
    ke[k] = kb[k] + LAP - 1;
    kb1[0] = B->ib1; ke1[0] = B->ie1 - 1; kb1[1] = B->jb1; ke1[1] = B->je1 - 1; kb1[2] = B->kb1; ke1[2] = B->ke1 - 1;
    k = (B->face1 - 1) % 3;
    if (B->face1 <= 3) kb1[k] = kb1[k] - LAP;
    ke1[k] = kb1[k] + LAP - 1;
    L[0] = abs(B->L1); P[0] = isign1(B->L1);
    L[1] = abs(B->L2); P[1] = isign1(B->L2);
    L[2] = abs(B->L3); P[2] = isign1(B->L3);
    for (int q = 0; q < 3; q++) ks[q] = P[q] > 0 ? kb[q] : ke[q];
}

/* update_Mesh_Center: Mesh_send_mpi (sub_udate_Meshlink_mpi.f90:30-133) + Mesh_Center_Periodic (:188-249). */
void orc_mesh_center_halo(orc_ctx *c) {
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc >= 0) continue;
            int kb[3], ke[3], kb1[3], ke1[3], L[3], P[3], ks[3];
            orc_halo_map(B, LAP, kb, ke, kb1, ke1, L, P, ks);
            orc_block *b1 = &c->blk[B->nb1 - 1];
            for (int k = kb1[2]; k <= ke1[2]; k++)
                for (int j = kb1[1]; j <= ke1[1]; j++)
                    for (int i = kb1[0]; i <= ke1[0]; i++) {
                        int ka[3] = {i - kb1[0], j - kb1[1], k - kb1[2]};
                        int i1 = ks[0] + ka[L[0] - 1] * P[0], j1 = ks[1] + ka[L[1] - 1] * P[1], k1 = ks[2] + ka[L[2] - 1] * P[2];
                        b1->xc[IX(b1, i, j, k)] = b->xc[IX(b, i1, j1, k1)];
                        b1->yc[IX(b1, i, j, k)] = b->yc[IX(b, i1, j1, k1)];
                        b1->zc[IX(b1, i, j, k)] = b->zc[IX(b, i1, j1, k1)];
                    }
        }
    }
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc != -2 && B->bc != -3) continue;
            double Xp, Yp, Zp;
            if (B->bc == -2) { Xp = -c->Periodic_dX; Yp = -c->Periodic_dY; Zp = -c->Periodic_dZ; }
            else { Xp = c->Periodic_dX; Yp = c->Periodic_dY; Zp = c->Periodic_dZ; }
            int kb[3] = {B->ib, B->jb, B->kb}, ke[3] = {B->ie - 1, B->je - 1, B->ke - 1};
            int kd = (B->face - 1) % 3;
            if (B->face <= 3) kb[kd] -= LAP;
            ke[kd] = kb[kd] + LAP - 1;
            for (int k = kb[2]; k <= ke[2]; k++)
                for (int j = kb[1]; j <= ke[1]; j++)
                    for (int i = kb[0]; i <= ke[0]; i++) {
                        size_t s = IX(b, i, j, k);
                        b->xc[s] = b->xc[s] + Xp; b->yc[s] = b->yc[s] + Yp; b->zc[s] = b->zc[s] + Zp;
                    }
        }
    }
}

/* check_mesh_quality_onemesh: the dtime_mesh factor (sub_geometry.f90:361-515). */
void orc_mesh_quality(orc_ctx *c) {
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        double *xc = b->xc, *yc = b->yc, *zc = b->zc;
#define XC(i, j, k) xc[IX(b, i, j, k)]
#define YC(i, j, k) yc[IX(b, i, j, k)]
#define ZC(i, j, k) zc[IX(b, i, j, k)]
#define SQ(a) ((a) * (a))
        for (int k = 1; k <= b->nz - 1; k++)
            for (int j = 1; j <= b->ny - 1; j++)
                for (int i = 1; i <= b->nx - 1; i++) {
                    double fl1 = sqrt((SQ(XC(i, j, k) - XC(i - 1, j, k)) + SQ(YC(i, j, k) - YC(i - 1, j, k))) + SQ(ZC(i, j, k) - ZC(i - 1, j, k)))
                               / sqrt((SQ(XC(i, j, k) - XC(i + 1, j, k)) + SQ(YC(i, j, k) - YC(i + 1, j, k))) + SQ(ZC(i, j, k) - ZC(i + 1, j, k)));
                    double fl2 = sqrt((SQ(XC(i, j, k) - XC(i, j - 1, k)) + SQ(YC(i, j, k) - YC(i, j - 1, k))) + SQ(ZC(i, j, k) - ZC(i, j - 1, k)))
                               / sqrt((SQ(XC(i, j, k) - XC(i, j + 1, k)) + SQ(YC(i, j, k) - YC(i, j + 1, k))) + SQ(ZC(i, j, k) - ZC(i, j + 1, k)));
                    double fl3 = sqrt((SQ(XC(i, j, k) - XC(i, j, k - 1)) + SQ(YC(i, j, k) - YC(i, j, k - 1))) + SQ(ZC(i, j, k) - ZC(i, j, k - 1)))
                               / sqrt((SQ(XC(i, j, k) - XC(i, j, k + 1)) + SQ(YC(i, j, k) - YC(i, j, k + 1))) + SQ(ZC(i, j, k) - ZC(i, j, k + 1)));
                    if (fl1 < 1.0) fl1 = 1.0 / fl1;
                    if (fl2 < 1.0) fl2 = 1.0 / fl2;
                    if (fl3 < 1.0) fl3 = 1.0 / fl3;
                    double fl = fl1; if (fl2 > fl) fl = fl2; if (fl3 > fl) fl = fl3;
                    double x1, x2, y1, y2, z1, z2, ft1, ft2, ft3;
                    x1 = XC(i, j, k) - XC(i - 1, j, k); x2 = XC(i + 1, j, k) - XC(i, j, k);
                    y1 = YC(i, j, k) - YC(i - 1, j, k); y2 = YC(i + 1, j, k) - YC(i, j, k);
                    z1 = ZC(i, j, k) - ZC(i - 1, j, k); z2 = ZC(i + 1, j, k) - ZC(i, j, k);
                    ft1 = ((x1 * x2 + y1 * y2) + z1 * z2) / sqrt(((x1 * x1 + y1 * y1) + z1 * z1) * ((x2 * x2 + y2 * y2) + z2 * z2));
                    x1 = XC(i, j, k) - XC(i, j - 1, k); x2 = XC(i, j + 1, k) - XC(i, j, k);
                    y1 = YC(i, j, k) - YC(i, j - 1, k); y2 = YC(i, j + 1, k) - YC(i, j, k);
                    z1 = ZC(i, j, k) - ZC(i, j - 1, k); z2 = ZC(i, j + 1, k) - ZC(i, j, k);
                    ft2 = ((x1 * x2 + y1 * y2) + z1 * z2) / sqrt(((x1 * x1 + y1 * y1) + z1 * z1) * ((x2 * x2 + y2 * y2) + z2 * z2));
                    x1 = XC(i, j, k) - XC(i, j, k - 1); x2 = XC(i, j, k + 1) - XC(i, j, k);
                    y1 = YC(i, j, k) - YC(i, j, k - 1); y2 = YC(i, j, k + 1) - YC(i, j, k);
                    z1 = ZC(i, j, k) - ZC(i, j, k - 1); z2 = ZC(i, j, k + 1) - ZC(i, j, k);
                    ft3 = ((x1 * x2 + y1 * y2) + z1 * z2) / sqrt(((x1 * x1 + y1 * y1) + z1 * z1) * ((x2 * x2 + y2 * y2) + z2 * z2));
                    double mn = ft1; if (ft2 < mn) mn = ft2; if (ft3 < mn) mn = ft3;
                    double ft = 1.0 * mn; if (ft > 1.0) ft = 1.0;
                    ft = acos(ft);
                    double Af = 0.9 * exp(-4.0 * SQ(fl - 1.0)) + 0.1;
                    double a = (4.0 / 3.1415926535) * ft;
                    double Aa = 0.9 * exp(-(a * a)) + 0.1;
                    b->dtime_mesh[IX(b, i, j, k)] = Af < Aa ? Af : Aa;
                }
#undef XC
#undef YC
#undef ZC
#undef SQ
    }
}

/* wall_point + comput_wall_dist (sub_comput_dw.f90:43-168).  The minimum of the Euclidean
 * distances equals the square root of the minimum squared distance (sqrt is monotone and
 * correctly rounded), which is what is evaluated here. */
void orc_wall_distance(orc_ctx *c) {
    long npw = 0;
    for (int m = 0; m < c->nblk; m++)
        for (int q = 0; q < c->blk[m].nsub; q++) {
            orc_bcmsg *B = &c->blk[m].bc[q];
            if (B->bc == BCW) npw += (long)(B->ie - B->ib + 1) * (B->je - B->jb + 1) * (B->ke - B->kb + 1);
        }
    double *Xw = (double *)malloc((size_t)(3 * npw + 3) * sizeof(double));
    long k0 = 0;
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc != BCW) continue;
            for (int k = B->kb; k <= B->ke; k++)
                for (int j = B->jb; j <= B->je; j++)
                    for (int i = B->ib; i <= B->ie; i++) {
                        Xw[3 * k0] = b->x[IX(b, i, j, k)]; Xw[3 * k0 + 1] = b->y[IX(b, i, j, k)]; Xw[3 * k0 + 2] = b->z[IX(b, i, j, k)];
                        k0++;
                    }
        }
    }
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        double *dis = zalloc(b->n);
        int nx = b->nx, ny = b->ny, nz = b->nz;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    double px = b->x[IX(b, i, j, k)], py = b->y[IX(b, i, j, k)], pz = b->z[IX(b, i, j, k)];
                    double best = 1.0e16;   /* d_init^2 with d_init = 1e8 */
                    for (long q = 0; q < npw; q++) {
                        double dx = px - Xw[3 * q], dy = py - Xw[3 * q + 1], dz = pz - Xw[3 * q + 2];
                        double s = (dx * dx + dy * dy) + dz * dz;
                        if (s < best) best = s;
                    }
                    dis[IX(b, i, j, k)] = npw > 0 ? sqrt(best) : 1.0e8;
                }
        for (int k = 1; k <= nz - 1; k++)
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx - 1; i++)
                    b->dw[IX(b, i, j, k)] = (((((((dis[IX(b, i, j, k)] + dis[IX(b, i, j + 1, k)]) + dis[IX(b, i, j, k + 1)]) + dis[IX(b, i, j + 1, k + 1)])
                                              + dis[IX(b, i + 1, j, k)]) + dis[IX(b, i + 1, j + 1, k)]) + dis[IX(b, i + 1, j, k + 1)]) + dis[IX(b, i + 1, j + 1, k + 1)]) * 0.125;
        free(dis);
    }
    free(Xw);
}

/* set_BcK (sub_init.f90:545-577) */
void orc_set_bck(orc_ctx *c) {
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (int q = 0; q < b->nsub; q++) {
            orc_bcmsg *B = &b->bc[q];
            if (B->bc < 0) continue;
            int side = B->face <= 3 ? 0 : 1;
            switch (B->face) {
            case 1: case 4:
                for (int k = B->kb; k <= B->ke - 1; k++) for (int j = B->jb; j <= B->je - 1; j++) b->BcI[side][JK(b, j, k)] = 1;
                break;
            case 2: case 5:
                for (int k = B->kb; k <= B->ke - 1; k++) for (int i = B->ib; i <= B->ie - 1; i++) b->BcJ[side][IK(b, i, k)] = 1;
                break;
            default:
                for (int j = B->jb; j <= B->je - 1; j++) for (int i = B->ib; i <= B->ie - 1; i++) b->BcK[side][IJ(b, i, j)] = 1;
            }
        }
    }
}

int orc_setup(orc_ctx *c) {
    if (link_subfaces(c)) return -1;
    orc_coordinate_ghosts(c);
    orc_geometry(c);
    orc_mesh_center_halo(c);
    if (c->turb != 0) orc_wall_distance(c);
    orc_set_bck(c);
    orc_mesh_quality(c);
    return 0;
}

/* init_flow_zero, freestream branch (sub_init.f90:302-410) */
int orc_init_flow(orc_ctx *c) {
    double d0 = 1.0, p0 = 1.0 / (c->gamma * c->Ma * c->Ma);
    double u0 = cos(c->A_alfa) * cos(c->A_beta), v0 = sin(c->A_alfa) * cos(c->A_beta), w0 = sin(c->A_beta);
    for (int m = 0; m < c->nblk; m++) {
        orc_block *b = &c->blk[m];
        for (size_t s = 0; s < b->n; s++) {
            b->U[0][s] = d0; b->U[1][s] = d0 * u0; b->U[2][s] = d0 * v0; b->U[3][s] = d0 * w0;
            b->U[4][s] = p0 / (c->gamma - 1.0) + (0.5 * d0) * ((u0 * u0 + v0 * v0) + w0 * w0);
            if (c->NVAR == 6) b->U[5][s] = 5.0;
            else if (c->NVAR == 7) { b->U[5][s] = 10.0 * c->Kt_inf; b->U[6][s] = c->Wt_inf; }
        }
    }
    orc_boundary(c);
    orc_halo(c);
    return 0;
}

int orc_set_interior_U(orc_ctx *c, int blk, const double *U) {
    orc_block *b = &c->blk[blk - 1];
    int nv = c->NVAR;
    for (int k = 1; k <= b->nz - 1; k++)
        for (int j = 1; j <= b->ny - 1; j++)
            for (int i = 1; i <= b->nx - 1; i++)
                for (int m = 0; m < nv; m++) {
                    size_t s = (size_t)m + (size_t)nv * ((size_t)(i - 1) + (size_t)(b->nx - 1) * ((size_t)(j - 1) + (size_t)(b->ny - 1) * (size_t)(k - 1)));
                    b->U[m][IX(b, i, j, k)] = U[s];
                }
    return 0;
}

int orc_bc_halo(orc_ctx *c) { orc_boundary(c); orc_halo(c); return 0; }

void orc_block_dims(orc_ctx *c, int blk, int *out) {
    orc_block *b = &c->blk[blk - 1];
    out[0] = b->nx; out[1] = b->ny; out[2] = b->nz; out[3] = b->NI; out[4] = b->NJ; out[5] = b->NK;
    out[6] = c->NVAR; out[7] = b->IF_OverLimit; out[8] = b->nneg_last;
}

/* copy an array out: names as in the reference (U1..U7 = U(m), Res1.., DU1.., metric sets as
 * ix1/.../kz0, Surf<f>_<c>).  Full padded box, i fastest. */
int orc_get_array(orc_ctx *c, int blk, const char *name, double *out) {
    orc_block *b = &c->blk[blk - 1];
    const double *src = NULL;
    static const char *mnames[9] = {"ix", "iy", "iz", "jx", "jy", "jz", "kx", "ky", "kz"};
    static const char *nnames[9] = {"ni1", "ni2", "ni3", "nj1", "nj2", "nj3", "nk1", "nk2", "nk3"};
    struct { const char *n; double *p; } tab[] = {
        {"x", b->x}, {"y", b->y}, {"z", b->z}, {"xc", b->xc}, {"yc", b->yc}, {"zc", b->zc}, {"vol", b->vol},
        {"Si", b->Si}, {"Sj", b->Sj}, {"Sk", b->Sk}, {"dw", b->dw}, {"dtime_mesh", b->dtime_mesh},
        {"dt", b->dt}, {"mu", b->mu}, {"mu_t", b->mut}, {"Lci", b->Lci}, {"Lcj", b->Lcj}, {"Lck", b->Lck},
        {"Lvi", b->Lvi}, {"Lvj", b->Lvj}, {"Lvk", b->Lvk}};
    for (size_t q = 0; q < sizeof(tab) / sizeof(tab[0]); q++) if (!strcmp(name, tab[q].n)) src = tab[q].p;
    for (int q = 0; q < 9 && !src; q++) if (!strcmp(name, nnames[q])) src = b->nv[q];
    for (int s = 0; s < 4 && !src; s++)
        for (int q = 0; q < 9 && !src; q++) {
            char nm[8]; snprintf(nm, sizeof nm, "%s%d", mnames[q], s);
            if (!strcmp(name, nm)) src = b->met[s][q];
        }
    if (!src && (name[0] == 'U' || name[0] == 'R' || name[0] == 'D') ) {
        int m = -1;
        if (!strncmp(name, "Un", 2) && name[2] >= '1' && name[2] <= '7' && !name[3]) { m = name[2] - '1'; src = m < c->NVAR ? b->Un[m] : NULL; }
        else if (name[0] == 'U' && name[1] >= '1' && name[1] <= '7' && !name[2]) { m = name[1] - '1'; src = m < c->NVAR ? b->U[m] : NULL; }
        else if (!strncmp(name, "Res", 3) && name[3] >= '1' && name[3] <= '7' && !name[4]) { m = name[3] - '1'; src = m < c->NVAR ? b->Res[m] : NULL; }
        else if (!strncmp(name, "DU", 2) && name[2] >= '1' && name[2] <= '7' && !name[3]) { m = name[2] - '1'; src = m < c->NVAR ? b->DU[m] : NULL; }
    }
    if (src) { memcpy(out, src, b->n * sizeof(double)); return 0; }
    if (!strncmp(name, "Surf", 4) && name[4] >= '1' && name[4] <= '6' && name[5] == '_' && name[6] >= '1' && name[6] <= '3') {
        size_t n2 = (size_t)(b->NI > b->NJ ? b->NI : b->NJ) * (size_t)(b->NJ > b->NK ? b->NJ : b->NK) * 2;
        memcpy(out, b->Surf[name[4] - '1'][name[6] - '1'], n2 * sizeof(double));
        return 0;
    }
    return -1;
}

int orc_set_array(orc_ctx *c, int blk, const char *name, const double *in) {
    orc_block *b = &c->blk[blk - 1];
    if (name[0] == 'U' && name[1] >= '1' && name[1] <= '7' && !name[2]) {
        int m = name[1] - '1';
        if (m >= c->NVAR) return -1;
        memcpy(b->U[m], in, b->n * sizeof(double));
        return 0;
    }
    return -1;
}

int orc_get_state(orc_ctx *c, int *kstep, double *tt, double *res_max, double *res_rms, double *force12) {
    *kstep = c->Kstep; *tt = c->tt;
    for (int m = 0; m < c->NVAR; m++) { res_max[m] = c->Res_max[m]; res_rms[m] = c->Res_rms[m]; }
    for (int q = 0; q < 12; q++) force12[q] = c->force[q];
    return 0;
}
