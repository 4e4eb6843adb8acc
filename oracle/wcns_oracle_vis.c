/*
 * wcns_oracle_vis.c -- CPU oracle (TEST INFRASTRUCTURE ONLY): laminar viscous
 * residual and viscous spectral radii, restating
 *   sub_tgh.f90:1511-1810  WCNSE5_VIS_virtual
 *   sub_tgh.f90:1909-2029  UVWT_DER_4th_virtual -> :2217-2272 DUVWT_NODE_LINE_virtual
 *   sub_tgh.f90:2032-2155  UVWT_DER_4th_half_virtual -> :2277-2330 DUVWT_half_line_virtual
 *   sub_tgh.f90:2335-2387  VALUE_line_half_virtual
 *   newvis.f90:421-554     VALUE_HALF_NODE_IJK
 *   WCNSE5.f90:854-890     DUVWT_DXYZ
 *   newvis.f90:283-349     FLUX_VIS_LINE_new
 *   LHS.f90:243-343        spectvisl
 * in the reference's operation order.
 */
#include "wcns_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMALL 1.0e-38
#define NIJK2ND 4

#include "wcns_oracle_priv.h"

static inline long G(const oblock_view *b, int i, int j, int k)
{
    return (long)(i + 2) + (long)(b->ni + 6) * ((long)(j + 2) + (long)(b->nj + 6) * (long)(k + 2));
}

/* DUVWT_NODE_LINE_virtual: uvwt(4,-2:n+3) -> duvwt(4,1:n), n > 4 */
static void duvwt_node_line(int n, const double *uv /* [4*(n+6)] at (m, i+2) */, double *d /* [4*n] */)
{
#define U(m,i) uv[(m) + 4 * ((i) + 2)]
#define D(m,i) d[(m) + 4 * ((i) - 1)]
    const double A1 = 8.0, B1 = -1.0, dd12 = 12.0, dd6 = 6.0;
    const double A2 = -3.0, B2 = -10.0, C2 = 18.0, D2 = -6.0, E2 = 1.0;
    int st = 1, ed = n;
    if (U(3, -2) < 0.0) {
        st = 3;
        for (int m = 0; m < 4; m++) {
            D(m, 1) = (-11. * U(m, 1) + 18. * U(m, 2) - 9. * U(m, 3) + 2. * U(m, 4)) / dd6;
            D(m, 2) = (A2 * U(m, 1) + B2 * U(m, 2) + C2 * U(m, 3) + D2 * U(m, 4) + E2 * U(m, 5)) / dd12;
        }
    }
    if (U(3, n + 3) < 0.0) {
        ed = n - 2;
        for (int m = 0; m < 4; m++) {
            D(m, n) = -(-11. * U(m, n) + 18. * U(m, n - 1) - 9.0 * U(m, n - 2) + 2. * U(m, n - 3)) / dd6;
            D(m, n - 1) = -(A2 * U(m, n) + B2 * U(m, n - 1) + C2 * U(m, n - 2) + D2 * U(m, n - 3) + E2 * U(m, n - 4)) / dd12;
        }
    }
    for (int i = st; i <= ed; i++)
        for (int m = 0; m < 4; m++)
            D(m, i) = (A1 * (U(m, i + 1) - U(m, i - 1)) + B1 * (U(m, i + 2) - U(m, i - 2))) / 12.0;
#undef U
#undef D
}

/* DUVWT_half_line_virtual: uvwt(4,-2:n+3) -> duvwt(4,0:n), n > 4 */
static void duvwt_half_line(int n, const double *uv, double *d /* [4*(n+1)] */)
{
#define U(m,i) uv[(m) + 4 * ((i) + 2)]
#define D(m,i) d[(m) + 4 * (i)]
    const double A1 = 27.0, B1 = -1.0, dd24 = 24.0;
    const double A2 = -22.0, B2 = 17.0, C2 = 9.0, D2 = -5.0, E2 = 1.0;
    const double A3 = -71.0, B3 = 141.0, C3 = -93.0, D3 = 23.0;
    int st = 0, ed = n;
    if (U(3, -2) < 0.0) {
        st = 2;
        for (int m = 0; m < 4; m++) {
            D(m, 0) = (A3 * U(m, 1) + B3 * U(m, 2) + C3 * U(m, 3) + D3 * U(m, 4)) / dd24;
            D(m, 1) = (A2 * U(m, 1) + B2 * U(m, 2) + C2 * U(m, 3) + D2 * U(m, 4) + E2 * U(m, 5)) / dd24;
        }
    }
    if (U(3, n + 3) < 0.0) {
        ed = n - 2;
        for (int m = 0; m < 4; m++) {
            D(m, n) = -(A3 * U(m, n) + B3 * U(m, n - 1) + C3 * U(m, n - 2) + D3 * U(m, n - 3)) / dd24;
            D(m, n - 1) = -(A2 * U(m, n) + B2 * U(m, n - 1) + C2 * U(m, n - 2) + D2 * U(m, n - 3) + E2 * U(m, n - 4)) / dd24;
        }
    }
    for (int i = st; i <= ed; i++)
        for (int m = 0; m < 4; m++)
            D(m, i) = (A1 * (U(m, i + 1) - U(m, i)) + B1 * (U(m, i + 2) - U(m, i - 1))) / dd24;
#undef U
#undef D
}

/* VALUE_line_half_virtual: q(4,-2:n+3) -> q_half(4,0:n), n > 4 */
static void value_line_half(int n, const double *q, double *qh)
{
#define Q(m,i) q[(m) + 4 * ((i) + 2)]
#define H(m,i) qh[(m) + 4 * (i)]
    const double A1 = 9.0, B1 = -1.0, dd16 = 16.0;
    const double A2 = 5.0, B2 = 15.0, C2 = -5.0, D2 = 1.0;
    const double A3 = 35.0, B3 = -35.0, C3 = 21.0, D3 = -5.0;
    int st = 0, ed = n;
    if (Q(3, -2) < 0.0) {
        st = 2;
        for (int m = 0; m < 4; m++) {
            H(m, 0) = (A3 * Q(m, 1) + B3 * Q(m, 2) + C3 * Q(m, 3) + D3 * Q(m, 4)) / dd16;
            H(m, 1) = (A2 * Q(m, 1) + B2 * Q(m, 2) + C2 * Q(m, 3) + D2 * Q(m, 4)) / dd16;
        }
    }
    if (Q(3, n + 3) < 0.0) {
        ed = n - 2;
        for (int m = 0; m < 4; m++) {
            H(m, n) = (A3 * Q(m, n) + B3 * Q(m, n - 1) + C3 * Q(m, n - 2) + D3 * Q(m, n - 3)) / dd16;
            H(m, n - 1) = (A2 * Q(m, n) + B2 * Q(m, n - 1) + C2 * Q(m, n - 2) + D2 * Q(m, n - 3)) / dd16;
        }
    }
    for (int i = st; i <= ed; i++)
        for (int m = 0; m < 4; m++)
            H(m, i) = (A1 * (Q(m, i) + Q(m, i + 1)) + B1 * (Q(m, i + 2) + Q(m, i - 1))) / dd16;
#undef Q
#undef H
}

/* DUVWT_DXYZ (12 derivatives, 9 metrics, vol) at one half node */
static void duvwt_dxyz(const double *du /*12*/, const double *kx /*9*/, double vol, double *out /*12*/)
{
    for (int m = 1; m <= 3; m++) {
        int m1 = m + m + m - 2, m2 = m1 + 1, m3 = m2 + 1;
        out[m - 1] = (kx[m1 - 1] * du[0] + kx[m2 - 1] * du[4] + kx[m3 - 1] * du[8]) / vol;
        out[m + 2] = (kx[m1 - 1] * du[1] + kx[m2 - 1] * du[5] + kx[m3 - 1] * du[9]) / vol;
        out[m + 5] = (kx[m1 - 1] * du[2] + kx[m2 - 1] * du[6] + kx[m3 - 1] * du[10]) / vol;
        out[m + 8] = (kx[m1 - 1] * du[3] + kx[m2 - 1] * du[7] + kx[m3 - 1] * du[11]) / vol;
    }
}

/* FLUX_VIS_LINE_new at one half node */
static void flux_vis(const double *uvwt, const double *nxyz, const double *d, double vs, double kcp, double *fv)
{
    double CC = 2.0 / 3.0;
    double vscc = vs * CC;
    double dudx = d[0], dudy = d[1], dudz = d[2], dvdx = d[3], dvdy = d[4], dvdz = d[5];
    double dwdx = d[6], dwdy = d[7], dwdz = d[8];
    double t1 = vscc * (2.0 * dudx - dvdy - dwdz);
    double t5 = vscc * (2.0 * dvdy - dwdz - dudx);
    double t9 = vscc * (2.0 * dwdz - dudx - dvdy);
    double t2 = vs * (dudy + dvdx);
    double t3 = vs * (dudz + dwdx);
    double t6 = vs * (dvdz + dwdy);
    double t4 = t2, t7 = t3, t8 = t6;
    double nx = nxyz[0], ny = nxyz[1], nz = nxyz[2];
    fv[0] = 0.0;
    fv[1] = t1 * nx + t2 * ny + t3 * nz;
    fv[2] = t4 * nx + t5 * ny + t6 * nz;
    fv[3] = t7 * nx + t8 * ny + t9 * nz;
    fv[4] = uvwt[0] * fv[1] + uvwt[1] * fv[2] + uvwt[2] * fv[3] + kcp * (d[9] * nx + d[10] * ny + d[11] * nz);
}

int o_rhs_vis(o_case *c)
{
    o_params prm;
    double visc;
    o_view_params(c, &prm, &visc);
    double re = 1.0 / prm.reynolds;
    double cp = 1.0 / ((prm.gama - 1.0) * prm.moo * prm.moo);
    double cp_prl = cp / prm.prl, cp_prt = cp / prm.prt;
    const int nthr = o_view_threads(c);
    for (int nb = 0;; nb++) {
        oblock_view B;
        if (o_view_block(c, nb, &B)) break;
        oblock_view *b = &B;
        int ni = b->ni, nj = b->nj, nk = b->nk;
        if (ni <= NIJK2ND || nj <= NIJK2ND || nk <= NIJK2ND) return -100;
        /* RHS.f90:22: the turbulence model advances first, in the same block loop */
        if (prm.nlamtur >= 0 && o_turbulence(c, nb)) return -101;
        int nmax = ni > nj ? ni : nj; nmax = nmax > nk ? nmax : nk;
        long N = (long)ni * nj * nk;
        double *duvwt = (double *)calloc(12 * N, sizeof(double));        /* duvwt(12,i,j,k) */
#define DNI(m,i,j,k) duvwt[(m) + 12 * ((long)((i) - 1) + (long)ni * ((long)((j) - 1) + (long)nj * ((k) - 1)))]
#define DN(m, ...) DNI(m, __VA_ARGS__)
        /* UVWT_DER_4th_virtual; the pencils of one direction write disjoint
         * nodes, so threads take whole pencils with line buffers of their own */
        for (int dir = 0; dir < 3; dir++) {
            int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
            int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
            #pragma omp parallel num_threads(nthr) if (nthr > 1)
            {
            double *uv = (double *)calloc(4 * (nmax + 6), sizeof(double));
            double *dt = (double *)calloc(4 * (nmax + 1), sizeof(double));
            #pragma omp for collapse(2) schedule(static)
            for (int b2 = 1; b2 <= nbb; b2++) for (int b1 = 1; b1 <= na; b1++) {
#define IJK(l) (dir == 0 ? (l) : b1), (dir == 1 ? (l) : (dir == 0 ? b1 : b2)), (dir == 2 ? (l) : b2)
                int st, ed;
                if (b->r[G(b, IJK(-2))] < 0.0) { uv[3 + 4 * 0] = -1.0; st = 0; } else st = -2;
                if (b->r[G(b, IJK(n + 3))] < 0.0) { uv[3 + 4 * (n + 5)] = -1.0; ed = n + 1; } else ed = n + 3;
                for (int l = st; l <= ed; l++) {
                    long q = G(b, IJK(l));
                    uv[0 + 4 * (l + 2)] = b->u[q]; uv[1 + 4 * (l + 2)] = b->v[q];
                    uv[2 + 4 * (l + 2)] = b->w[q]; uv[3 + 4 * (l + 2)] = b->t[q];
                }
                duvwt_node_line(n, uv, dt);
                for (int l = 1; l <= n; l++)
                    for (int m = 0; m < 4; m++) DN(m + 4 * dir, IJK(l)) = dt[m + 4 * (l - 1)];
            }
            free(uv); free(dt);
            }
        }
        /* the three directions of WCNSE5_VIS_virtual */
        for (int dir = 0; dir < 3; dir++) {
            int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
            int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
            #pragma omp parallel num_threads(nthr) if (nthr > 1)
            {
        double *uv = (double *)calloc(4 * (nmax + 6), sizeof(double));
        double *uvh = (double *)calloc(4 * (nmax + 1), sizeof(double));
        double *dmid = (double *)calloc(4 * (nmax + 1), sizeof(double));
        double *kx = (double *)calloc(9 * nmax, sizeof(double)), *kxh = (double *)calloc(9 * (nmax + 1), sizeof(double));
        double *vl = (double *)calloc(nmax, sizeof(double)), *vlh = (double *)calloc(nmax + 1, sizeof(double));
        double *s1 = (double *)calloc(nmax, sizeof(double)), *s1h = (double *)calloc(nmax + 1, sizeof(double));
        double *s2 = (double *)calloc(nmax, sizeof(double)), *s2h = (double *)calloc(nmax + 1, sizeof(double));
        double *dl = (double *)calloc(12 * nmax, sizeof(double)), *dh = (double *)calloc(12 * (nmax + 1), sizeof(double));
        double *fv = (double *)calloc(5 * (nmax + 1), sizeof(double)), *dfv = (double *)calloc(5 * nmax, sizeof(double));
            #pragma omp for collapse(2) schedule(static)
            for (int b2 = 1; b2 <= nbb; b2++) for (int b1 = 1; b1 <= na; b1++) {
                for (int l = 1; l <= n; l++) {
                    long q = G(b, IJK(l));
                    double v1 = b->visl[q], v2 = b->visl[q] * cp_prl;
                    for (int mv = 1; mv <= prm.nvis; mv++) { v1 = v1 + b->vist[q]; v2 = v2 + b->vist[q] * cp_prt; }
                    s1[l - 1] = v1; s2[l - 1] = v2;
                    double *k9 = &kx[9 * (l - 1)];
                    k9[0] = b->kcx[q]; k9[1] = b->etx[q]; k9[2] = b->ctx[q];
                    k9[3] = b->kcy[q]; k9[4] = b->ety[q]; k9[5] = b->cty[q];
                    k9[6] = b->kcz[q]; k9[7] = b->etz[q]; k9[8] = b->ctz[q];
                    vl[l - 1] = b->vol[q];
                    for (int m = 0; m < 12; m++) dl[12 * (l - 1) + m] = DN(m, IJK(l));
                }
                /* the uvwt line (also used for the direct half-node derivative, UVWT_DER_4th_half_virtual) */
                int st, ed;
                memset(uv, 0, sizeof(double) * 4 * (nmax + 6));
                if (b->r[G(b, IJK(-2))] < 0.0) { uv[3] = -1.0; st = 0; } else st = -2;
                if (b->r[G(b, IJK(n + 3))] < 0.0) { uv[3 + 4 * (n + 5)] = -1.0; ed = n + 1; } else ed = n + 3;
                for (int l = st; l <= ed; l++) {
                    long q = G(b, IJK(l));
                    uv[4 * (l + 2)] = b->u[q]; uv[1 + 4 * (l + 2)] = b->v[q];
                    uv[2 + 4 * (l + 2)] = b->w[q]; uv[3 + 4 * (l + 2)] = b->t[q];
                }
                duvwt_half_line(n, uv, dmid);
                /* uvwt_lvir: all of -2..n+3, then the flags (sub_tgh.f90:1597-1607) */
                for (int l = -2; l <= n + 3; l++) {
                    long q = G(b, IJK(l));
                    uv[4 * (l + 2)] = b->u[q]; uv[1 + 4 * (l + 2)] = b->v[q];
                    uv[2 + 4 * (l + 2)] = b->w[q]; uv[3 + 4 * (l + 2)] = b->t[q];
                }
                if (b->r[G(b, IJK(-2))] < 0.0) uv[3] = -1.0;
                if (b->r[G(b, IJK(n + 3))] < 0.0) uv[3 + 4 * (n + 5)] = -1.0;
                value_line_half(n, uv, uvh);
                /* cross-direction derivatives: VALUE_HALF_NODE_IJK == VALUE_HALF_NODE per component */
                o_value_half_node(12, nmax, n, dl, dh);
                o_value_half_node(9, nmax, n, kx, kxh);
                o_value_half_node(1, nmax, n, vl, vlh);
                o_value_half_node(1, nmax, n, s1, s1h);
                o_value_half_node(1, nmax, n, s2, s2h);
                for (int f = 0; f <= n; f++)
                    for (int m = 0; m < 4; m++) dh[12 * f + 4 * dir + m] = dmid[m + 4 * f];
                for (int f = 0; f <= n; f++) {
                    double d12[12], nxyz[3];
                    duvwt_dxyz(&dh[12 * f], &kxh[9 * f], vlh[f], d12);
                    nxyz[0] = kxh[9 * f + dir]; nxyz[1] = kxh[9 * f + 3 + dir]; nxyz[2] = kxh[9 * f + 6 + dir];
                    flux_vis(&uvh[4 * f], nxyz, d12, s1h[f], s2h[f], &fv[5 * f]);
                }
                o_flux_dxyz(5, nmax, n, fv, dfv);
                for (int l = 1; l <= n; l++) {
                    long q = G(b, IJK(l));
                    for (int m = 0; m < 5; m++) b->dq[5 * q + m] = b->dq[5 * q + m] - re * dfv[5 * (l - 1) + m];
                }
#undef IJK
            }
        free(uv); free(uvh); free(dmid); free(kx); free(kxh); free(vl); free(vlh);
        free(s1); free(s1h); free(s2); free(s2h); free(dl); free(dh); free(fv); free(dfv);
            }
        }
        free(duvwt);
#undef DN
    }
    return 0;
}

/* LHS.f90:243-343 spectvisl (laminar); LHS.f90:141-241 spectvist (mu + mu_t) */
int o_spectvis(o_case *c, int nb)
{
    o_params prm;
    double visc;
    o_view_params(c, &prm, &visc);
    oblock_view B;
    if (o_view_block(c, nb, &B)) return -1;
    oblock_view *b = &B;
    int ni = b->ni, nj = b->nj, nk = b->nk;
#define PT(i, j, k, ii, jj, kk) do {                                                   \
        long q = G(b, i, j, k), qq = G(b, ii, jj, kk);                                 \
        double rm_vol = b->r[q] * b->vol[qq];                                           \
        double kx = b->kcx[qq], ky = b->kcy[qq], kz = b->kcz[qq];                       \
        double ex = b->etx[qq], ey = b->ety[qq], ez = b->etz[qq];                       \
        double cx = b->ctx[qq], cy = b->cty[qq], cz = b->ctz[qq];                       \
        double ri = kx * kx + ky * ky + kz * kz, rj = ex * ex + ey * ey + ez * ez;      \
        double rk = cx * cx + cy * cy + cz * cz;                                        \
        double vis = prm.nlamtur >= 0 ? b->visl[qq] + b->vist[qq] : b->visl[qq];          \
        double coef = 2.0 * vis / (prm.reynolds * rm_vol + SMALL);                      \
        double coef1 = prm.csrv * coef;                                                  \
        b->srva[q] = coef1 * ri; b->srvb[q] = coef1 * rj; b->srvc[q] = coef1 * rk;       \
    } while (0)
    const int nthr = o_view_threads(c);
    #pragma omp parallel for collapse(2) num_threads(nthr) if (nthr > 1)
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) PT(i, j, k, i, j, k);
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) { PT(0, j, k, 1, j, k); PT(ni + 1, j, k, ni, j, k); }
    for (int k = 1; k <= nk; k++) for (int i = 1; i <= ni; i++) { PT(i, 0, k, i, 1, k); PT(i, nj + 1, k, i, nj, k); }
    for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) { PT(i, j, 0, i, j, 1); PT(i, j, nk + 1, i, j, nk); }
#undef PT
    return 0;
}
