/*
 * wcns_oracle_sa.c -- CPU oracle (TEST INFRASTRUCTURE ONLY): the
 * Spalart-Allmaras one-equation model as the reference runs it inside
 * r_h_s_vis (RHS.f90:22 -> turblence.f90:15-103 turbulence ->
 * differential_model_turbulence) with nameturb = 'SA', ncmpcor = 0, ndes = 0,
 * ntrans = 0, method = 1 (finite differences), restating
 *   turblence.f90:2506-2699  cal_rhs_source_SA
 *   turblence.f90:753-849    dfidxyz_center
 *   turblence.f90:2785-2808  viscosity_equivalent_SA_r
 *   turblence.f90:1022-1185  cal_rhs_viscous_turbulence_dif_r
 *   turblence.f90:218-259    cal_spec_turbulence_r
 *   turblence.f90:291-459    cal_rhs_inviscous_turbulence, NND_turbulence, flux_turbulence
 *   Convect.f90:1281-1286    minmod
 *   turblence.f90:477-583    lusgs_turbulence, mxdq_std_turbulence
 *   turblence.f90:585-623    update_turbulence
 *   turblence.f90:2735-2760  cal_vist_SA
 *   turblence.f90:625-660    max_control_turbulence
 * in the reference's operation order.  x**y with a real exponent is pow(),
 * except x**2.0 and x**3.0, which the flang build of the reference evaluates
 * as x*x and x*x*x (checked against it bit for bit; a gfortran build calls pow
 * and may differ from this by one rounding).
 */
#include "wcns_oracle_priv.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cal_rhs_source_SA model constants (turblence.f90:2509-2513) */
#define SA_CB1 0.1355
#define SA_CB2 0.622
#define SA_CV1 7.1
#define SA_EPS 1.0e-20
#define SA_CGA (2. / 3.0)
#define SA_KI 0.419
#define SA_CW2 0.3
#define SA_CW3 2.0

static double amin1(double a, double b) { return a < b ? a : b; }
static double amax1(double a, double b) { return a > b ? a : b; }

/* Convect.f90:1281-1286 */
static double minmod(double x, double y)
{
    return 0.5 * (copysign(1.0, x) + copysign(1.0, y)) * amin1(fabs(x), fabs(y));
}

/* turblence.f90:753-849 dfidxyz_center (method = 1): central differences at
 * nodes 1..n through the metrics; out[(i,j,k)][3] */
static void dfidxyz_center(const oblock_view *b, const double *fi, double *out)
{
    int ni = b->ni, nj = b->nj;
    for (int k = 1; k <= b->nk; k++)
        for (int j = 1; j <= b->nj; j++)
            for (int i = 1; i <= b->ni; i++) {
                long q = OG(b, i, j, k);
                double dkc = fi[OG(b, i + 1, j, k)] - fi[OG(b, i - 1, j, k)];
                double det = fi[OG(b, i, j + 1, k)] - fi[OG(b, i, j - 1, k)];
                double dct = fi[OG(b, i, j, k + 1)] - fi[OG(b, i, j, k - 1)];
                double ddx = dkc * b->kcx[q] + det * b->etx[q] + dct * b->ctx[q];
                double ddy = dkc * b->kcy[q] + det * b->ety[q] + dct * b->cty[q];
                double ddz = dkc * b->kcz[q] + det * b->etz[q] + dct * b->ctz[q];
                double vol2 = 0.5 / b->vol[q];
                long n = (i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1));
                out[3 * n] = ddx * vol2;
                out[3 * n + 1] = ddy * vol2;
                out[3 * n + 2] = ddz * vol2;
            }
}

/* turblence.f90:437-459 flux_turbulence */
static double flux_turbulence(const double *prim, double nx, double ny, double nz, int npn)
{
    double l1 = nx * prim[1] + ny * prim[2] + nz * prim[3];
    l1 = 0.5 * (l1 + npn * sqrt(l1 * l1));
    return l1 * prim[4] * prim[0];
}

/* turblence.f90:371-435 NND_turbulence (method = 1, cmethd = 2): q(5, 0..n+1)
 * (index l+1), tr(3, 1..n) -> dfc(2..n-1) */
static void nnd_turbulence(double xk, double xb, int n, const double *q, const double *tr, double *dfc, double *fc,
                           double *dq, double *dqcl, double *dqcr)
{
#define Q(m, l) q[5 * ((l) + 1) + (m)]
#define DQ_(m, l) dq[5 * (l) + (m)]
#define CL(m, l) dqcl[5 * (l) + (m)]
#define CR(m, l) dqcr[5 * (l) + (m)]
#define TR(a, l) tr[3 * ((l) - 1) + (a)]
    double c1 = 0.25 * (1.0 + xk), c2 = 0.25 * (1.0 - xk);
    for (int i = 1; i <= n + 1; i++)
        for (int m = 0; m < 5; m++) DQ_(m, i) = Q(m, i) - Q(m, i - 1);
    for (int i = 1; i <= n; i++)
        for (int m = 0; m < 5; m++) {
            CL(m, i) = minmod(xb * DQ_(m, i + 1), DQ_(m, i));
            CR(m, i) = minmod(DQ_(m, i + 1), xb * DQ_(m, i));
        }
    for (int i = 2; i <= n; i++) {
        double nx = 0.50 * (TR(0, i - 1) + TR(0, i));
        double ny = 0.50 * (TR(1, i - 1) + TR(1, i));
        double nz = 0.50 * (TR(2, i - 1) + TR(2, i));
        double ql[5], qr[5];
        for (int m = 0; m < 5; m++) {
            ql[m] = Q(m, i - 1) + c1 * CR(m, i - 1) + c2 * CL(m, i - 1);
            qr[m] = Q(m, i) - c1 * CL(m, i) - c2 * CR(m, i);
        }
        double fl = flux_turbulence(ql, nx, ny, nz, 1);
        double fr = flux_turbulence(qr, nx, ny, nz, -1);
        fc[i] = fl + fr;
    }
    for (int i = 2; i <= n - 1; i++) dfc[i] = fc[i + 1] - fc[i];
#undef Q
#undef DQ_
#undef CL
#undef CR
#undef TR
}

/* turblence.f90:560-583 mxdq_std_turbulence */
static double mxdq_turb(const double *prim, const double *g, double dq, int npn)
{
    double l1 = g[0] * prim[1] + g[1] * prim[2] + g[2] * prim[3];
    l1 = 0.5 * (l1 + npn * fabs(l1));
    return l1 * dq;
}

static int sst_step(o_case *c, int nb);

int o_turbulence(o_case *c, int nb)
{
    o_params prm0;
    double visc0;
    o_view_params(c, &prm0, &visc0);
    if (prm0.nlamtur == 2) return sst_step(c, nb);
    if (prm0.nlamtur != 1) return -1;
    o_params prm;
    double visc;
    o_view_params(c, &prm, &visc);
    oblock_view B;
    if (o_view_block(c, nb, &B)) return -1;
    oblock_view *b = &B;
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const long N = (long)ni * nj * nk;
    const double re = 1.0 / prm.reynolds, reynolds = prm.reynolds, muoo = o_view_muoo(c);
    const double cw1 = SA_CB1 / SA_KI / SA_KI + (1.0 + SA_CB2) / SA_CGA;
    const double cv13 = pow(SA_CV1, 3.0), ki2 = SA_KI * SA_KI, cw36 = pow(SA_CW3, 6.0), onesix = 1.0 / 6.0;
#define N3(i, j, k) ((long)((i) - 1) + (long)ni * ((long)((j) - 1) + (long)nj * ((k) - 1)))
#define N3V(...) N3(__VA_ARGS__)
#define OGV(b, ...) OG(b, __VA_ARGS__)
    double *dqke = (double *)calloc(b->g3, sizeof(double));   /* -1..n+1 used */
    double *spec = (double *)calloc(b->g3, sizeof(double));
    double *dudxyz = (double *)calloc(3 * N, sizeof(double)), *dvdxyz = (double *)calloc(3 * N, sizeof(double));
    double *dwdxyz = (double *)calloc(3 * N, sizeof(double)), *dmudxyz = (double *)calloc(3 * N, sizeof(double));
    double *viseq = (double *)calloc(N, sizeof(double));

    /* cal_rhs_source_SA */
    dfidxyz_center(b, b->u, dudxyz);
    dfidxyz_center(b, b->v, dvdxyz);
    dfidxyz_center(b, b->w, dwdxyz);
    dfidxyz_center(b, b->qke, dmudxyz);
    const double ft2 = 0.0;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                long q = OG(b, i, j, k), n = N3(i, j, k);
                double ro = b->r[q], volume = b->vol[q], vsl = b->visl[q], vtbar = b->qke[q];
                double romu = ro * vtbar;
                double dn = amax1(b->dist[q], re);
                double dn2 = dn * dn;
                double reki2dn2 = reynolds * ki2 * dn2;
                double kafan = romu / vsl;
                double kf3 = kafan * kafan * kafan;
                double fv1 = kf3 / (kf3 + cv13);
                double fv2 = 1.0 - kafan / (1.0 + kafan * fv1);
                double a1 = dudxyz[3 * n + 1] - dvdxyz[3 * n], a2 = dvdxyz[3 * n + 2] - dwdxyz[3 * n + 1];
                double a3 = dwdxyz[3 * n] - dudxyz[3 * n + 2];
                double omg = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
                double sbar = omg + vtbar / reki2dn2 * fv2;
                sbar = amax1(sbar, SA_EPS);
                double rbar = vtbar / (sbar * reki2dn2);
                rbar = amin1(rbar, 10.0);
                double rbar5 = pow(rbar, 5.0);
                double gbar = rbar + SA_CW2 * rbar * (rbar5 - 1.0);
                double gbar6 = pow(gbar, 6.0);
                double fw = gbar * pow((1.0 + cw36) / (gbar6 + cw36), onesix);
                double cpmu = SA_CB1 * (1.0 - ft2) * sbar;
                double cdmu = (cw1 * fw - SA_CB1 * ft2 / ki2) * vtbar / dn2 / reynolds;
                double Sp = cpmu * vtbar, Sd = cdmu * vtbar;
                double dmudx = dmudxyz[3 * n], dmudy = dmudxyz[3 * n + 1], dmudz = dmudxyz[3 * n + 2];
                double Dbar = SA_CB2 / SA_CGA * (dmudx * dmudx + dmudy * dmudy + dmudz * dmudz) * re;
                dqke[q] = dqke[q] - ro * (Sp - Sd + Dbar) * volume;
                spec[q] = spec[q] - amin1(cpmu - 2.0 * cdmu, 0.0) * volume;
            }
    /* viscosity_equivalent_SA_r on 1..n */
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        long q = OG(b, i, j, k);
        viseq[N3(i, j, k)] = 1.5 * (b->visl[q] / b->r[q] + b->qke[q]);
    }

    /* cal_rhs_viscous_turbulence_dif_r: faces between l-1 and l, l = 2..n; nodes 2..n-1 */
    int nmax = ni > nj ? ni : nj;
    nmax = nmax > nk ? nmax : nk;
    double *fv = (double *)calloc(nmax + 2, sizeof(double));
    for (int dir = 0; dir < 3; dir++) {
        int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
        int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
        const double *mx = dir == 0 ? b->kcx : (dir == 1 ? b->etx : b->ctx);
        const double *my = dir == 0 ? b->kcy : (dir == 1 ? b->ety : b->cty);
        const double *mz = dir == 0 ? b->kcz : (dir == 1 ? b->etz : b->ctz);
        for (int b2 = 2; b2 <= nbb - 1; b2++)
            for (int b1 = 2; b1 <= na - 1; b1++) {
#define IJK(l) (dir == 0 ? (l) : b1), (dir == 1 ? (l) : (dir == 0 ? b1 : b2)), (dir == 2 ? (l) : b2)
                for (int l = 2; l <= n; l++) {
                    long qa = OGV(b, IJK(l)), qb = OGV(b, IJK(l - 1));
                    long na_ = N3V(IJK(l)), nb_ = N3V(IJK(l - 1));
                    double nx = 0.5 * (mx[qa] + mx[qb]), ny = 0.5 * (my[qa] + my[qb]), nz = 0.5 * (mz[qa] + mz[qb]);
                    double dmudx = dmudxyz[3 * na_] + dmudxyz[3 * nb_];
                    double dmudy = dmudxyz[3 * na_ + 1] + dmudxyz[3 * nb_ + 1];
                    double dmudz = dmudxyz[3 * na_ + 2] + dmudxyz[3 * nb_ + 2];
                    double vis = viseq[na_] + viseq[nb_];
                    fv[l] = (dmudx * nx + dmudy * ny + dmudz * nz) * vis * 0.25;
                }
                for (int l = 2; l <= n - 1; l++) {
                    long q = OGV(b, IJK(l));
                    dqke[q] = dqke[q] - re * b->r[q] * (fv[l + 1] - fv[l]);
                }
            }
    }

    /* cal_spec_turbulence_r on nodes 2..n-1 */
    const double re2 = re * 2.0;
    for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
        long q = OG(b, i, j, k);
        double um = b->u[q], vm = b->v[q], wm = b->w[q];
        double unx = um * b->kcx[q] + vm * b->kcy[q] + wm * b->kcz[q];
        double uny = um * b->etx[q] + vm * b->ety[q] + wm * b->etz[q];
        double unz = um * b->ctx[q] + vm * b->cty[q] + wm * b->ctz[q];
        spec[q] = spec[q] + fabs(unx) + fabs(uny) + fabs(unz) + 1.0 / (b->dtdt[q] * prm.timedt_turb);
        double nx = b->kcx[q] * b->kcx[q] + b->kcy[q] * b->kcy[q] + b->kcz[q] * b->kcz[q];
        double ny = b->etx[q] * b->etx[q] + b->ety[q] * b->ety[q] + b->etz[q] * b->etz[q];
        double nz = b->ctx[q] * b->ctx[q] + b->cty[q] * b->cty[q] + b->ctz[q] * b->ctz[q];
        double vis = viseq[N3(i, j, k)] / b->vol[q];
        spec[q] = spec[q] + re2 * vis * (nx + ny + nz);
    }

    /* cal_rhs_inviscous_turbulence: NND convection per pencil, rows 2..n-1 */
    double *ql = (double *)calloc(5 * (nmax + 3), sizeof(double)), *tr = (double *)calloc(3 * (nmax + 1), sizeof(double));
    double *dfc = (double *)calloc(nmax + 2, sizeof(double)), *fc = (double *)calloc(nmax + 2, sizeof(double));
    double *dqw = (double *)calloc(5 * (nmax + 3), sizeof(double)), *clw = (double *)calloc(5 * (nmax + 2), sizeof(double));
    double *crw = (double *)calloc(5 * (nmax + 2), sizeof(double));
    for (int dir = 0; dir < 3; dir++) {
        int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
        int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
        const double *mx = dir == 0 ? b->kcx : (dir == 1 ? b->etx : b->ctx);
        const double *my = dir == 0 ? b->kcy : (dir == 1 ? b->ety : b->cty);
        const double *mz = dir == 0 ? b->kcz : (dir == 1 ? b->etz : b->ctz);
        for (int b2 = 2; b2 <= nbb - 1; b2++)
            for (int b1 = 2; b1 <= na - 1; b1++) {
                for (int l = 0; l <= n + 1; l++) {
                    long q = OGV(b, IJK(l));
                    double *e = ql + 5 * (l + 1);
                    e[0] = b->r[q]; e[1] = b->u[q]; e[2] = b->v[q]; e[3] = b->w[q]; e[4] = b->qke[q];
                }
                for (int l = 1; l <= n; l++) {
                    long q = OGV(b, IJK(l));
                    tr[3 * (l - 1)] = mx[q]; tr[3 * (l - 1) + 1] = my[q]; tr[3 * (l - 1) + 2] = mz[q];
                }
                nnd_turbulence(prm.xk, prm.xb, n, ql, tr, dfc, fc, dqw, clw, crw);
                for (int l = 2; l <= n - 1; l++) {
                    long q = OGV(b, IJK(l));
                    dqke[q] = dqke[q] + dfc[l];
                }
            }
    }
#undef IJK

    /* lusgs_turbulence: forward then backward over nodes 2..n-1 */
    for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
        long q = OG(b, i, j, k);
        long qi = OG(b, i - 1, j, k), qj = OG(b, i, j - 1, k), qk = OG(b, i, j, k - 1);
        double pi[5] = {b->r[qi], b->u[qi], b->v[qi], b->w[qi], b->p[qi]};
        double pj[5] = {b->r[qj], b->u[qj], b->v[qj], b->w[qj], b->p[qj]};
        double pk[5] = {b->r[qk], b->u[qk], b->v[qk], b->w[qk], b->p[qk]};
        double gi[3] = {b->kcx[qi], b->kcy[qi], b->kcz[qi]};
        double gj[3] = {b->etx[qj], b->ety[qj], b->etz[qj]};
        double gk[3] = {b->ctx[qk], b->cty[qk], b->ctz[qk]};
        double de = mxdq_turb(pi, gi, dqke[qi], 1), df = mxdq_turb(pj, gj, dqke[qj], 1);
        double dg = mxdq_turb(pk, gk, dqke[qk], 1);
        double rhs0 = de + df + dg;
        dqke[q] = (rhs0 - dqke[q]) / spec[q];
    }
    for (int k = nk - 1; k >= 2; k--) for (int j = nj - 1; j >= 2; j--) for (int i = ni - 1; i >= 2; i--) {
        long q = OG(b, i, j, k);
        long qi = OG(b, i + 1, j, k), qj = OG(b, i, j + 1, k), qk = OG(b, i, j, k + 1);
        double pi[5] = {b->r[qi], b->u[qi], b->v[qi], b->w[qi], b->p[qi]};
        double pj[5] = {b->r[qj], b->u[qj], b->v[qj], b->w[qj], b->p[qj]};
        double pk[5] = {b->r[qk], b->u[qk], b->v[qk], b->w[qk], b->p[qk]};
        double gi[3] = {b->kcx[qi], b->kcy[qi], b->kcz[qi]};
        double gj[3] = {b->etx[qj], b->ety[qj], b->etz[qj]};
        double gk[3] = {b->ctx[qk], b->cty[qk], b->ctz[qk]};
        double de = mxdq_turb(pi, gi, dqke[qi], -1), df = mxdq_turb(pj, gj, dqke[qj], -1);
        double dg = mxdq_turb(pk, gk, dqke[qk], -1);
        double rhs0 = de + df + dg;
        dqke[q] = dqke[q] - rhs0 / spec[q];
    }

    /* update_turbulence (nodes 2..n-1), cal_vist_SA (1..n) */
    const double minlimiter = 1.0e-20;
    for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
        long q = OG(b, i, j, k);
        b->qke[q] = amax1(b->qke[q] + dqke[q] / b->r[q], minlimiter * muoo);
    }
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        long q = OG(b, i, j, k);
        double romu = b->r[q] * b->qke[q];
        double kafan = romu / b->visl[q];
        double kf3 = kafan * kafan * kafan;
        double fv1 = kf3 / (kf3 + cv13);
        b->vist[q] = romu * fv1;
    }
    /* max_control_turbulence: nodes 0..n (method - 1 .. n + method - 1) */
    if (prm.kmaxlim > 1.0e-15)
        for (int k = 0; k <= nk; k++) for (int j = 0; j <= nj; j++) for (int i = 0; i <= ni; i++) {
            long q = OG(b, i, j, k);
            b->vist[q] = amin1(b->vist[q], prm.kmaxlim * b->visl[q]);
        }
#undef N3
#undef N3V
#undef OGV
    free(dqke); free(spec); free(dudxyz); free(dvdxyz); free(dwdxyz); free(dmudxyz); free(viseq);
    free(fv); free(ql); free(tr); free(dfc); free(fc); free(dqw); free(clw); free(crw);
    return 0;
}

/* ===================================================================== SST */
/* cal_rhs_source_SST constants (turblence.f90:2813-2818) */
static const double SST_CGAK[2] = {0.85, 1.0}, SST_CGAW[2] = {0.5, 0.856}, SST_BETA[2] = {0.075, 0.0828};
#define SST_CDMIN 1.0e-20
#define SST_KI 0.41

/* The SST step of one block (turblence.f90:15-103 with nameturb 'SST', nlamtur 2,
 * ncmpcor 0, ndes 0): cal_rhs_source_SST (:2810-3089), cal_rhs_viscous_turbulence_dif
 * (:851-1019), cal_spec_turbulence (:175-215), cal_rhs_inviscous_turbulence per
 * variable, lusgs_turbulence per variable, update_turbulence with cal_vist_SST
 * (:3119-3164), max_control_turbulence. */
static int sst_step(o_case *c, int nb)
{
    o_params prm;
    double visc;
    o_view_params(c, &prm, &visc);
    oblock_view B;
    if (o_view_block(c, nb, &B)) return -1;
    oblock_view *b = &B;
    const int ni = b->ni, nj = b->nj, nk = b->nk;
    const long N = (long)ni * nj * nk;
    const double re = 1.0 / prm.reynolds, reynolds = prm.reynolds;
    double koo, omgoo;
    o_view_turb_inflow(c, &koo, &omgoo);
    const double tinf[2] = {koo, omgoo};
    double *q2[2] = {b->qke, b->qkw};
    const double ki2 = SST_KI * SST_KI, twothd = 0.2 / 0.3;
    const double betax0 = 0.09, betax = betax0, obetx = 1.0 / betax, ki2osqrtbetx = ki2 / sqrt(betax);
#define N3(i, j, k) ((long)((i) - 1) + (long)ni * ((long)((j) - 1) + (long)nj * ((k) - 1)))
#define N3V(...) N3(__VA_ARGS__)
#define OGV(b, ...) OG(b, __VA_ARGS__)
    double *dqke[2], *spec[2], *dmu[2], *viseq[2];
    for (int m = 0; m < 2; m++) {
        dqke[m] = (double *)calloc(b->g3, sizeof(double));
        spec[m] = (double *)calloc(b->g3, sizeof(double));
        dmu[m] = (double *)calloc(3 * N, sizeof(double));
        viseq[m] = (double *)calloc(N, sizeof(double));
    }
    double *dudxyz = (double *)calloc(3 * N, sizeof(double)), *dvdxyz = (double *)calloc(3 * N, sizeof(double));
    double *dwdxyz = (double *)calloc(3 * N, sizeof(double)), *F1 = (double *)calloc(N, sizeof(double));

    /* cal_rhs_source_SST */
    dfidxyz_center(b, b->u, dudxyz);
    dfidxyz_center(b, b->v, dvdxyz);
    dfidxyz_center(b, b->w, dwdxyz);
    dfidxyz_center(b, b->qke, dmu[0]);
    dfidxyz_center(b, b->qkw, dmu[1]);
    const double re500 = re * 500.0, cgaw24 = SST_CGAW[1] * 4.0, cgaw22 = SST_CGAW[1] * 2.0;
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        long q = OG(b, i, j, k), n = N3(i, j, k);
        double ro = b->r[q], qk = b->qke[q], qw = b->qkw[q], vsl = b->visl[q], vst = b->vist[q];
        double rok = ro * qk, row = ro * qw;
        double dn = b->dist[q], dn2 = dn * dn;
        double dkxdw = dmu[0][3 * n] * dmu[1][3 * n] + dmu[0][3 * n + 1] * dmu[1][3 * n + 1] +
                       dmu[0][3 * n + 2] * dmu[1][3 * n + 2];
        double rowdkdw = ro / qw * dkxdw * cgaw22;
        double CDkw = amax1(SST_CDMIN, rowdkdw);
        double sqrtkwdn = obetx * sqrt(qk) / qw / dn;
        double vlrodn2w = re500 * vsl / row / dn2;
        double rokdn2 = cgaw24 * rok / CDkw / dn2;
        double arg2 = amax1(sqrtkwdn, vlrodn2w);
        double arg1 = amin1(arg2, rokdn2);
        F1[n] = arg1 >= 2.0 ? 1.0 : tanh(pow(arg1, 4.0));
        double F2 = 1.0 - F1[n];
        double cga0 = F1[n] * SST_CGAK[0] + F2 * SST_CGAK[1];
        double cga1 = F1[n] * SST_CGAW[0] + F2 * SST_CGAW[1];
        viseq[0][n] = vsl + cga0 * vst;
        viseq[1][n] = vsl + cga1 * vst;
    }
    for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
        long q = OG(b, i, j, k), n = N3(i, j, k);
        double ro = b->r[q], qk = b->qke[q], qw = b->qkw[q], volume = b->vol[q], vst = b->vist[q];
        double rok = ro * qk, row = ro * qw;
        double F2 = 1.0 - F1[n];
        double beita = F1[n] * SST_BETA[0] + F2 * SST_BETA[1];
        double cga1 = F1[n] * SST_CGAW[0] + F2 * SST_CGAW[1];
        double ga = beita * obetx - cga1 * ki2osqrtbetx;
        double ux = dudxyz[3 * n], uy = dudxyz[3 * n + 1], uz = dudxyz[3 * n + 2];
        double vx = dvdxyz[3 * n], vy = dvdxyz[3 * n + 1], vz = dvdxyz[3 * n + 2];
        double wx = dwdxyz[3 * n], wy = dwdxyz[3 * n + 1], wz = dwdxyz[3 * n + 2];
        double div_u = ux + vy + wz;
        double dkxdw = dmu[0][3 * n] * dmu[1][3 * n] + dmu[0][3 * n + 1] * dmu[1][3 * n + 1] +
                       dmu[0][3 * n + 2] * dmu[1][3 * n + 2];
        double omg = vst * ((2.0 * (ux * ux + vy * vy + wz * wz + uy * vx + uz * wx + wy * vz) +
                             uy * uy + uz * uz + vx * vx + vz * vz + wx * wx + wy * wy) -
                            twothd * div_u * div_u) * re -
                     twothd * div_u * rok;
        double cdk = betax * qw, cdw = beita * qw;
        double Sd0 = cdk * rok, Sd1 = cdw * row;
        double Sp0 = amin1(omg, 20.0 * Sd0);
        double Sp1 = ga * ro * omg / vst * reynolds;
        double dbw = cgaw22 * F2 * dkxdw / qw / qw;
        double Dbar0 = 0.0, Dbar1 = dbw * row;
        spec[0][q] = spec[0][q] + 2.0 * (cdk + twothd * amax1(ux + vy + wz, 0.0)) * volume;
        spec[1][q] = spec[1][q] + 2.0 * (2.0 * cdw + fabs(dbw)) * volume;
        dqke[0][q] = dqke[0][q] - (Sp0 - Sd0 + Dbar0) * volume;
        dqke[1][q] = dqke[1][q] - (Sp1 - Sd1 + Dbar1) * volume;
    }

    /* cal_rhs_viscous_turbulence_dif: faces between l-1 and l, l = 2..n; nodes 2..n-1 */
    int nmax = ni > nj ? ni : nj;
    nmax = nmax > nk ? nmax : nk;
    double *fv = (double *)calloc(2 * (nmax + 2), sizeof(double));
    for (int dir = 0; dir < 3; dir++) {
        int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
        int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
        const double *mx = dir == 0 ? b->kcx : (dir == 1 ? b->etx : b->ctx);
        const double *my = dir == 0 ? b->kcy : (dir == 1 ? b->ety : b->cty);
        const double *mz = dir == 0 ? b->kcz : (dir == 1 ? b->etz : b->ctz);
        for (int b2 = 2; b2 <= nbb - 1; b2++)
            for (int b1 = 2; b1 <= na - 1; b1++) {
#define IJK(l) (dir == 0 ? (l) : b1), (dir == 1 ? (l) : (dir == 0 ? b1 : b2)), (dir == 2 ? (l) : b2)
                for (int l = 2; l <= n; l++) {
                    long qa = OGV(b, IJK(l)), qb = OGV(b, IJK(l - 1));
                    long na_ = N3V(IJK(l)), nb_ = N3V(IJK(l - 1));
                    double nx = 0.5 * (mx[qa] + mx[qb]), ny = 0.5 * (my[qa] + my[qb]), nz = 0.5 * (mz[qa] + mz[qb]);
                    for (int m = 0; m < 2; m++) {
                        double dmudx = dmu[m][3 * na_] + dmu[m][3 * nb_];
                        double dmudy = dmu[m][3 * na_ + 1] + dmu[m][3 * nb_ + 1];
                        double dmudz = dmu[m][3 * na_ + 2] + dmu[m][3 * nb_ + 2];
                        double vis = viseq[m][na_] + viseq[m][nb_];
                        fv[2 * l + m] = (dmudx * nx + dmudy * ny + dmudz * nz) * vis * 0.25;
                    }
                }
                for (int m = 0; m < 2; m++)
                    for (int l = 2; l <= n - 1; l++) {
                        long q = OGV(b, IJK(l));
                        dqke[m][q] = dqke[m][q] - re * (fv[2 * (l + 1) + m] - fv[2 * l + m]);
                    }
            }
    }

    /* cal_spec_turbulence on nodes 2..n-1 */
    const double re2 = re * 2.0;
    for (int m = 0; m < 2; m++)
        for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
            long q = OG(b, i, j, k);
            double um = b->u[q], vm = b->v[q], wm = b->w[q];
            double unx = um * b->kcx[q] + vm * b->kcy[q] + wm * b->kcz[q];
            double uny = um * b->etx[q] + vm * b->ety[q] + wm * b->etz[q];
            double unz = um * b->ctx[q] + vm * b->cty[q] + wm * b->ctz[q];
            spec[m][q] = spec[m][q] + fabs(unx) + fabs(uny) + fabs(unz) + 1.0 / (b->dtdt[q] * prm.timedt_turb);
            double nx = b->kcx[q] * b->kcx[q] + b->kcy[q] * b->kcy[q] + b->kcz[q] * b->kcz[q];
            double ny = b->etx[q] * b->etx[q] + b->ety[q] * b->ety[q] + b->etz[q] * b->etz[q];
            double nz = b->ctx[q] * b->ctx[q] + b->cty[q] * b->cty[q] + b->ctz[q] * b->ctz[q];
            double vis = viseq[m][N3(i, j, k)] / b->vol[q] / b->r[q];
            spec[m][q] = spec[m][q] + re2 * vis * (nx + ny + nz);
        }

    /* cal_rhs_inviscous_turbulence, variable by variable */
    double *ql = (double *)calloc(5 * (nmax + 3), sizeof(double)), *tr = (double *)calloc(3 * (nmax + 1), sizeof(double));
    double *dfc = (double *)calloc(nmax + 2, sizeof(double)), *fc = (double *)calloc(nmax + 2, sizeof(double));
    double *dqw = (double *)calloc(5 * (nmax + 3), sizeof(double)), *clw = (double *)calloc(5 * (nmax + 2), sizeof(double));
    double *crw = (double *)calloc(5 * (nmax + 2), sizeof(double));
    for (int m = 0; m < 2; m++)
        for (int dir = 0; dir < 3; dir++) {
            int n = dir == 0 ? ni : (dir == 1 ? nj : nk);
            int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
            const double *mx = dir == 0 ? b->kcx : (dir == 1 ? b->etx : b->ctx);
            const double *my = dir == 0 ? b->kcy : (dir == 1 ? b->ety : b->cty);
            const double *mz = dir == 0 ? b->kcz : (dir == 1 ? b->etz : b->ctz);
            for (int b2 = 2; b2 <= nbb - 1; b2++)
                for (int b1 = 2; b1 <= na - 1; b1++) {
                    for (int l = 0; l <= n + 1; l++) {
                        long q = OGV(b, IJK(l));
                        double *e = ql + 5 * (l + 1);
                        e[0] = b->r[q]; e[1] = b->u[q]; e[2] = b->v[q]; e[3] = b->w[q]; e[4] = q2[m][q];
                    }
                    for (int l = 1; l <= n; l++) {
                        long q = OGV(b, IJK(l));
                        tr[3 * (l - 1)] = mx[q]; tr[3 * (l - 1) + 1] = my[q]; tr[3 * (l - 1) + 2] = mz[q];
                    }
                    nnd_turbulence(prm.xk, prm.xb, n, ql, tr, dfc, fc, dqw, clw, crw);
                    for (int l = 2; l <= n - 1; l++) {
                        long q = OGV(b, IJK(l));
                        dqke[m][q] = dqke[m][q] + dfc[l];
                    }
                }
        }
#undef IJK

    /* lusgs_turbulence, variable by variable: forward then backward over nodes 2..n-1 */
    for (int m = 0; m < 2; m++) {
        double *dq = dqke[m], *sp = spec[m];
        for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
            long q = OG(b, i, j, k);
            long qi = OG(b, i - 1, j, k), qj = OG(b, i, j - 1, k), qk = OG(b, i, j, k - 1);
            double pi[5] = {b->r[qi], b->u[qi], b->v[qi], b->w[qi], b->p[qi]};
            double pj[5] = {b->r[qj], b->u[qj], b->v[qj], b->w[qj], b->p[qj]};
            double pk[5] = {b->r[qk], b->u[qk], b->v[qk], b->w[qk], b->p[qk]};
            double gi[3] = {b->kcx[qi], b->kcy[qi], b->kcz[qi]};
            double gj[3] = {b->etx[qj], b->ety[qj], b->etz[qj]};
            double gk[3] = {b->ctx[qk], b->cty[qk], b->ctz[qk]};
            double de = mxdq_turb(pi, gi, dq[qi], 1), df = mxdq_turb(pj, gj, dq[qj], 1);
            double dg = mxdq_turb(pk, gk, dq[qk], 1);
            double rhs0 = de + df + dg;
            dq[q] = (rhs0 - dq[q]) / sp[q];
        }
        for (int k = nk - 1; k >= 2; k--) for (int j = nj - 1; j >= 2; j--) for (int i = ni - 1; i >= 2; i--) {
            long q = OG(b, i, j, k);
            long qi = OG(b, i + 1, j, k), qj = OG(b, i, j + 1, k), qk = OG(b, i, j, k + 1);
            double pi[5] = {b->r[qi], b->u[qi], b->v[qi], b->w[qi], b->p[qi]};
            double pj[5] = {b->r[qj], b->u[qj], b->v[qj], b->w[qj], b->p[qj]};
            double pk[5] = {b->r[qk], b->u[qk], b->v[qk], b->w[qk], b->p[qk]};
            double gi[3] = {b->kcx[qi], b->kcy[qi], b->kcz[qi]};
            double gj[3] = {b->etx[qj], b->ety[qj], b->etz[qj]};
            double gk[3] = {b->ctx[qk], b->cty[qk], b->ctz[qk]};
            double de = mxdq_turb(pi, gi, dq[qi], -1), df = mxdq_turb(pj, gj, dq[qj], -1);
            double dg = mxdq_turb(pk, gk, dq[qk], -1);
            double rhs0 = de + df + dg;
            dq[q] = dq[q] - rhs0 / sp[q];
        }
    }

    /* update_turbulence (nodes 2..n-1), cal_vist_SST (1..n) */
    const double minlimiter = 1.0e-20;
    for (int m = 0; m < 2; m++)
        for (int k = 2; k <= nk - 1; k++) for (int j = 2; j <= nj - 1; j++) for (int i = 2; i <= ni - 1; i++) {
            long q = OG(b, i, j, k);
            q2[m][q] = amax1(q2[m][q] + dqke[m][q] / b->r[q], minlimiter * tinf[m]);
        }
    {
        const double a1 = 0.31, obetx_v = 2.0 / 0.09;
        const double a1reynolds = a1 * reynolds, b1 = 500.0 * re;
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
            long q = OG(b, i, j, k), n = N3(i, j, k);
            double ro = b->r[q], qk = b->qke[q], qw = b->qkw[q], vsl = b->visl[q];
            double dn = b->dist[q], dn2 = dn * dn;
            double sqrtkwdn = obetx_v * sqrt(qk) / qw / dn;
            double vlrodn2w = b1 * vsl / ro / dn2 / qw;
            double arg2 = amax1(sqrtkwdn, vlrodn2w);
            double F2 = arg2 >= 4.0 ? 1.0 : tanh(arg2 * arg2);
            double a_ = dudxyz[3 * n + 1] - dvdxyz[3 * n], b_ = dvdxyz[3 * n + 2] - dwdxyz[3 * n + 1];
            double c_ = dwdxyz[3 * n] - dudxyz[3 * n + 2];
            double omg = sqrt(a_ * a_ + b_ * b_ + c_ * c_);
            b->vist[q] = a1reynolds * ro * qk / amax1(a1 * qw, omg * F2);
        }
    }
    /* max_control_turbulence: nodes 0..n */
    if (prm.kmaxlim > 1.0e-15)
        for (int k = 0; k <= nk; k++) for (int j = 0; j <= nj; j++) for (int i = 0; i <= ni; i++) {
            long q = OG(b, i, j, k);
            b->vist[q] = amin1(b->vist[q], prm.kmaxlim * b->visl[q]);
        }
#undef N3
#undef N3V
#undef OGV
    for (int m = 0; m < 2; m++) { free(dqke[m]); free(spec[m]); free(dmu[m]); free(viseq[m]); }
    free(dudxyz); free(dvdxyz); free(dwdxyz); free(F1);
    free(fv); free(ql); free(tr); free(dfc); free(fc); free(dqw); free(clw); free(crw);
    return 0;
}
