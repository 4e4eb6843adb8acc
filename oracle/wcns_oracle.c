/*
 * wcns_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the WCNS
 * pseudo-time-step hot path.  A plain-C restatement of the reference Fortran
 * routines, each function citing the reference file:line it follows
 * (paths relative to the reference's src/).  Build with -O2 -ffp-contract=off
 * so that every product/sum is rounded exactly as the reference's -O2 build on
 * x86-64 (no FMA in the baseline ISA).
 *
 * This file is the checker: the GPU library never links or calls it.
 */
#include "wcns_oracle_priv.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* Globalvar.f90:215-229 */
#define SMALL      1.0e-38
#define SML_SSS    1.0e-30
#define SML_VOL    1.0e-30
#define PMIN_LIMIT 1.0e-8
#define PMAX_LIMIT 1.0e+3
#define RMIN_LIMIT 1.0e-8
#define RMAX_LIMIT 1.0e+3
#define TMIN_LIMIT 1.0e-8
#define PAI        3.14159265      /* Globalvar.f90:236, truncated on purpose */
#define NIJK2ND    4               /* Main.f90:20-21 */
#define NIJK2D     4
#define NL 5

#define MAXREG 64

typedef struct {
    int bctype;
    int raw_s[6], raw_t[6];
    int s_st[3], s_ed[3], s_lr3d[3], s_nd, s_lr;
    int nbt, ibcwin;                       /* 1-based, as in the .inp file */
    int t_st[3], t_ed[3], t_lr3d[3], t_nd, t_lr;
    int *image, *jmage, *kmage;            /* over the face box */
    int bi, bj, bk;                        /* box extents */
    int iused;
} oregion;

typedef struct {
    int ni, nj, nk;
    long g3;                               /* number of G3 nodes */
    double *x, *y, *z;
    double *r, *u, *v, *w, *p, *t, *c;
    double *q, *dq, *rhs_nb;
    double *kcx, *kcy, *kcz, *etx, *ety, *etz, *ctx, *cty, *ctz, *vol;
    double *sra, *srb, *src, *srva, *srvb, *srvc, *dtdt, *visl, *vist;
    double *qke, *dist;                    /* SA nu~ / SST k (G3), wall distance (G3 shape, 1..n used) */
    double *qkw;                           /* SST omega (G3) */
    int nregions;
    oregion reg[MAXREG];
    int bcindex[MAXREG];
} oblock;

typedef struct { int nb, i, j, k; } opnt;
typedef struct { int n, cap; opnt *p; } ogroup;

struct o_case {
    o_params prm;
    double roo, uoo, voo, woo, poo, coo, too, eoo, visc, re, attack_r, sideslip_r;
    double muoo;                           /* SA freestream value, turblence.f90:3166-3181 */
    double koo, omgoo, vtoo;               /* SST freestream k, omega and mu_t, same place */
    int nblocks, cap;
    oblock *blk;
    int ngroups;          /* groups with more than two points (multi-points) */
    ogroup *groups;
    long n_fallback, n_repair;
    int nthreads;         /* OpenMP threads of the pencil, node and hyperplane loops (1: serial) */
};

/* ---------------------------------------------------------------- indexing */
static inline long G(const oblock *b, int i, int j, int k)
{
    return (long)(i + 2) + (long)(b->ni + 6) * ((long)(j + 2) + (long)(b->nj + 6) * (long)(k + 2));
}
#define R_(b,i,j,k)   (b)->r[G(b,i,j,k)]
#define U_(b,i,j,k)   (b)->u[G(b,i,j,k)]
#define V_(b,i,j,k)   (b)->v[G(b,i,j,k)]
#define W_(b,i,j,k)   (b)->w[G(b,i,j,k)]
#define P_(b,i,j,k)   (b)->p[G(b,i,j,k)]
#define T_(b,i,j,k)   (b)->t[G(b,i,j,k)]
#define DQ(b,m,i,j,k) (b)->dq[5 * G(b,i,j,k) + (m) - 1]
#define Q_(b,m,i,j,k) (b)->q[5 * G(b,i,j,k) + (m) - 1]
#define F3(b,a,i,j,k) (b)->a[G(b,i,j,k)]

/* ============================================================ gas model */
/* GasModel.f90:358-373 getem */
static inline double getem(const double *prim, double gama)
{
    double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    return pm / (gama - 1.0) + 0.5 * rm * (um * um + vm * vm + wm * wm);
}
/* GasModel.f90:385-399 gethmgm_ns: hm = gama/(gama-1)*pm/rm */
static inline double gethm(const double *prim, double gama)
{
    return gama / (gama - 1.0) * prim[4] / prim[0];
}

/* ================================================== pencil-level routines */
/* WCNSE5.f90:799-852 VALUE_HALF_NODE; q(n,nmax) -> q_half(n,0:nmax) */
void o_value_half_node(int n, int nmax, int ni, const double *q, double *q_half)
{
    (void)nmax;
#define QQ(m,i)  q[((i) - 1) * n + (m) - 1]
#define QH(m,i)  q_half[(i) * n + (m) - 1]
    const double A1 = 9.0, B1 = -1.0, A2 = 5.0, B2 = 15.0, C2 = -5.0, D2 = 1.0;
    if (ni <= NIJK2ND) {
        for (int i = 1; i <= ni - 1; i++)
            for (int m = 1; m <= n; m++)
                QH(m, i) = 0.5 * (QQ(m, i) + QQ(m, i + 1));
        for (int m = 1; m <= n; m++) {
            int i2 = ni < 2 ? ni : 2, i3 = ni < 3 ? ni : 3;
            int j1 = ni - 1 > 1 ? ni - 1 : 1, j2 = ni - 2 > 1 ? ni - 2 : 1;
            QH(m, 0)  = (15.0 * QQ(m, 1) - 10.0 * QQ(m, i2) + 3.0 * QQ(m, i3)) / 8.0;
            QH(m, ni) = (15.0 * QQ(m, ni) - 10.0 * QQ(m, j1) + 3.0 * QQ(m, j2)) / 8.0;
        }
    } else {
        for (int i = 2; i <= ni - 2; i++)
            for (int m = 1; m <= n; m++)
                QH(m, i) = (A1 * (QQ(m, i) + QQ(m, i + 1)) + B1 * (QQ(m, i + 2) + QQ(m, i - 1))) / 16.0;
        for (int m = 1; m <= n; m++) {
            QH(m, 1)      = (A2 * QQ(m, 1) + B2 * QQ(m, 2) + C2 * QQ(m, 3) + D2 * QQ(m, 4)) / 16.0;
            QH(m, ni - 1) = (A2 * QQ(m, ni) + B2 * QQ(m, ni - 1) + C2 * QQ(m, ni - 2) + D2 * QQ(m, ni - 3)) / 16.0;
            QH(m, 0)      = (35.0 * QQ(m, 1) - 35.0 * QQ(m, 2) + 21.0 * QQ(m, 3) - 5.0 * QQ(m, 4)) / 16.0;
            QH(m, ni)     = (35.0 * QQ(m, ni) - 35.0 * QQ(m, ni - 1) + 21.0 * QQ(m, ni - 2) - 5.0 * QQ(m, ni - 3)) / 16.0;
        }
    }
#undef QQ
#undef QH
}

/* WCNSE5.f90:209-254 FLUX_DXYZ; f(nl,0:nmax) -> df(nl,nmax) */
void o_flux_dxyz(int nl, int nmax, int ni, const double *f, double *df)
{
    (void)nmax;
#define FF(m,i)  f[(i) * nl + (m) - 1]
#define DF(m,i)  df[((i) - 1) * nl + (m) - 1]
    if (ni <= NIJK2ND) {
        for (int i = 1; i <= ni; i++)
            for (int m = 1; m <= nl; m++)
                DF(m, i) = FF(m, i) - FF(m, i - 1);
    } else {
        for (int i = 3; i <= ni - 2; i++)
            for (int m = 1; m <= nl; m++)
                DF(m, i) = (2250.0 * (FF(m, i) - FF(m, i - 1)) - 125.0 * (FF(m, i + 1) - FF(m, i - 2))
                            + 9.0 * (FF(m, i + 2) - FF(m, i - 3))) / 1920.0;
        for (int m = 1; m <= nl; m++) {
            DF(m, 2)      = (FF(m, 0) - 27.0 * FF(m, 1) + 27.0 * FF(m, 2) - FF(m, 3)) / 24.0;
            DF(m, ni - 1) = -(FF(m, ni) - 27.0 * FF(m, ni - 1) + 27.0 * FF(m, ni - 2) - FF(m, ni - 3)) / 24.0;
            DF(m, 1)      = (-23.0 * FF(m, 0) + 21.0 * FF(m, 1) + 3.0 * FF(m, 2) - FF(m, 3)) / 24.0;
            DF(m, ni)     = -(-23.0 * FF(m, ni) + 21.0 * FF(m, ni - 1) + 3.0 * FF(m, ni - 2) - FF(m, ni - 3)) / 24.0;
        }
    }
#undef FF
#undef DF
}

/* flux.f90:17-77 flux_SW_pn */
static void flux_sw_pn(const double *prim, double nx, double ny, double nz, double nt,
                       double *f, double efix, double gama, int npn)
{
    double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    double u2 = 0.5 * (um * um + vm * vm + wm * wm);
    double hint = gethm(prim, gama);
    double eint = hint - pm / rm;
    double em = eint + u2;
    double hm;
    double cm = sqrt(gama * pm / rm);
    double ct = nx * um + ny * vm + nz * wm + nt;
    double cgm = sqrt(nx * nx + ny * ny + nz * nz);
    if (cgm < SML_SSS) cgm = SML_SSS;
    double l1 = ct, l4 = ct + cm * cgm, l5 = ct - cm * cgm;
    double eps = efix * efix * cgm * cgm;
    l1 = 0.5 * (l1 + npn * sqrt(l1 * l1 + eps));
    l4 = 0.5 * (l4 + npn * sqrt(l4 * l4 + eps));
    l5 = 0.5 * (l5 + npn * sqrt(l5 * l5 + eps));
    double cgm1 = 1.0 / cgm;
    double nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1;
    double cta = (ct - nt) * cgm1;
    double c2r = cm * cm / gama;
    double x1 = c2r * (2.0 * l1 - l4 - l5) / (2.0 * cm * cm);
    double x2 = c2r * (l4 - l5) / (2.0 * cm);
    hm = em + pm / rm;
    f[0] = (l1 - x1) * rm;
    f[1] = (l1 * um - x1 * um + nxa * x2) * rm;
    f[2] = (l1 * vm - x1 * vm + nya * x2) * rm;
    f[3] = (l1 * wm - x1 * wm + nza * x2) * rm;
    f[4] = (l1 * em - x1 * hm + cta * x2) * rm;
}

/* flux.f90:2-15 flux_SW */
void o_flux_sw(const double *priml, const double *primr, double nx, double ny, double nz,
               double nt, double *f, double efix, double gama)
{
    double fl[5], fr[5];
    flux_sw_pn(priml, nx, ny, nz, nt, fl, efix, gama, 1);
    flux_sw_pn(primr, nx, ny, nz, nt, fr, efix, gama, -1);
    for (int m = 0; m < 5; m++) f[m] = fl[m] + fr[m];
}

/* Convect.f90:1245-1272 flux */
static void flux_phys(const double *prim, double gama, double nx, double ny, double nz,
                      double nt, double *f)
{
    double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    double em = getem(prim, gama);
    double cit = nx * um + ny * vm + nz * wm + nt;
    f[0] = rm * cit;
    f[1] = f[0] * um + nx * pm;
    f[2] = f[0] * vm + ny * pm;
    f[3] = f[0] * wm + nz * pm;
    f[4] = (em + pm) * cit - pm * nt;
}

/* flux.f90:116-244 flux_Roe.  Returns 1 when c2<=0 (the reference stops). */
int o_flux_roe(const double *qlin, const double *qrin, double nx, double ny, double nz,
               double nt, double *f, double efix, double gama)
{
    double ql[5], qr[5], fl[5], fr[5], dq[5], qroe[5], df[5];
    memcpy(ql, qlin, sizeof ql);
    memcpy(qr, qrin, sizeof qr);
    flux_phys(ql, gama, nx, ny, nz, nt, fl);
    flux_phys(qr, gama, nx, ny, nz, nt, fr);
    for (int m = 0; m < 5; m++) dq[m] = qr[m] - ql[m];
    double gama1 = gama - 1.0;
    double gama2 = gama / gama1;
    ql[4] = gama2 * ql[4] / ql[0] + 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
    qr[4] = gama2 * qr[4] / qr[0] + 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    qroe[0] = sqrt(ql[0] * qr[0]);
    double rrl = sqrt(qr[0] / ql[0]);
    double rrl_1 = rrl + 1.0;
    for (int m = 1; m < 5; m++) qroe[m] = (ql[m] + qr[m] * rrl) / rrl_1;
    double rm = qroe[0], um = qroe[1], vm = qroe[2], wm = qroe[3], hm = qroe[4];
    double c2 = (hm - 0.5 * (um * um + vm * vm + wm * wm)) * gama1;
    if (c2 <= 0.0) return 1;
    double c = sqrt(c2);
    double cta = um * nx + vm * ny + wm * nz + nt;
    double cgm = sqrt(nx * nx + ny * ny + nz * nz);
    if (cgm < SML_SSS) cgm = SML_SSS;
    double cgm1 = 1.0 / cgm;
    nx = nx * cgm1; ny = ny * cgm1; nz = nz * cgm1;
    double ccgm = c * cgm;
    double l1 = fabs(cta - ccgm), l4 = fabs(cta), l5 = fabs(cta + ccgm);
    double l12 = l1 * l1, l42 = l4 * l4, l52 = l5 * l5;
    double deti = efix * cgm;
    double det2 = deti * deti;
    l1 = sqrt(l12 + det2);
    l4 = sqrt(l42 + det2);
    l5 = sqrt(l52 + det2);
    double dcta = dq[1] * nx + dq[2] * ny + dq[3] * nz;
    double cta1 = cta * cgm1;
    double rodcta = qroe[0] * dcta;
    double dp_c2 = dq[4] / c2;
    double rodctac = rodcta / c;
    double a1 = l4 * (dq[0] - dp_c2);
    double a2 = l5 * (dp_c2 + rodctac) * 0.5;
    double a3 = l1 * (dp_c2 - rodctac) * 0.5;
    double a4 = a1 + a2 + a3;
    double a5 = c * (a2 - a3);
    double a6 = l4 * (rm * dq[1] - nx * rodcta);
    double a7 = l4 * (rm * dq[2] - ny * rodcta);
    double a8 = l4 * (rm * dq[3] - nz * rodcta);
    df[0] = a4;
    df[1] = um * a4 + nx * a5 + a6;
    df[2] = vm * a4 + ny * a5 + a7;
    df[3] = wm * a4 + nz * a5 + a8;
    df[4] = hm * a4 + cta1 * a5 + um * a6 + vm * a7 + wm * a8 - c2 * a1 / gama1;
    for (int m = 0; m < 5; m++) f[m] = 0.5 * (fr[m] + fl[m] - df[m]);
    return 0;
}

/* flux.f90:410-501 flux_vanleer_pn: van Leer's split flux of one state,
 * npn = 1 (F+) or -1 (F-).  Returns -1 where gama p / rho < 0 (the reference
 * stops there, flux.f90:437-441). */
static int flux_vanleer_pn(const double *prim, double nx, double ny, double nz, double nt, double *f,
                           double gama, int npn)
{
    double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    double u2 = 0.5 * (um * um + vm * vm + wm * wm);
    double hint = gethm(prim, gama);
    double hm = hint + u2;
    double c2 = gama * pm / rm;
    if (c2 < 0.0) return -1;
    double cm = sqrt(c2);
    double ct = nx * um + ny * vm + nz * wm + nt;
    double cgm = sqrt(nx * nx + ny * ny + nz * nz) + SML_SSS;
    double cgm1 = 1.0 / cgm;
    double nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1, cta = ct * cgm1;
    double bmvl = cta / cm;
    if (npn == 1) {
        if (fabs(bmvl) < 1.0) {
            double ctcm = (-cta + 2.0 * cm) / gama;
            f[0] = 0.25 * cgm * rm * cm * ((bmvl + 1.0) * (bmvl + 1.0));
            f[1] = f[0] * (um + nxa * ctcm); f[2] = f[0] * (vm + nya * ctcm);
            f[3] = f[0] * (wm + nza * ctcm); f[4] = f[0] * hm;
        } else if (bmvl >= 1.0) {
            double rmct = rm * cta;
            f[0] = cgm * rmct; f[1] = cgm * (rmct * um + pm * nxa); f[2] = cgm * (rmct * vm + pm * nya);
            f[3] = cgm * (rmct * wm + pm * nza); f[4] = cgm * rmct * hm;
        } else {
            for (int n = 0; n < 5; n++) f[n] = 0.0;
        }
    } else {
        if (fabs(bmvl) < 1.0) {
            double ctcm = (-cta - 2.0 * cm) / gama;
            f[0] = -0.25 * cgm * rm * cm * ((bmvl - 1.0) * (bmvl - 1.0));
            f[1] = f[0] * (um + nxa * ctcm); f[2] = f[0] * (vm + nya * ctcm);
            f[3] = f[0] * (wm + nza * ctcm); f[4] = f[0] * hm;
        } else if (bmvl <= -1.0) {
            double rmct = rm * cta;
            f[0] = cgm * rmct; f[1] = cgm * (rmct * um + pm * nxa); f[2] = cgm * (rmct * vm + pm * nya);
            f[3] = cgm * (rmct * wm + pm * nza); f[4] = cgm * rmct * hm;
        } else {
            for (int n = 0; n < 5; n++) f[n] = 0.0;
        }
    }
    return 0;
}

/* flux.f90:395-408 flux_vanleer: F+(qL) + F-(qR) */
int o_flux_vanleer(const double *priml, const double *primr, double nx, double ny, double nz, double nt,
                   double *f, double efix, double gama)
{
    double fl[5], fr[5];
    (void)efix;
    if (flux_vanleer_pn(priml, nx, ny, nz, nt, fl, gama, 1)) return -1;
    if (flux_vanleer_pn(primr, nx, ny, nz, nt, fr, gama, -1)) return -1;
    for (int m = 0; m < 5; m++) f[m] = fl[m] + fr[m];
    return 0;
}

/* WCNSE5.f90:4858-4912 pave: Roe average of two primitive states */
void o_pave(const double *ql, const double *qr, double *qa, double gama)
{
    double gam1 = gama - 1.0, gam2 = gama / gam1;
    double rl = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4];
    double rr = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4];
    double vl2 = ul * ul + vl * vl + wl * wl;
    double vr2 = ur * ur + vr * vr + wr * wr;
    double hl = gam2 * pl / rl + 0.5 * vl2;
    double hr = gam2 * pr / rr + 0.5 * vr2;
    double rls = sqrt(rl), rrs = sqrt(rr);
    double ra = rls * rrs;
    double oors = 1.0 / (rls + rrs);
    rls = rls * oors;
    rrs = rrs * oors;
    double ua = rls * ul + rrs * ur, va = rls * vl + rrs * vr, wa = rls * wl + rrs * wr;
    double ha = rls * hl + rrs * hr;
    double va2 = ua * ua + va * va + wa * wa;
    qa[0] = ra; qa[1] = ua; qa[2] = va; qa[3] = wa;
    qa[4] = (ha - 0.5 * va2) * ra / gam2;
}

/* WCNSE5.f90:4960-5004 LxQ: left eigenvectors of the face normal gykb(1:3)
 * at the primitive state prim, applied to dq (gykb(4) is normalised there
 * and never used) */
void o_lxq(const double *prim, const double *gykb, const double *dq, double *f, double gama)
{
    double nx = gykb[0], ny = gykb[1], nz = gykb[2];
    double sn = sqrt(nx * nx + ny * ny + nz * nz);
    double oon = 1.0 / (sn > SML_SSS ? sn : SML_SSS);
    nx = nx * oon; ny = ny * oon; nz = nz * oon;
    double rm = prim[0], pm = prim[4];
    double c2 = gama * pm / rm;
    double cm = sqrt(c2);
    double dqn = nx * dq[1] + ny * dq[2] + nz * dq[3];
    double wp1 = dq[0] - dq[4] / c2;
    double wp2 = dq[4] / (rm * cm);
    f[0] = nz * dq[2] - ny * dq[3] + nx * wp1;
    f[1] = -nz * dq[1] + nx * dq[3] + ny * wp1;
    f[2] = ny * dq[1] - nx * dq[2] + nz * wp1;
    f[3] = dqn + wp2;
    f[4] = -dqn + wp2;
}

/* WCNSE5.f90:5006-5050 RxQ: the right eigenvectors, the inverse of LxQ */
void o_rxq(const double *prim, const double *gykb, const double *dq, double *f, double gama)
{
    double nx = gykb[0], ny = gykb[1], nz = gykb[2];
    double sn = sqrt(nx * nx + ny * ny + nz * nz);
    double oon = 1.0 / (sn > SML_SSS ? sn : SML_SSS);
    nx = nx * oon; ny = ny * oon; nz = nz * oon;
    double rm = prim[0], pm = prim[4];
    double c2 = gama * pm / rm;
    double cm = sqrt(c2);
    double dqn = nx * dq[0] + ny * dq[1] + nz * dq[2];
    double wp1 = 0.5 * rm * (dq[3] + dq[4]);
    double wp2 = 0.5 * (dq[3] - dq[4]);
    f[0] = dqn + wp1 / cm;
    f[1] = -nz * dq[1] + ny * dq[2] + nx * wp2;
    f[2] = nz * dq[0] - nx * dq[2] + ny * wp2;
    f[3] = -ny * dq[0] + nx * dq[1] + nz * wp2;
    f[4] = wp1 * cm;
}

/* flux.f90:246-393 flux_roemx: the shock-stable Roe flux of Kim and Kim
 * (J. Comput. Phys. 2005; "RoeM"), whose pressure-wiggle control uses the
 * face's minimum pressure ratio rpmin (Convect.f90:78-200) */
void o_flux_roemx(const double *ql, const double *qr, double sx, double sy, double sz, double rpmin,
                  double *flr, double gama)
{
    double sn = sqrt(sx * sx + sy * sy + sz * sz);
    if (sn < SML_SSS) sn = SML_SSS;
    double osn = 1.0 / sn;
    double nx = sx * osn, ny = sy * osn, nz = sz * osn;
    double rl = ql[0] + SMALL, ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4] + SMALL;
    double rr = qr[0] + SMALL, ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4] + SMALL;
    double vl2 = ul * ul + vl * vl + wl * wl;
    double vr2 = ur * ur + vr * vr + wr * wr;
    double ael = gama - 1.0, aer = gama - 1.0;
    double el = pl / (rl * ael), er = pr / (rr * aer);
    el = el + 0.5 * vl2;
    double hl = el + pl / rl;
    er = er + 0.5 * vr2;
    double hr = er + pr / rr;
    double vnl = nx * ul + ny * vl + nz * wl;
    double vnr = nx * ur + ny * vr + nz * wr;
    double dq[5], qi[5], mxdq[5];
    dq[0] = rr - rl;
    dq[1] = rr * ur - rl * ul;
    dq[2] = rr * vr - rl * vl;
    dq[3] = rr * wr - rl * wl;
    dq[4] = rr * hr - rl * hl;
    double rls = sqrt(rl), rrs = sqrt(rr);
    qi[0] = rls * rrs;
    double oors = 1.0 / (rls + rrs);
    rls = rls * oors;
    rrs = rrs * oors;
    qi[1] = rls * ul + rrs * ur;
    qi[2] = rls * vl + rrs * vr;
    qi[3] = rls * wl + rrs * wr;
    qi[4] = rls * hl + rrs * hr;
    double gami = rls * gama + rrs * gama;
    double vi2 = qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3];
    double vi2p2 = 0.5 * vi2;
    double ci2 = (gami - 1.0) * (qi[4] - vi2p2);
    double ci = sqrt(ci2);
    double vni = qi[1] * nx + qi[2] * ny + qi[3] * nz;
    double b1 = 0.0, b2 = 0.0;
    if (vni + ci > b1) b1 = vni + ci;
    if (vnr + ci > b1) b1 = vnr + ci;
    if (vni - ci < b2) b2 = vni - ci;
    if (vnl - ci < b2) b2 = vnl - ci;
    double oob12 = 1.0 / (b1 - b2);
    b1 = b1 * oob12;
    double b1b2 = b1 * b2;
    b2 = b2 * oob12;
    double mi = vni / ci, mia = fabs(mi);
    double fden = vi2 <= SMALL ? 1.0 : pow(mia, 1.0 - rpmin);
    double rpmlr = pl / pr < pr / pl ? pl / pr : pr / pl;
    double gcdc = mia <= SMALL ? 1.0 : pow(mia, 1.0 - rpmlr);
    double dvn = vnr - vnl, dvx = ur - ul, dvy = vr - vl, dvz = wr - wl, dps = pr - pl, dh = hr - hl;
    double rmp = dq[0] - fden * dps / ci2;
    mxdq[0] = rmp;
    mxdq[1] = rmp * qi[1] + qi[0] * (dvx - nx * dvn);
    mxdq[2] = rmp * qi[2] + qi[0] * (dvy - ny * dvn);
    mxdq[3] = rmp * qi[3] + qi[0] * (dvz - nz * dvn);
    mxdq[4] = rmp * qi[4] + qi[0] * dh;
    double rvnl = b1 * rl * vnl;
    double rvnr = -b2 * rr * vnr;
    double plr = b1 * pl - b2 * pr;
    flr[0] = rvnl + rvnr + b1b2 * dq[0];
    flr[1] = rvnl * ul + rvnr * ur + nx * plr + b1b2 * dq[1];
    flr[2] = rvnl * vl + rvnr * vr + ny * plr + b1b2 * dq[2];
    flr[3] = rvnl * wl + rvnr * wr + nz * plr + b1b2 * dq[3];
    flr[4] = rvnl * hl + rvnr * hr + b1b2 * dq[4];
    double cmxd = -gcdc * b1b2 / (1.0 + mia);
    for (int m = 0; m < 5; m++) flr[m] = (flr[m] + cmxd * mxdq[m]) * sn;
}

static int face_flux_at(int flux, const double *ql, const double *qr, double kx, double ky,
                        double kz, double kt, double *f, double efix, double gama, const double *rpmin, int face);

static int face_flux(int flux, const double *ql, const double *qr, double kx, double ky,
                     double kz, double kt, double *f, double efix, double gama)
{
    if (flux == 3) return o_flux_roe(ql, qr, kx, ky, kz, kt, f, efix, gama);
    if (flux == 2) return o_flux_vanleer(ql, qr, kx, ky, kz, kt, f, efix, gama);
    o_flux_sw(ql, qr, kx, ky, kz, kt, f, efix, gama);
    return 0;
}

static int face_flux_at(int flux, const double *ql, const double *qr, double kx, double ky,
                        double kz, double kt, double *f, double efix, double gama, const double *rpmin, int face)
{
    if (flux == 4) {
        /* the face index reaches flux_roemx through common /rpind/ i, the
         * loop index of every caller (WCNSE5.f90:2797, 4186) */
        o_flux_roemx(ql, qr, kx, ky, kz, rpmin[face], f, gama);
        return 0;
    }
    return face_flux(flux, ql, qr, kx, ky, kz, kt, f, efix, gama);
}

/* WCNSE5.f90:5052-5123 VALUE_HALF_NODE_45 for n > 4: VALUE_HALF_NODE's faces
 * 0..ni plus faces -2, -1, ni+1, ni+2 from the ghost metrics.  q(5,-3:ni+4),
 * q_half(5,-2:ni+2) (index i+2). */
void o_value_half_node_45(int ni, const double *q, double *q_half)
{
#define QQ(m,i) q[(m) + 5 * ((i) + 3)]
#define QH(m,i) q_half[(m) + 5 * ((i) + 2)]
    for (int m = 0; m < 5; m++) {
        for (int i = 2; i <= ni - 2; i++)
            QH(m, i) = (9.0 * (QQ(m, i) + QQ(m, i + 1)) + -1.0 * (QQ(m, i + 2) + QQ(m, i - 1))) / 16.0;
        QH(m, 1) = (5.0 * QQ(m, 1) + 15.0 * QQ(m, 2) + -5.0 * QQ(m, 3) + 1.0 * QQ(m, 4)) / 16.0;
        QH(m, ni - 1) = (5.0 * QQ(m, ni) + 15.0 * QQ(m, ni - 1) + -5.0 * QQ(m, ni - 2) + 1.0 * QQ(m, ni - 3)) / 16.0;
        QH(m, 0) = (35.0 * QQ(m, 1) - 35.0 * QQ(m, 2) + 21.0 * QQ(m, 3) - 5.0 * QQ(m, 4)) / 16.0;
        QH(m, ni) = (35.0 * QQ(m, ni) - 35.0 * QQ(m, ni - 1) + 21.0 * QQ(m, ni - 2) - 5.0 * QQ(m, ni - 3)) / 16.0;
        QH(m, -1) = (5.0 * QQ(m, 0) + 15.0 * QQ(m, -1) + -5.0 * QQ(m, -2) + 1.0 * QQ(m, -3)) / 16.0;
        QH(m, -2) = (9.0 * (QQ(m, -2) + QQ(m, -1)) + -1.0 * (QQ(m, 0) + QQ(m, -3))) / 16.0;
        QH(m, ni + 1) = (5.0 * QQ(m, ni + 1) + 15.0 * QQ(m, ni + 2) + -5.0 * QQ(m, ni + 3) + 1.0 * QQ(m, ni + 4)) / 16.0;
        QH(m, ni + 2) = (9.0 * (QQ(m, ni + 2) + QQ(m, ni + 3)) + -1.0 * (QQ(m, ni + 4) + QQ(m, ni + 1))) / 16.0;
    }
#undef QQ
#undef QH
}

static int wcns_line_core(int scheme, int flux, int ni, int nmax, double efix, double gama,
                          const double *trxyz, const double *trx1, const double *q, const double *rpmin, double *df);

/*
 * WCNSE5.f90:2-206 WCNS_E_5 (scheme 4) and WCNSE5.f90:2087-2759 WCNS_E_5_41.
 * q(1:6,-2:nmax+3,1:2) (only the first plane is read), trxyz(5,nmax), df(5,nmax).
 * Returns the number of positivity fallbacks (nerror) or -1 on a Roe c2<=0.
 */
int o_wcns_line(int scheme, int flux, int ni, int nmax, double efix, double gama,
                const double *trxyz, const double *q, double *df)
{
    return wcns_line_core(scheme, flux, ni, nmax, efix, gama, trxyz, NULL, q, NULL, df);
}

/*
 * The line routine of every scheme with every flux: schemes 4, 41, 42 as
 * o_wcns_line, and WCNS_E_5_45 (WCNSE5.f90:3442-4155) and WCNS_E_5_47
 * (WCNSE5.f90:4158-4854), whose metric line trx1(5,-3:ni+4) carries the four
 * ghost metrics inviscd3d gathers at either end (Convect.f90:180-297); flux 4
 * (flux_roemx) reads rpmin(0:ni).  trx1 and rpmin may be NULL when unused.
 */
int o_wcns_line_ex(int scheme, int flux, int ni, int nmax, double efix, double gama,
                   const double *trx1, const double *q, const double *rpmin, double *df)
{
    double *tr = (double *)malloc(sizeof(double) * 5 * (ni > 0 ? ni : 1));
    for (int i = 1; i <= ni; i++)
        for (int m = 0; m < 5; m++) tr[5 * (i - 1) + m] = trx1[5 * (i + 3) + m];
    int r = wcns_line_core(scheme, flux, ni, nmax, efix, gama, tr, trx1, q, rpmin, df);
    free(tr);
    return r;
}

static int wcns_line_core(int scheme, int flux, int ni, int nmax, double efix, double gama,
                          const double *trxyz, const double *trx1, const double *q, const double *rpmin, double *df)
{
#define QL(m,i)  q[((m) - 1) + 6 * ((i) + 2)]
    const double CL[3] = {1.0 / 16.0, 10.0 / 16.0, 5.0 / 16.0};
    const double CR[3] = {5.0 / 16.0, 10.0 / 16.0, 1.0 / 16.0};
    const double EPS = 1.0e-6;
    int np = nmax + 1;
    double *nxbuf = (double *)malloc(sizeof(double) * 5 * (np + 4));
    double *nx   = nxbuf + 10;           /* faces -2..nmax+2 */
    double *f    = (double *)malloc(sizeof(double) * 5 * np);
    double *qwl  = (double *)malloc(sizeof(double) * 5 * np);
    double *qwr  = (double *)malloc(sizeof(double) * 5 * np);
    int ng = ni + 4;                     /* g,s,wl,wr over -1..ni+2 */
    double *g  = (double *)malloc(sizeof(double) * 5 * 3 * ng);
    double *s  = (double *)malloc(sizeof(double) * 5 * 3 * ng);
    double *wl = (double *)malloc(sizeof(double) * 5 * 3 * ng);
    double *wr = (double *)malloc(sizeof(double) * 5 * 3 * ng);
#define NX(m,i)   nx[(i) * 5 + (m) - 1]
#define FL(m,i)   f[(i) * 5 + (m) - 1]
#define QWL(m,i)  qwl[(i) * 5 + (m) - 1]
#define QWR(m,i)  qwr[(i) * 5 + (m) - 1]
#define GG(m,n,i) g[((m) - 1) + 5 * (((n) - 1) + 3 * ((i) + 1))]
#define SS(m,n,i) s[((m) - 1) + 5 * (((n) - 1) + 3 * ((i) + 1))]
#define WL(m,n,i) wl[((m) - 1) + 5 * (((n) - 1) + 3 * ((i) + 1))]
#define WR(m,n,i) wr[((m) - 1) + 5 * (((n) - 1) + 3 * ((i) + 1))]
    int nerror = 0, bad = 0;
    double ql[5], qr[5], flr[5];
    double u_l_old[5][9], u_r_old[5][9];

    if (scheme == 45) o_value_half_node_45(ni, trx1, nxbuf);
    else o_value_half_node(5, nmax, ni, trxyz, nx);
    int ist = (QL(1, -2) < SMALL) ? 2 : 0;
    int ied = (QL(1, ni + 3) < SMALL) ? ni - 2 : ni;
    const int charw = scheme == 45 || scheme == 47;

    if (scheme == 45) {
        /* WCNSE5.f90:3492-3545: characteristic jumps cq(i) = L(face i)
         * (q(i+1) - q(i)) with |rho| in the Roe average, i = ist-2..ied+2;
         * the differences and weights of cq at i = ist..ied+1 */
        double *cq = (double *)malloc(sizeof(double) * 5 * (ni + 6));
#define CQ(m,i) cq[((m) - 1) + 5 * ((i) + 2)]
        for (int i = ist - 2; i <= ied + 2; i++) {
            double qal[5], qar[5], qtmp[5], qave[5];
            for (int m = 1; m <= NL; m++) { qal[m - 1] = QL(m, i); qar[m - 1] = QL(m, i + 1); }
            qal[0] = fabs(qal[0]);
            qar[0] = fabs(qar[0]);
            for (int m = 0; m < NL; m++) qtmp[m] = qar[m] - qal[m];
            o_pave(qal, qar, qave, gama);
            o_lxq(qave, &NX(1, i), qtmp, &CQ(1, i), gama);
        }
        for (int i = ist; i <= ied + 1; i++)
            for (int m = 1; m <= NL; m++) {
                double bl[3], br[3], IS;
                GG(m, 1, i) = 0.5 * (3.0 * CQ(m, i - 1) - CQ(m, i - 2));
                GG(m, 2, i) = 0.5 * (CQ(m, i) + CQ(m, i - 1));
                GG(m, 3, i) = 0.5 * (3.0 * CQ(m, i) - CQ(m, i + 1));
                SS(m, 1, i) = CQ(m, i - 1) - CQ(m, i - 2);
                SS(m, 2, i) = CQ(m, i) - CQ(m, i - 1);
                SS(m, 3, i) = CQ(m, i + 1) - CQ(m, i);
                for (int n = 1; n <= 3; n++) {
                    IS = GG(m, n, i) * GG(m, n, i) + SS(m, n, i) * SS(m, n, i);
                    double EIS2 = (EPS + IS) * (EPS + IS);
                    bl[n - 1] = CL[n - 1] / EIS2;
                    br[n - 1] = CR[n - 1] / EIS2;
                }
                IS = bl[0] + bl[1] + bl[2];
                for (int n = 1; n <= 3; n++) WL(m, n, i) = bl[n - 1] / IS;
                IS = br[0] + br[1] + br[2];
                for (int n = 1; n <= 3; n++) WR(m, n, i) = br[n - 1] / IS;
            }
#undef CQ
        free(cq);
        /* WCNSE5.f90:3550-3580: both states of face i back through R(face i) */
        for (int i = ist; i <= ied; i++) {
            int i1 = i + 1;
            double qal[5], qar[5], qa[5], qb[5], qave[5];
            for (int m = 1; m <= NL; m++) {
                qal[m - 1] = 0.125 * (WL(m, 1, i) * (SS(m, 1, i) + 4.0 * GG(m, 1, i)) +
                                      WL(m, 2, i) * (SS(m, 2, i) + 4.0 * GG(m, 2, i)) +
                                      WL(m, 3, i) * (SS(m, 3, i) + 4.0 * GG(m, 3, i)));
                qar[m - 1] = 0.125 * (WR(m, 1, i1) * (SS(m, 1, i1) - 4.0 * GG(m, 1, i1)) +
                                      WR(m, 2, i1) * (SS(m, 2, i1) - 4.0 * GG(m, 2, i1)) +
                                      WR(m, 3, i1) * (SS(m, 3, i1) - 4.0 * GG(m, 3, i1)));
            }
            double pa[5], pb[5];
            for (int m = 1; m <= NL; m++) { pa[m - 1] = QL(m, i); pb[m - 1] = QL(m, i1); }
            o_pave(pa, pb, qave, gama);
            o_rxq(qave, &NX(1, i), qal, qa, gama);
            o_rxq(qave, &NX(1, i), qar, qb, gama);
            for (int m = 1; m <= NL; m++) { QWL(m, i) = QL(m, i) + qa[m - 1]; QWR(m, i) = QL(m, i1) + qb[m - 1]; }
        }
    }
    if (scheme == 47) {
        /* WCNSE5.f90:4200-4278: differences in primitive variables, then
         * everything at node i in the characteristic variables of node i's
         * own state and metric; both states back through R(node) */
        for (int m = 1; m <= NL; m++) {
            int i = ist - 1;
            SS(m, 2, i) = QL(m, i - 1) - 2.0 * QL(m, i) + QL(m, i + 1);
            SS(m, 3, i) = QL(m, i) - 2.0 * QL(m, i + 1) + QL(m, i + 2);
            for (i = ist; i <= ied + 1; i++) {
                GG(m, 1, i) = 0.5 * (QL(m, i - 2) - 4.0 * QL(m, i - 1) + 3.0 * QL(m, i));
                GG(m, 2, i) = 0.5 * (QL(m, i + 1) - QL(m, i - 1));
                GG(m, 3, i) = 0.5 * (-3.0 * QL(m, i) + 4.0 * QL(m, i + 1) - QL(m, i + 2));
                SS(m, 1, i) = SS(m, 2, i - 1);
                SS(m, 2, i) = SS(m, 3, i - 1);
                SS(m, 3, i) = QL(m, i) - 2.0 * QL(m, i + 1) + QL(m, i + 2);
            }
        }
        double *qch = (double *)malloc(sizeof(double) * 5 * (ni + 4));   /* L q at i = -1..ni+2 */
#define QC(m,i) qch[((m) - 1) + 5 * ((i) + 1)]
        for (int i = ist; i <= ied + 1; i++) {
            double prim[5], v[5], IS0[5][3];
            const double *mt = &trx1[5 * (i + 3)];
            for (int m = 1; m <= NL; m++) prim[m - 1] = QL(m, i);
            o_lxq(prim, mt, prim, &QC(1, i), gama);
            for (int n = 1; n <= 3; n++) {
                double gch[5], sch[5];
                for (int m = 1; m <= NL; m++) v[m - 1] = GG(m, n, i);
                o_lxq(prim, mt, v, gch, gama);
                for (int m = 1; m <= NL; m++) v[m - 1] = SS(m, n, i);
                o_lxq(prim, mt, v, sch, gama);
                for (int m = 1; m <= NL; m++) {
                    GG(m, n, i) = gch[m - 1];
                    SS(m, n, i) = sch[m - 1];
                    IS0[m - 1][n - 1] = gch[m - 1] * gch[m - 1] + sch[m - 1] * sch[m - 1];
                }
            }
            for (int m = 1; m <= NL; m++) {
                double bl[3], br[3], IS;
                for (int n = 1; n <= 3; n++) {
                    double EIS2 = (EPS + IS0[m - 1][n - 1]) * (EPS + IS0[m - 1][n - 1]);
                    bl[n - 1] = CL[n - 1] / EIS2;
                    br[n - 1] = CR[n - 1] / EIS2;
                }
                IS = bl[0] + bl[1] + bl[2];
                for (int n = 1; n <= 3; n++) WL(m, n, i) = bl[n - 1] / IS;
                IS = br[0] + br[1] + br[2];
                for (int n = 1; n <= 3; n++) WR(m, n, i) = br[n - 1] / IS;
            }
        }
        for (int i = ist; i <= ied; i++) {
            int i1 = i + 1;
            double a[5], b[5], pa[5], pb[5], oa[5], ob[5];
            for (int m = 1; m <= NL; m++) {
                a[m - 1] = QC(m, i) + 0.125 * (WL(m, 1, i) * (SS(m, 1, i) + 4.0 * GG(m, 1, i)) +
                                               WL(m, 2, i) * (SS(m, 2, i) + 4.0 * GG(m, 2, i)) +
                                               WL(m, 3, i) * (SS(m, 3, i) + 4.0 * GG(m, 3, i)));
                b[m - 1] = QC(m, i1) + 0.125 * (WR(m, 1, i1) * (SS(m, 1, i1) - 4.0 * GG(m, 1, i1)) +
                                                WR(m, 2, i1) * (SS(m, 2, i1) - 4.0 * GG(m, 2, i1)) +
                                                WR(m, 3, i1) * (SS(m, 3, i1) - 4.0 * GG(m, 3, i1)));
                pa[m - 1] = QL(m, i);
                pb[m - 1] = QL(m, i1);
            }
            o_rxq(pa, &trx1[5 * (i + 3)], a, oa, gama);
            o_rxq(pb, &trx1[5 * (i1 + 3)], b, ob, gama);
            for (int m = 1; m <= NL; m++) { QWL(m, i) = oa[m - 1]; QWR(m, i) = ob[m - 1]; }
        }
#undef QC
        free(qch);
    }

    for (int m = 1; m <= NL && !charw; m++) {
        int i = ist - 1;
        SS(m, 2, i) = QL(m, i - 1) - 2.0 * QL(m, i) + QL(m, i + 1);
        SS(m, 3, i) = QL(m, i) - 2.0 * QL(m, i + 1) + QL(m, i + 2);
        for (i = ist; i <= ied + 1; i++) {
            double bl[3], br[3], IS;
            GG(m, 1, i) = 0.5 * (QL(m, i - 2) - 4.0 * QL(m, i - 1) + 3.0 * QL(m, i));
            GG(m, 2, i) = 0.5 * (QL(m, i + 1) - QL(m, i - 1));
            GG(m, 3, i) = 0.5 * (-3.0 * QL(m, i) + 4.0 * QL(m, i + 1) - QL(m, i + 2));
            SS(m, 1, i) = SS(m, 2, i - 1);
            SS(m, 2, i) = SS(m, 3, i - 1);
            SS(m, 3, i) = QL(m, i) - 2.0 * QL(m, i + 1) + QL(m, i + 2);
            for (int n = 1; n <= 3; n++) {
                IS = GG(m, n, i) * GG(m, n, i) + SS(m, n, i) * SS(m, n, i);
                double EIS2 = (EPS + IS) * (EPS + IS);
                bl[n - 1] = CL[n - 1] / EIS2;
                br[n - 1] = CR[n - 1] / EIS2;
            }
            IS = bl[0] + bl[1] + bl[2];
            for (int n = 1; n <= 3; n++) WL(m, n, i) = bl[n - 1] / IS;
            IS = br[0] + br[1] + br[2];
            for (int n = 1; n <= 3; n++) WR(m, n, i) = br[n - 1] / IS;
        }
    }
    for (int i = ist; i <= ied && !charw; i++) {
        int i1 = i + 1;
        for (int m = 1; m <= NL; m++) {
            QWL(m, i) = QL(m, i) + 0.125 * (WL(m, 1, i) * (SS(m, 1, i) + 4.0 * GG(m, 1, i)) +
                                            WL(m, 2, i) * (SS(m, 2, i) + 4.0 * GG(m, 2, i)) +
                                            WL(m, 3, i) * (SS(m, 3, i) + 4.0 * GG(m, 3, i)));
            QWR(m, i) = QL(m, i1) + 0.125 * (WR(m, 1, i1) * (SS(m, 1, i1) - 4.0 * GG(m, 1, i1)) +
                                             WR(m, 2, i1) * (SS(m, 2, i1) - 4.0 * GG(m, 2, i1)) +
                                             WR(m, 3, i1) * (SS(m, 3, i1) - 4.0 * GG(m, 3, i1)));
        }
    }
    if (ist > 1) {
        for (int m = 1; m <= NL; m++) {
            QWL(m, 0) = (35.0 * QL(m, 1) - 35.0 * QL(m, 2) + 21.0 * QL(m, 3) - 5.0 * QL(m, 4)) / 16.0;
            QWL(m, 1) = (5.0 * QL(m, 1) + 15.0 * QL(m, 2) - 5.0 * QL(m, 3) + QL(m, 4)) / 16.0;
            QWL(m, 2) = (-QL(m, 1) + 9.0 * QL(m, 2) + 9.0 * QL(m, 3) - QL(m, 4)) / 16.0;
            QWR(m, 0) = QWL(m, 0);
            QWR(m, 1) = QWL(m, 1);
        }
    }
    if (ied < ni) {
        for (int m = 1; m <= NL; m++) {
            QWR(m, ni)     = (35.0 * QL(m, ni) - 35.0 * QL(m, ni - 1) + 21.0 * QL(m, ni - 2) - 5.0 * QL(m, ni - 3)) / 16.0;
            QWR(m, ni - 1) = (5.0 * QL(m, ni) + 15.0 * QL(m, ni - 1) - 5.0 * QL(m, ni - 2) + QL(m, ni - 3)) / 16.0;
            QWR(m, ni - 2) = (-QL(m, ni) + 9.0 * QL(m, ni - 1) + 9.0 * QL(m, ni - 2) - QL(m, ni - 3)) / 16.0;
            QWL(m, ni)     = QWR(m, ni);
            QWL(m, ni - 1) = QWR(m, ni - 1);
        }
    }
    for (int i = 0; i <= ni; i++) {
        for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
        if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
            nerror++;
            int ii = i > 1 ? i : 1;
            for (int m = 1; m <= NL; m++) ql[m - 1] = QL(m, ii);
        }
        if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
            nerror++;
            int ii = i + 1 < ni ? i + 1 : ni;
            for (int m = 1; m <= NL; m++) qr[m - 1] = QL(m, ii);
        }
        bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
        for (int m = 1; m <= NL; m++) FL(m, i) = flr[m - 1];
    }
    o_flux_dxyz(5, nmax, ni, f, df);
#define DFO(m,i) df[((i) - 1) * 5 + (m) - 1]

    if ((scheme == 41 || scheme >= 42) && ist > 1) {
        /* WCNSE5.f90:2269-2513: left physical boundary, rows 1-3; scheme 42
         * (WCNS_E_5_42, WCNSE5.f90:2766-3441) differs only in the four
         * linear face states below (WCNSE5.f90:2967-2986) */
        for (int i = 1; i <= 4; i++)
            for (int m = 1; m <= NL; m++) { u_l_old[m - 1][i] = QWL(m, i); u_r_old[m - 1][i] = QWR(m, i); }
        for (int m = 1; m <= NL; m++) {
            if (scheme != 41) {
                QWL(m, 0) = (677.0 * QL(m, 1) - 421.0 * QL(m, 2) + 99.0 * QL(m, 3) + 13.0 * QL(m, 4)) / 368.0;
                QWL(m, 1) = (5.0 * QL(m, 1) + 15.0 * QL(m, 2) - 5.0 * QL(m, 3) + QL(m, 4)) / 16.0;
                QWL(m, 2) = (-QL(m, 1) + 9.0 * QL(m, 2) + 9.0 * QL(m, 3) - QL(m, 4)) / 16.0;
                QWL(m, 3) = (QL(m, 1) - 5.0 * QL(m, 2) + 15.0 * QL(m, 3) + 5.0 * QL(m, 4)) / 16.0;
            } else {
                QWL(m, 0) = (3.0 * QL(m, 1) - 1.0 * QL(m, 2) - 0. * QL(m, 3)) / 2.0;
                QWL(m, 1) = (19.0 * QL(m, 1) + 25.0 * QL(m, 2) - 2. * QL(m, 3)) / 42.0;
                QWL(m, 2) = (4.0 * QL(m, 1) + 3.0 * QL(m, 2) + 9.0 * QL(m, 3) + 2. * QL(m, 4)) / 18.0;
                QWL(m, 3) = (-2.0 * QL(m, 1) + 3.0 * QL(m, 2) + 3.0 * QL(m, 3) + 2. * QL(m, 4)) / 6.0;
            }
            QWR(m, 0) = QWL(m, 0);
            QWR(m, 1) = QWL(m, 1);
            QWR(m, 2) = QWL(m, 2);
            QWR(m, 3) = QWL(m, 3);
        }
        for (int i = 0; i <= 3; i++) {
            for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
            int ii = i > 1 ? i : 1;
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                for (int m = 1; m <= NL; m++) ql[m - 1] = (QL(m, ii) + QL(m, i + 1)) / 2.;
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                for (int m = 1; m <= NL; m++) qr[m - 1] = (QL(m, ii) + QL(m, i + 1)) / 2.;
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, 1) = (-23.0 * FL(m, 0) + 21.0 * FL(m, 1) + 3.0 * FL(m, 2) - 1.0 * FL(m, 3)) / 24.0;

        for (int m = 1; m <= NL; m++) {
            QWL(m, 1) = (3.0 * QL(m, 1) + 6.0 * QL(m, 2) - QL(m, 3)) / 8.0;
            QWR(m, 1) = QWL(m, 1);
            QWL(m, 2) = (-30.0 * QL(m, 1) + 145.0 * QL(m, 2) + 29.0 * QL(m, 3) - 8.0 * QL(m, 4)) / 136.0;
            QWR(m, 2) = (2.0 * QL(m, 1) + 49.0 * QL(m, 2) + 125.0 * QL(m, 3) - 40.0 * QL(m, 4)) / 136.0;
            QWL(m, 3) = (-QL(m, 2) + 6.0 * QL(m, 3) + 3.0 * QL(m, 4)) / 8.0;
            QWR(m, 3) = QWL(m, 3);
            QWL(m, 4) = (QL(m, 2) + 2.0 * QL(m, 3) + 5.0 * QL(m, 4)) / 8.0;
            QWR(m, 4) = QWL(m, 4);
            QWL(m, 5) = (QL(m, 2) + QL(m, 3) + 6.0 * QL(m, 4)) / 8.0;
            QWR(m, 5) = QWL(m, 5);
        }
        for (int i = 1; i <= 5; i++) {
            for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i > 1 ? i : 1;
                for (int m = 1; m <= NL; m++) ql[m - 1] = QL(m, ii);
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i + 1 < ni ? i + 1 : ni;
                for (int m = 1; m <= NL; m++) qr[m - 1] = QL(m, ii);
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, 2) = (-22.0 * FL(m, 1) + 17.0 * FL(m, 2) + 9.0 * FL(m, 3) - 5.0 * FL(m, 4) + FL(m, 5)) / 24.0;

        for (int m = 1; m <= NL; m++) {
            QWR(m, 1) = u_r_old[m - 1][1];
            QWL(m, 2) = (-29.0 * QL(m, 1) + 170.0 * QL(m, 2) + 63.0 * QL(m, 3) + 12.0 * QL(m, 4)) / 216.0;
            QWR(m, 2) = u_r_old[m - 1][2];
            QWR(m, 3) = u_r_old[m - 1][3];
            QWR(m, 4) = u_r_old[m - 1][4];
        }
        for (int i = 1; i <= 4; i++) {
            for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i > 1 ? i : 1;
                for (int m = 1; m <= NL; m++) ql[m - 1] = QL(m, ii);
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i + 1 < ni ? i + 1 : ni;
                for (int m = 1; m <= NL; m++) qr[m - 1] = QL(m, ii);
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, 3) = (FL(m, 1) - 27.0 * FL(m, 2) + 27.0 * FL(m, 3) - FL(m, 4)) / 24.0;
    }

    if ((scheme == 41 || scheme >= 42) && ied < ni) {
        /* WCNSE5.f90:2518-2751: right physical boundary, rows ni-2..ni; scheme
         * 42's face states WCNSE5.f90:3207-3226 */
        for (int i = 1; i <= 4; i++) {
            int i1 = ni - i;
            for (int m = 1; m <= NL; m++) { u_l_old[m - 1][9 - i] = QWL(m, i1); u_r_old[m - 1][9 - i] = QWR(m, i1); }
        }
        for (int m = 1; m <= NL; m++) {
            if (scheme != 41) {
                QWR(m, ni)     = (677.0 * QL(m, ni) - 421.0 * QL(m, ni - 1) + 99.0 * QL(m, ni - 2) + 13.0 * QL(m, ni - 3)) / 368.0;
                QWR(m, ni - 1) = (5.0 * QL(m, ni) + 15.0 * QL(m, ni - 1) - 5.0 * QL(m, ni - 2) + QL(m, ni - 3)) / 16.0;
                QWR(m, ni - 2) = (-QL(m, ni) + 9.0 * QL(m, ni - 1) + 9.0 * QL(m, ni - 2) - QL(m, ni - 3)) / 16.0;
                QWR(m, ni - 3) = (QL(m, ni) - 5.0 * QL(m, ni - 1) + 15.0 * QL(m, ni - 2) + 5.0 * QL(m, ni - 3)) / 16.0;
            } else {
                QWR(m, ni)     = (3.0 * QL(m, ni) - 1.0 * QL(m, ni - 1) - 0. * QL(m, ni - 2)) / 2.0;
                QWR(m, ni - 1) = (19.0 * QL(m, ni) + 25.0 * QL(m, ni - 1) - 2. * QL(m, ni - 2)) / 42.0;
                QWR(m, ni - 2) = (4.0 * QL(m, ni) + 3.0 * QL(m, ni - 1) + 9.0 * QL(m, ni - 2) + 2. * QL(m, ni - 3)) / 18.0;
                QWR(m, ni - 3) = (-2.0 * QL(m, ni) + 3.0 * QL(m, ni - 1) + 3.0 * QL(m, ni - 2) + 2. * QL(m, ni - 3)) / 6.0;
            }
            QWL(m, ni)     = QWR(m, ni);
            QWL(m, ni - 1) = QWR(m, ni - 1);
            QWL(m, ni - 2) = QWR(m, ni - 2);
            QWL(m, ni - 3) = QWR(m, ni - 3);
        }
        for (int i = ni; i >= ni - 3; i--) {
            for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
            /* WCNSE5.f90:2572,2580: min(i+1,1) is always 1 -- kept as written */
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                for (int m = 1; m <= NL; m++) ql[m - 1] = (QL(m, 1) + QL(m, i)) / 2.0;
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                for (int m = 1; m <= NL; m++) qr[m - 1] = (QL(m, 1) + QL(m, i)) / 2.0;
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, ni - i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, ni) = -(-23.0 * FL(m, 0) + 21.0 * FL(m, 1) + 3.0 * FL(m, 2) - 1.0 * FL(m, 3)) / 24.0;

        for (int m = 1; m <= NL; m++) {
            QWR(m, ni - 1) = (3.0 * QL(m, ni) + 6.0 * QL(m, ni - 1) - QL(m, ni - 2)) / 8.0;
            QWL(m, ni - 1) = QWR(m, ni - 1);
            QWR(m, ni - 2) = (-30.0 * QL(m, ni) + 145.0 * QL(m, ni - 1) + 29.0 * QL(m, ni - 2) - 8.0 * QL(m, ni - 3)) / 136.0;
            QWL(m, ni - 2) = (2.0 * QL(m, ni) + 49.0 * QL(m, ni - 1) + 125.0 * QL(m, ni - 2) - 40.0 * QL(m, ni - 3)) / 136.0;
            QWR(m, ni - 3) = (-QL(m, ni - 1) + 6.0 * QL(m, ni - 2) + 3.0 * QL(m, ni - 3)) / 8.0;
            QWL(m, ni - 3) = QWR(m, ni - 3);
            QWR(m, ni - 4) = (QL(m, ni - 1) + 2.0 * QL(m, ni - 2) + 5.0 * QL(m, ni - 3)) / 8.0;
            QWL(m, ni - 4) = QWR(m, ni - 4);
            QWR(m, ni - 5) = (QL(m, ni - 1) + QL(m, ni - 2) + 6.0 * QL(m, ni - 3)) / 8.0;
            QWL(m, ni - 5) = QWR(m, ni - 5);
        }
        for (int i = ni - 1; i >= ni - 5; i--) {
            for (int m = 1; m <= NL; m++) { ql[m - 1] = QWL(m, i); qr[m - 1] = QWR(m, i); }
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i > 1 ? i : 1;
                for (int m = 1; m <= NL; m++) ql[m - 1] = QL(m, ii);
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i + 1 < ni ? i + 1 : ni;
                for (int m = 1; m <= NL; m++) qr[m - 1] = QL(m, ii);
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, ni - i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, ni - 1) = -(-22.0 * FL(m, 1) + 17.0 * FL(m, 2) + 9.0 * FL(m, 3) - 5.0 * FL(m, 4) + FL(m, 5)) / 24.0;

        for (int m = 1; m <= NL; m++) {
            QWL(m, ni - 1) = u_l_old[m - 1][8];
            QWR(m, ni - 2) = (-29.0 * QL(m, ni) + 170.0 * QL(m, ni - 1) + 63.0 * QL(m, ni - 2) + 12.0 * QL(m, ni - 3)) / 216.0;
            QWL(m, ni - 2) = u_l_old[m - 1][7];
            QWL(m, ni - 3) = u_l_old[m - 1][6];
            QWL(m, ni - 4) = u_l_old[m - 1][5];
        }
        for (int i = ni - 1; i >= ni - 4; i--) {
            for (int m = 1; m <= NL; m++) { qr[m - 1] = QWR(m, i); ql[m - 1] = QWL(m, i); }
            if (ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i > 1 ? i : 1;
                for (int m = 1; m <= NL; m++) ql[m - 1] = QL(m, ii);
            }
            if (qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT) {
                nerror++;
                int ii = i + 1 < ni ? i + 1 : ni;
                for (int m = 1; m <= NL; m++) qr[m - 1] = QL(m, ii);
            }
            bad |= face_flux_at(flux, ql, qr, NX(1, i), NX(2, i), NX(3, i), NX(4, i), flr, efix, gama, rpmin, i);
            for (int m = 1; m <= NL; m++) FL(m, ni - i) = flr[m - 1];
        }
        for (int m = 1; m <= NL; m++)
            DFO(m, ni - 2) = -(FL(m, 1) - 27.0 * FL(m, 2) + 27.0 * FL(m, 3) - FL(m, 4)) / 24.0;
    }
#undef DFO
    free(nxbuf); free(f); free(qwl); free(qwr); free(g); free(s); free(wl); free(wr);
#undef QL
#undef NX
#undef FL
#undef QWL
#undef QWR
#undef GG
#undef SS
#undef WL
#undef WR
    return bad ? -1 : nerror;
}

/* LHS.f90:756-828 mxdq_std (gettmhmgm_ns for hint) */
void o_mxdq_std(const double *prim, const double *gykb, const double *dq, double *f,
                double rad, int npn, double gama)
{
    double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    double hint = gethm(prim, gama);
    double c2 = gama * pm / rm;
    double cm = sqrt(c2);
    double v2 = um * um + vm * vm + wm * wm;
    double hm = hint + 0.5 * v2;
    double nx = gykb[0], ny = gykb[1], nz = gykb[2], nt = gykb[3];
    double ct = nx * um + ny * vm + nz * wm + nt;
    double cgm = sqrt(nx * nx + ny * ny + nz * nz);
    if (cgm < SML_SSS) cgm = SML_SSS;
    double l1 = ct, l4 = ct + cm * cgm, l5 = ct - cm * cgm;
    l1 = 0.5 * (l1 + npn * rad);
    l4 = 0.5 * (l4 + npn * rad);
    l5 = 0.5 * (l5 + npn * rad);
    double x1 = (2.0 * l1 - l4 - l5) / (2.0 * c2);
    double x2 = (l4 - l5) / (2.0 * cm);
    double cgm1 = 1.0 / cgm;
    nx = nx * cgm1; ny = ny * cgm1; nz = nz * cgm1;
    ct = (ct - nt) * cgm1;
    double ae = gama - 1.0;
    double af = 0.5 * ae * v2;
    double dc = ct * dq[0] - nx * dq[1] - ny * dq[2] - nz * dq[3];
    double dh = af * dq[0] - ae * (um * dq[1] + vm * dq[2] + wm * dq[3] - dq[4]);
    double c2dc = c2 * dc;
    f[0] = l1 * dq[0] - dh * x1 - dc * x2;
    f[1] = l1 * dq[1] + (nx * c2dc - um * dh) * x1 + (nx * dh - um * dc) * x2;
    f[2] = l1 * dq[2] + (ny * c2dc - vm * dh) * x1 + (ny * dh - vm * dc) * x2;
    f[3] = l1 * dq[3] + (nz * c2dc - wm * dh) * x1 + (nz * dh - wm * dc) * x2;
    f[4] = l1 * dq[4] + (ct * c2dc - hm * dh) * x1 + (ct * dh - hm * dc) * x2;
}

/* =========================================================== case setup */
o_case *o_case_create(const o_params *p)
{
    o_case *c = (o_case *)calloc(1, sizeof(o_case));
    c->prm = *p;
    c->cap = 8;
    c->blk = (oblock *)calloc(c->cap, sizeof(oblock));
    c->nthreads = 1;
    return c;
}

/* Threads of the CPU timing run.  Every parallel loop writes disjoint nodes
 * and sums nothing in floating point across threads, and the sweeps run
 * hyperplane by hyperplane (i+j+k = const), which hands every node the same
 * operands as the lexicographic loops of LHS.f90; the results are therefore
 * the serial ones bit for bit. */
void o_set_threads(o_case *c, int n) { c->nthreads = n < 1 ? 1 : n; }
int o_view_threads(const o_case *c) { return c->nthreads; }

static void block_free(oblock *b)
{
    double **fs[] = {&b->x, &b->y, &b->z, &b->r, &b->u, &b->v, &b->w, &b->p, &b->t, &b->c,
                     &b->q, &b->dq, &b->rhs_nb, &b->kcx, &b->kcy, &b->kcz, &b->etx, &b->ety,
                     &b->etz, &b->ctx, &b->cty, &b->ctz, &b->vol, &b->sra, &b->srb, &b->src,
                     &b->srva, &b->srvb, &b->srvc, &b->dtdt, &b->visl, &b->vist, &b->qke, &b->dist, &b->qkw};
    for (size_t n = 0; n < sizeof fs / sizeof fs[0]; n++) { free(*fs[n]); *fs[n] = NULL; }
    for (int r = 0; r < b->nregions; r++) {
        free(b->reg[r].image); free(b->reg[r].jmage); free(b->reg[r].kmage);
    }
}

void o_case_destroy(o_case *c)
{
    if (!c) return;
    for (int n = 0; n < c->nblocks; n++) block_free(&c->blk[n]);
    free(c->blk);
    for (int g = 0; g < c->ngroups; g++) free(c->groups[g].p);
    free(c->groups);
    free(c);
}

int o_block_add(o_case *c, int ni, int nj, int nk, const double *x, const double *y, const double *z)
{
    if (c->nblocks == c->cap) {
        c->cap *= 2;
        c->blk = (oblock *)realloc(c->blk, sizeof(oblock) * c->cap);
        memset(c->blk + c->nblocks, 0, sizeof(oblock) * (c->cap - c->nblocks));
    }
    oblock *b = &c->blk[c->nblocks];
    memset(b, 0, sizeof *b);
    b->ni = ni; b->nj = nj; b->nk = nk;
    b->g3 = (long)(ni + 6) * (nj + 6) * (nk + 6);
    long n = (long)ni * nj * nk;
    b->x = (double *)malloc(sizeof(double) * n);
    b->y = (double *)malloc(sizeof(double) * n);
    b->z = (double *)malloc(sizeof(double) * n);
    memcpy(b->x, x, sizeof(double) * n);
    memcpy(b->y, y, sizeof(double) * n);
    memcpy(b->z, z, sizeof(double) * n);
    double **g3f[] = {&b->r, &b->u, &b->v, &b->w, &b->p, &b->t, &b->c, &b->kcx, &b->kcy, &b->kcz,
                      &b->etx, &b->ety, &b->etz, &b->ctx, &b->cty, &b->ctz, &b->vol, &b->sra,
                      &b->srb, &b->src, &b->srva, &b->srvb, &b->srvc, &b->dtdt, &b->visl, &b->vist,
                      &b->qke, &b->dist, &b->qkw};
    for (size_t f = 0; f < sizeof g3f / sizeof g3f[0]; f++) *g3f[f] = (double *)calloc(b->g3, sizeof(double));
    b->q = (double *)calloc(5 * b->g3, sizeof(double));
    b->dq = (double *)calloc(5 * b->g3, sizeof(double));
    b->rhs_nb = (double *)calloc(5 * b->g3, sizeof(double));
    return c->nblocks++;
}

int o_region_add(o_case *c, int nb, const int s_raw[6], int bctype, const int t_raw[6], int nbt, int ibcwin)
{
    if (nb < 0 || nb >= c->nblocks) return -1;
    oblock *b = &c->blk[nb];
    if (b->nregions >= MAXREG) return -2;
    oregion *g = &b->reg[b->nregions];
    memset(g, 0, sizeof *g);
    g->bctype = bctype;
    memcpy(g->raw_s, s_raw, sizeof(int) * 6);
    if (bctype < 0) { memcpy(g->raw_t, t_raw, sizeof(int) * 6); g->nbt = nbt; g->ibcwin = ibcwin; }
    return b->nregions++;
}

/* parallel.f90:1123-1301 analyze_bc + analyze_bc_connect (method = 1) */
static int analyze_region(o_case *c, oblock *b, oregion *g)
{
    (void)c;
    int s_st[3], s_ed[3];
    for (int m = 0; m < 3; m++) { s_st[m] = g->raw_s[2 * m]; s_ed[m] = g->raw_s[2 * m + 1]; }
    g->s_nd = -1;
    for (int m = 0; m < 3; m++) {
        g->s_st[m] = abs(s_st[m]) < abs(s_ed[m]) ? abs(s_st[m]) : abs(s_ed[m]);
        g->s_ed[m] = abs(s_st[m]) > abs(s_ed[m]) ? abs(s_st[m]) : abs(s_ed[m]);
        g->s_lr3d[m] = 0;
        if (s_st[m] == s_ed[m]) {
            g->s_nd = m + 1;
            if (s_st[m] == 1) { g->s_lr = -1; g->s_lr3d[m] = -1; }
            else              { g->s_lr = 1;  g->s_lr3d[m] = 1; }
        }
    }
    if (g->s_nd < 0) return -3;
    (void)b;
    if (g->bctype >= 0) return 0;
    int t_st[3], t_ed[3];
    for (int m = 0; m < 3; m++) { t_st[m] = g->raw_t[2 * m]; t_ed[m] = g->raw_t[2 * m + 1]; }
    g->t_nd = -1;
    for (int m = 0; m < 3; m++) {
        g->t_lr3d[m] = 0;
        if (t_st[m] == t_ed[m]) {
            g->t_nd = m + 1;
            if (t_st[m] == 1) { g->t_lr = -1; g->t_lr3d[m] = -1; }
            else              { g->t_lr = 1;  g->t_lr3d[m] = 1; }
        }
    }
    if (g->t_nd < 0) return -4;
    int sdir[3][3] = {{0}};     /* s_t_dirction(n,m) -> sdir[n-1][m-1] */
    int found = 0;
    for (int m = 1; m <= 3 && !found; m++)
        for (int n = 1; n <= 3 && !found; n++) {
            if (m != g->s_nd && n != g->t_nd) {
                int js1 = s_st[m - 1];
                if (abs(js1) < abs(s_ed[m - 1])) js1 = s_ed[m - 1];
                int js2 = t_st[n - 1];
                if (abs(js2) < abs(t_ed[n - 1])) js2 = t_ed[n - 1];
                if (js1 * js2 > 0) {
                    sdir[n - 1][m - 1] = 1;
                    sdir[g->t_nd - 1][g->s_nd - 1] = 1;
                    sdir[6 - n - g->t_nd - 1][6 - m - g->s_nd - 1] = 1;
                    found = 1;
                }
            }
        }
    int s_sign[3], t_sign[3], st_sign[3];
    for (int m = 0; m < 3; m++) {
        s_sign[m] = 1; s_st[m] = abs(s_st[m]); s_ed[m] = abs(s_ed[m]);
        if (m + 1 != g->s_nd && s_st[m] > s_ed[m]) s_sign[m] = -1;
        t_sign[m] = 1; t_st[m] = abs(t_st[m]); t_ed[m] = abs(t_ed[m]);
        if (m + 1 != g->t_nd && t_st[m] > t_ed[m]) t_sign[m] = -1;
    }
    for (int m = 0; m < 3; m++) {
        int co = 0;
        for (int n = 0; n < 3; n++) if (sdir[n][m] == 1) co = n + 1;
        st_sign[m] = (co > 0 ? t_sign[co - 1] : 1) * s_sign[m];
    }
    g->bi = g->s_ed[0] - g->s_st[0] + 1;
    g->bj = g->s_ed[1] - g->s_st[1] + 1;
    g->bk = g->s_ed[2] - g->s_st[2] + 1;
    long nbx = (long)g->bi * g->bj * g->bk;
    g->image = (int *)malloc(sizeof(int) * nbx);
    g->jmage = (int *)malloc(sizeof(int) * nbx);
    g->kmage = (int *)malloc(sizeof(int) * nbx);
    for (int i = s_st[0];; i += s_sign[0]) {
        int idelt = (i - s_st[0]) * st_sign[0];
        for (int j = s_st[1];; j += s_sign[1]) {
            int jdelt = (j - s_st[1]) * st_sign[1];
            for (int k = s_st[2];; k += s_sign[2]) {
                int kdelt = (k - s_st[2]) * st_sign[2];
                long o = (i - g->s_st[0]) + (long)g->bi * ((j - g->s_st[1]) + (long)g->bj * (k - g->s_st[2]));
                g->image[o] = t_st[0] + sdir[0][0] * idelt + sdir[0][1] * jdelt + sdir[0][2] * kdelt;
                g->jmage[o] = t_st[1] + sdir[1][0] * idelt + sdir[1][1] * jdelt + sdir[1][2] * kdelt;
                g->kmage[o] = t_st[2] + sdir[2][0] * idelt + sdir[2][1] * jdelt + sdir[2][2] * kdelt;
                if (k == s_ed[2]) break;
            }
            if (j == s_ed[1]) break;
        }
        if (i == s_ed[0]) break;
    }
    for (int m = 0; m < 3; m++) {
        g->t_st[m] = t_st[m] < t_ed[m] ? t_st[m] : t_ed[m];
        g->t_ed[m] = t_st[m] > t_ed[m] ? t_st[m] : t_ed[m];
    }
    return 0;
}

static inline long rbox(const oregion *g, int i, int j, int k)
{
    return (i - g->s_st[0]) + (long)g->bi * ((j - g->s_st[1]) + (long)g->bj * (k - g->s_st[2]));
}

/* WCNSE5.f90:1478-1728 GRID_DERIVATIVE_gcl (gcl = 1) */
static void grid_derivative_gcl(oblock *b)
{
    int ni = b->ni, nj = b->nj, nk = b->nk;
    int nmax = ni > nj ? ni : nj; nmax = nmax > nk ? nmax : nk;
    long n = (long)ni * nj * nk;
#define XI(a,i,j,k) a[((i) - 1) + (long)ni * (((j) - 1) + (long)nj * ((k) - 1))]
#define D18(m,i,j,k) d[((m) - 1) + 18 * (((i) - 1) + (long)ni * (((j) - 1) + (long)nj * ((k) - 1)))]
    double *d = (double *)calloc(18 * n, sizeof(double));
    double *xk = (double *)calloc(3 * n, sizeof(double));
    double *tem = (double *)malloc(sizeof(double) * 6 * (nmax + 1));
    double *temh = (double *)malloc(sizeof(double) * 6 * (nmax + 1));
    double *dtem = (double *)malloc(sizeof(double) * 6 * (nmax + 1));
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) {
        for (int i = 1; i <= ni; i++) { tem[(i-1)*3+0] = XI(b->x,i,j,k); tem[(i-1)*3+1] = XI(b->y,i,j,k); tem[(i-1)*3+2] = XI(b->z,i,j,k); }
        o_value_half_node(3, nmax, ni, tem, temh);
        o_flux_dxyz(3, nmax, ni, temh, tem);
        for (int i = 1; i <= ni; i++) { D18(1,i,j,k) = tem[(i-1)*3+0]; D18(4,i,j,k) = tem[(i-1)*3+1]; D18(7,i,j,k) = tem[(i-1)*3+2]; }
    }
    for (int k = 1; k <= nk; k++) for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) { tem[(j-1)*3+0] = XI(b->x,i,j,k); tem[(j-1)*3+1] = XI(b->y,i,j,k); tem[(j-1)*3+2] = XI(b->z,i,j,k); }
        o_value_half_node(3, nmax, nj, tem, temh);
        o_flux_dxyz(3, nmax, nj, temh, tem);
        for (int j = 1; j <= nj; j++) { D18(2,i,j,k) = tem[(j-1)*3+0]; D18(5,i,j,k) = tem[(j-1)*3+1]; D18(8,i,j,k) = tem[(j-1)*3+2]; }
    }
    for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        for (int k = 1; k <= nk; k++) { tem[(k-1)*3+0] = XI(b->x,i,j,k); tem[(k-1)*3+1] = XI(b->y,i,j,k); tem[(k-1)*3+2] = XI(b->z,i,j,k); }
        o_value_half_node(3, nmax, nk, tem, temh);
        o_flux_dxyz(3, nmax, nk, temh, tem);
        for (int k = 1; k <= nk; k++) { D18(3,i,j,k) = tem[(k-1)*3+0]; D18(6,i,j,k) = tem[(k-1)*3+1]; D18(9,i,j,k) = tem[(k-1)*3+2]; }
    }
    /* with GCL=1 the cross-product metrics of WCNSE5.f90:1565-1582 are overwritten below,
     * so only x_kc,x_et,x_ct (xkcetct) and the x*y style products are needed. */
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        long o = (i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1));
        xk[3*o+0] = D18(1,i,j,k); xk[3*o+1] = D18(2,i,j,k); xk[3*o+2] = D18(3,i,j,k);
        double xm = XI(b->x,i,j,k), ym = XI(b->y,i,j,k), zm = XI(b->z,i,j,k);
        F3(b,kcx,i,j,k) = D18(1,i,j,k) * ym; F3(b,kcy,i,j,k) = D18(2,i,j,k) * ym; F3(b,kcz,i,j,k) = D18(3,i,j,k) * ym;
        F3(b,etx,i,j,k) = D18(4,i,j,k) * zm; F3(b,ety,i,j,k) = D18(5,i,j,k) * zm; F3(b,etz,i,j,k) = D18(6,i,j,k) * zm;
        F3(b,ctx,i,j,k) = D18(7,i,j,k) * xm; F3(b,cty,i,j,k) = D18(8,i,j,k) * xm; F3(b,ctz,i,j,k) = D18(9,i,j,k) * xm;
    }
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) {
        for (int i = 1; i <= ni; i++) {
            double *t6 = &tem[(i - 1) * 6];
            t6[0] = F3(b,cty,i,j,k); t6[1] = F3(b,ctz,i,j,k); t6[2] = F3(b,kcy,i,j,k);
            t6[3] = F3(b,kcz,i,j,k); t6[4] = F3(b,ety,i,j,k); t6[5] = F3(b,etz,i,j,k);
        }
        o_value_half_node(6, nmax, ni, tem, temh);
        o_flux_dxyz(6, nmax, ni, temh, dtem);
        for (int i = 1; i <= ni; i++) for (int m = 1; m <= 6; m++) D18(m,i,j,k) = dtem[(i-1)*6 + m - 1];
    }
    for (int k = 1; k <= nk; k++) for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) {
            double *t6 = &tem[(j - 1) * 6];
            t6[0] = F3(b,ctx,i,j,k); t6[1] = F3(b,ctz,i,j,k); t6[2] = F3(b,kcx,i,j,k);
            t6[3] = F3(b,kcz,i,j,k); t6[4] = F3(b,etx,i,j,k); t6[5] = F3(b,etz,i,j,k);
        }
        o_value_half_node(6, nmax, nj, tem, temh);
        o_flux_dxyz(6, nmax, nj, temh, dtem);
        for (int j = 1; j <= nj; j++) for (int m = 1; m <= 6; m++) D18(m + 6,i,j,k) = dtem[(j-1)*6 + m - 1];
    }
    for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        for (int k = 1; k <= nk; k++) {
            double *t6 = &tem[(k - 1) * 6];
            t6[0] = F3(b,ctx,i,j,k); t6[1] = F3(b,cty,i,j,k); t6[2] = F3(b,kcx,i,j,k);
            t6[3] = F3(b,kcy,i,j,k); t6[4] = F3(b,etx,i,j,k); t6[5] = F3(b,ety,i,j,k);
        }
        o_value_half_node(6, nmax, nk, tem, temh);
        o_flux_dxyz(6, nmax, nk, temh, dtem);
        for (int k = 1; k <= nk; k++) for (int m = 1; m <= 6; m++) D18(m + 12,i,j,k) = dtem[(k-1)*6 + m - 1];
    }
    for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
        long o = (i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1));
        F3(b,kcx,i,j,k) = D18(18,i,j,k) - D18(12,i,j,k);
        F3(b,kcy,i,j,k) = D18(14,i,j,k) - D18(8,i,j,k);
        F3(b,kcz,i,j,k) = D18(16,i,j,k) - D18(10,i,j,k);
        F3(b,etx,i,j,k) = D18(6,i,j,k) - D18(17,i,j,k);
        F3(b,ety,i,j,k) = D18(2,i,j,k) - D18(13,i,j,k);
        F3(b,etz,i,j,k) = D18(4,i,j,k) - D18(15,i,j,k);
        F3(b,ctx,i,j,k) = D18(11,i,j,k) - D18(5,i,j,k);
        F3(b,cty,i,j,k) = D18(7,i,j,k) - D18(1,i,j,k);
        F3(b,ctz,i,j,k) = D18(9,i,j,k) - D18(3,i,j,k);
        F3(b,vol,i,j,k) = xk[3*o+0] * F3(b,kcx,i,j,k) + xk[3*o+1] * F3(b,etx,i,j,k) + xk[3*o+2] * F3(b,ctx,i,j,k);
    }
    free(d); free(xk); free(tem); free(temh); free(dtem);
#undef XI
#undef D18
}

/* parallel.f90:177-393 mod_comp_connects + parallel.f90:4202-4311 connect_pre_parallel */
static int pnt_eq(opnt a, opnt b) { return a.nb == b.nb && a.i == b.i && a.j == b.j && a.k == b.k; }
static void group_push(ogroup *g, opnt p)
{
    if (g->n == g->cap) { g->cap = g->cap ? 2 * g->cap : 4; g->p = (opnt *)realloc(g->p, sizeof(opnt) * g->cap); }
    g->p[g->n++] = p;
}
static void build_groups(o_case *c)
{
    int ncomp = 0, capc = 0;
    ogroup *comp = NULL;
    for (int nb = 0; nb < c->nblocks; nb++)
        for (int r = 0; r < c->blk[nb].nregions; r++) c->blk[nb].reg[r].iused = 0;
    for (int nb = 0; nb < c->nblocks; nb++)
        for (int r = 0; r < c->blk[nb].nregions; r++) {
            oregion *g = &c->blk[nb].reg[r];
            if (g->bctype < 0 && g->iused == 0) {
                g->iused = 1;
                c->blk[g->nbt - 1].reg[g->ibcwin - 1].iused = -1;
            }
        }
    const int mcyc[5] = {1, 2, 3, 1, 2};
    for (int nb = 0; nb < c->nblocks; nb++)
        for (int r = 0; r < c->blk[nb].nregions; r++) {
            oregion *g = &c->blk[nb].reg[r];
            if (g->bctype >= 0 || g->iused != 1) continue;
            int ib[3], ie[3];
            for (int m = 0; m < 3; m++) { ib[m] = g->s_st[m]; ie[m] = g->s_ed[m]; }
            int id2 = mcyc[g->s_nd] - 1, id3 = mcyc[g->s_nd + 1] - 1;
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
                for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
                    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                        int idx[3] = {i, j, k};
                        int inner = 1;
                        for (int m = 0; m < 3; m++) {
                            int lo = ib[m], hi = ie[m];
                            if (m == id2 || m == id3) { lo++; hi--; }
                            if (idx[m] < lo || idx[m] > hi) inner = 0;
                        }
                        if (inner) continue;
                        long o = rbox(g, i, j, k);
                        opnt a = {nb + 1, i, j, k};
                        opnt bt = {g->nbt, g->image[o], g->jmage[o], g->kmage[o]};
                        int done = 0;
                        for (int q = 0; q < ncomp && !done; q++) {
                            ogroup *cg = &comp[q];
                            int ea = 0, eb = 0;
                            for (int s = 0; s < cg->n; s++) { if (pnt_eq(cg->p[s], a)) { ea = 1; break; } }
                            for (int s = 0; s < cg->n; s++) { if (pnt_eq(cg->p[s], bt)) { eb = 1; break; } }
                            if (ea || eb) {
                                if (!ea) group_push(cg, a);
                                if (!eb) group_push(cg, bt);
                                done = 1;
                            }
                        }
                        if (!done) {
                            if (ncomp == capc) { capc = capc ? 2 * capc : 64; comp = (ogroup *)realloc(comp, sizeof(ogroup) * capc); }
                            memset(&comp[ncomp], 0, sizeof(ogroup));
                            group_push(&comp[ncomp], a);
                            group_push(&comp[ncomp], bt);
                            ncomp++;
                        }
                    }
        }
    c->ngroups = 0;
    c->groups = NULL;
    for (int q = 0; q < ncomp; q++) {
        if (comp[q].n > 2) {
            c->groups = (ogroup *)realloc(c->groups, sizeof(ogroup) * (c->ngroups + 1));
            c->groups[c->ngroups++] = comp[q];
        } else {
            free(comp[q].p);
        }
    }
    free(comp);
}

/* PreProcess.f90:968-1113 init_inflow (height<0 branch; non-dimensional freestream) */
static void init_inflow(o_case *c)
{
    double gama = c->prm.gama, moo = c->prm.moo;
    c->attack_r = c->prm.attack * PAI / 180.0;
    c->sideslip_r = c->prm.sideslip * PAI / 180.0;
    c->uoo = cos(c->attack_r) * cos(c->sideslip_r);
    c->voo = sin(c->attack_r) * cos(c->sideslip_r);
    c->woo = sin(c->sideslip_r);
    c->roo = 1.0;
    c->too = 1.0;
    c->poo = 1.0 / (gama * moo * moo);
    c->coo = sqrt(gama * c->poo / c->roo);
    c->eoo = c->poo / (gama - 1.0) + 0.5 * c->roo * (c->uoo * c->uoo + c->voo * c->voo + c->woo * c->woo);
    c->visc = 110.4 / c->prm.tref;
    /* PreProcess.f90:994: Reynolds number per grid length unit */
    c->prm.reynolds = (c->prm.reynolds / c->prm.lfref) * c->prm.lref;
    c->re = 1.0 / c->prm.reynolds;
    /* turblence.f90:3166-3181 inflow_turbulence */
    c->vtoo = 1.0e-6;
    c->muoo = 1.0e-1;
    if (c->prm.nlamtur == 2) {
        double lref_lfref = c->prm.lref / c->prm.lfref;
        c->omgoo = 10.0 * lref_lfref;
        c->koo = c->omgoo * c->vtoo / c->prm.reynolds;
    }
}

/* turblence.f90:3091-3117 init_SST: k = koo, omega = omgoo, mu_t = vtoo on -1..n+1 */
static void init_sst(o_case *c)
{
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int k = -1; k <= b->nk + 1; k++) for (int j = -1; j <= b->nj + 1; j++) for (int i = -1; i <= b->ni + 1; i++) {
            long q = G(b, i, j, k);
            b->qke[q] = c->koo;
            b->qkw[q] = c->omgoo;
            b->vist[q] = c->vtoo;
        }
    }
}

/* turblence.f90:2710-2734 init_SA: nu~ = muoo, mu_t = muoo fv1(muoo) on -1..n+1 */
static void init_sa(o_case *c)
{
    const double cv13 = pow(7.1, 3.0);
    double kafan = c->muoo;
    double kf3 = kafan * kafan * kafan;     /* kafan**3.0, as the flang build evaluates it */
    double fv1 = kf3 / (kf3 + cv13);
    double vt = c->muoo * fv1;
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int k = -1; k <= b->nk + 1; k++) for (int j = -1; j <= b->nj + 1; j++) for (int i = -1; i <= b->ni + 1; i++) {
            long q = G(b, i, j, k);
            b->qke[q] = c->muoo;
            b->vist[q] = vt;
        }
    }
}

/* parallel.f90:4313-4468 calc_wall_dists_parallel: squared distance to the
 * nearest node of any no-slip wall region (bctype 2, 20, 21), then sqrt + small */
int o_wall_distance(o_case *c)
{
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++)
            b->dist[G(b, i, j, k)] = 1.0e30;
    }
    for (int wb = 0; wb < c->nblocks; wb++) {
        oblock *w = &c->blk[wb];
        for (int r = 0; r < w->nregions; r++) {
            oregion *g = &w->reg[r];
            if (!(g->bctype == 2 || g->bctype == 20 || g->bctype == 21)) continue;
            for (int kw = g->s_st[2]; kw <= g->s_ed[2]; kw++)
                for (int jw = g->s_st[1]; jw <= g->s_ed[1]; jw++)
                    for (int iw = g->s_st[0]; iw <= g->s_ed[0]; iw++) {
                        long qw = (iw - 1) + (long)w->ni * ((jw - 1) + (long)w->nj * (kw - 1));
                        double xw = w->x[qw], yw = w->y[qw], zw = w->z[qw];
                        for (int nb = 0; nb < c->nblocks; nb++) {
                            oblock *b = &c->blk[nb];
                            for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
                                long q = (i - 1) + (long)b->ni * ((j - 1) + (long)b->nj * (k - 1));
                                double dx = b->x[q] - xw, dy = b->y[q] - yw, dz = b->z[q] - zw;
                                double dist = dx * dx + dy * dy + dz * dz;
                                long qg = G(b, i, j, k);
                                if (dist < b->dist[qg]) b->dist[qg] = dist;
                            }
                        }
                    }
        }
    }
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
            long q = G(b, i, j, k);
            b->dist[q] = sqrt(b->dist[q]) + SMALL;
        }
    }
    return 0;
}

int o_setup(o_case *c)
{
    static const int order[6] = {-1, 4, 5, 6, 2, 3};   /* PreProcess.f90:2543-2573 */
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        int n = 0;
        for (int o = 0; o < 6; o++)
            for (int r = 0; r < b->nregions; r++) {
                int bt = b->reg[r].bctype < 0 ? -1 : b->reg[r].bctype;
                if (bt == order[o]) b->bcindex[n++] = r;
            }
        if (n != b->nregions) return -10;
        for (int r = 0; r < b->nregions; r++) {
            int e = analyze_region(c, b, &b->reg[r]);
            if (e) return e;
        }
        grid_derivative_gcl(b);
        /* check_grid_derivative PreProcess.f90:2575-2678: negative Jacobian is fatal */
        for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++)
            if (F3(b,vol,i,j,k) <= 0.0) return -11;
    }
    build_groups(c);
    init_inflow(c);
    if (c->prm.nlamtur >= 0) {
        /* turblence.f90:1735-1785 init_turbulence: wall distance, uniform field
         * (nlamtur 1: SA, 2: SST) */
        if ((c->prm.nlamtur != 1 && c->prm.nlamtur != 2) || c->prm.nvis != 1) return -12;
        o_wall_distance(c);
        if (c->prm.nlamtur == 1) init_sa(c);
        else init_sst(c);
    }
    return 0;
}

long o_g3_size(const o_case *c, int nb) { return c->blk[nb].g3; }

void o_freestream(const o_case *c, double out[8])
{
    out[0] = c->roo; out[1] = c->uoo; out[2] = c->voo; out[3] = c->woo;
    out[4] = c->poo; out[5] = c->coo; out[6] = c->too; out[7] = c->visc;
}

int o_set_prim(o_case *c, int nb, const double *r, const double *u, const double *v,
               const double *w, const double *p)
{
    oblock *b = &c->blk[nb];
    size_t n = sizeof(double) * b->g3;
    memcpy(b->r, r, n); memcpy(b->u, u, n); memcpy(b->v, v, n); memcpy(b->w, w, n); memcpy(b->p, p, n);
    /* sub_tgh.f90:36-58 initial_t -> RHS.f90:412-434 get_t_ini (interior only, nvis == 1) */
    double moo2 = c->prm.moo * c->prm.moo;
    memset(b->t, 0, n);
    if (c->prm.nvis == 1)
        for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
            long q = G(b, i, j, k);
            b->t[q] = moo2 * (c->prm.gama * b->p[q] / b->r[q]);
        }
    return 0;
}

static double *field_ptr(const oblock *b, int which)
{
    switch (which) {
    case 0: return b->r;   case 1: return b->u;   case 2: return b->v;   case 3: return b->w;
    case 4: return b->p;   case 5: return b->t;   case 6: return b->c;   case 7: return b->dtdt;
    case 8: return b->sra; case 9: return b->srb; case 10: return b->src;
    case 11: return b->srva; case 12: return b->srvb; case 13: return b->srvc;
    case 14: return b->visl; case 15: return b->vist; case 16: return b->vol;
    case 17: return b->kcx; case 18: return b->kcy; case 19: return b->kcz;
    case 20: return b->etx; case 21: return b->ety; case 22: return b->etz;
    case 23: return b->ctx; case 24: return b->cty; case 25: return b->ctz;
    case 26: return b->qke; case 27: return b->dist; case 28: return b->qkw;
    default: return NULL;
    }
}

int o_get_field(const o_case *c, int nb, int which, double *out)
{
    const oblock *b = &c->blk[nb];
    if (which == 100) { memcpy(out, b->dq, sizeof(double) * 5 * b->g3); return 0; }
    if (which == 101) { memcpy(out, b->q, sizeof(double) * 5 * b->g3); return 0; }
    if (which == 102) { memcpy(out, b->rhs_nb, sizeof(double) * 5 * b->g3); return 0; }
    double *f = field_ptr(b, which);
    if (!f) return -1;
    memcpy(out, f, sizeof(double) * b->g3);
    return 0;
}

int o_set_field(o_case *c, int nb, int which, const double *in)
{
    oblock *b = &c->blk[nb];
    if (which == 100) { memcpy(b->dq, in, sizeof(double) * 5 * b->g3); return 0; }
    double *f = field_ptr(b, which);
    if (!f) return -1;
    memcpy(f, in, sizeof(double) * b->g3);
    return 0;
}

int o_region_images(const o_case *c, int nb, int nr, int *img, int *jmg, int *kmg, int *box)
{
    const oregion *g = &c->blk[nb].reg[nr];
    for (int m = 0; m < 3; m++) { box[m] = g->s_st[m]; box[3 + m] = g->s_ed[m]; }
    if (g->bctype >= 0) return 1;
    long n = (long)g->bi * g->bj * g->bk;
    memcpy(img, g->image, sizeof(int) * n);
    memcpy(jmg, g->jmage, sizeof(int) * n);
    memcpy(kmg, g->kmage, sizeof(int) * n);
    return 0;
}

int o_multipoint(const o_case *c, int *ngroups, int *npts, int *out, int cap)
{
    *ngroups = c->ngroups;
    int tot = 0;
    for (int g = 0; g < c->ngroups; g++) {
        for (int s = 0; s < c->groups[g].n; s++) {
            if (4 * (tot + 1) <= cap) {
                out[4 * tot + 0] = c->groups[g].p[s].nb; out[4 * tot + 1] = c->groups[g].p[s].i;
                out[4 * tot + 2] = c->groups[g].p[s].j;  out[4 * tot + 3] = c->groups[g].p[s].k;
            }
            tot++;
        }
        npts[g] = c->groups[g].n;
    }
    return tot;
}

long o_counter(const o_case *c, int which) { return which == 0 ? c->n_fallback : c->n_repair; }

/* ====================================================== boundary conditions */
/* Sub1.f90:750-780 getnxyz_mml */
static void getnxyz(const oblock *b, int i, int j, int k, int is, int js, int ks,
                    double *nx, double *ny, double *nz)
{
    double x = 0, y = 0, z = 0;
    for (int n = 0; n < abs(is); n++) { x = x + F3(b,kcx,i,j,k); y = y + F3(b,kcy,i,j,k); z = z + F3(b,kcz,i,j,k); }
    for (int n = 0; n < abs(js); n++) { x = x + F3(b,etx,i,j,k); y = y + F3(b,ety,i,j,k); z = z + F3(b,etz,i,j,k); }
    for (int n = 0; n < abs(ks); n++) { x = x + F3(b,ctx,i,j,k); y = y + F3(b,cty,i,j,k); z = z + F3(b,ctz,i,j,k); }
    double s = sqrt(x * x + y * y + z * z);
    if (s < SML_SSS) s = SML_SSS;
    *nx = x / s; *ny = y / s; *nz = z / s;
}

/* the turbulence variables are carried through the boundary conditions (nvis = 1,
 * nlamtur >= 0): nu~ (SA), or k and omega (SST) */
#define TURB(c) ((c)->prm.nvis == 1 && (c)->prm.nlamtur >= 0)
static int nturb(const o_case *c) { return TURB(c) ? c->prm.nlamtur : 0; }
static double *tvar(const oblock *b, int m) { return m == 0 ? b->qke : b->qkw; }
/* BC_farfield_turbulence's inflow values (turblence.f90:2321-2338) */
static double tinf(const o_case *c, int m) { return c->prm.nlamtur == 1 ? c->muoo : (m == 0 ? c->koo : c->omgoo); }
#define SETP(b,is,js,ks,rv,uv,vv,wv,pv) do { long _g = G(b,is,js,ks); (b)->r[_g]=(rv); (b)->u[_g]=(uv); (b)->v[_g]=(vv); (b)->w[_g]=(wv); (b)->p[_g]=(pv); } while (0)

/* Sub2.f90:206-322 boundary_n1_vir3 (cic1 = 0; turbulence copies when nvis==1 & nlamtur>=0) */
static void bc_interface(o_case *c, oblock *b, oregion *g)
{
    oblock *bt = &c->blk[g->nbt - 1];
    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++) {
                long o = rbox(g, i, j, k);
                int it = g->image[o] - g->t_lr3d[0], jt = g->jmage[o] - g->t_lr3d[1], kt = g->kmage[o] - g->t_lr3d[2];
                int is = i + g->s_lr3d[0], js = j + g->s_lr3d[1], ks = k + g->s_lr3d[2];
                if (TURB(c)) {
                    /* Sub2.f90:266-280: first ghost from the partner, face node = mean of the two neighbours */
                    long gt = G(bt, it, jt, kt), gs = G(b, is, js, ks), g0 = G(b, i, j, k);
                    long gi = G(b, i - g->s_lr3d[0], j - g->s_lr3d[1], k - g->s_lr3d[2]);
                    b->vist[gs] = bt->vist[gt];
                    b->vist[g0] = (b->vist[gs] + b->vist[gi]) * 0.5;
                    for (int m = 0; m < nturb(c); m++) {
                        double *q = tvar(b, m);
                        q[gs] = tvar(bt, m)[gt];
                        q[g0] = (q[gs] + q[gi]) * 0.5;
                    }
                }
                for (int l = 1; l <= 3; l++) {
                    long gt = G(bt, it, jt, kt), gs = G(b, is, js, ks);
                    b->r[gs] = bt->r[gt]; b->u[gs] = bt->u[gt]; b->v[gs] = bt->v[gt];
                    b->w[gs] = bt->w[gt]; b->p[gs] = bt->p[gt];
                    is += g->s_lr3d[0]; js += g->s_lr3d[1]; ks += g->s_lr3d[2];
                    it -= g->t_lr3d[0]; jt -= g->t_lr3d[1]; kt -= g->t_lr3d[2];
                }
            }
}

/* BC.f90:460-607 boundary_4 (method = 1) */
static void bc_farfield(o_case *c, oblock *b, oregion *g)
{
    double gama = c->prm.gama;
    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++) {
                const int *l3 = g->s_lr3d;
                int it = i - l3[0], jt = j - l3[1], kt = k - l3[2];
                double nx, ny, nz;
                getnxyz(b, i, j, k, l3[0], l3[1], l3[2], &nx, &ny, &nz);
                double rin = R_(b,it,jt,kt), uin = U_(b,it,jt,kt), vin = V_(b,it,jt,kt);
                double win = W_(b,it,jt,kt), pin = P_(b,it,jt,kt);
                double s_in = pin / pow(rin, gama);
                double s_oo = c->poo / pow(c->roo, gama);
                double cin = sqrt(gama * pin / rin);
                double sq = sqrt(nx * nx + ny * ny + nz * nz);
                if (sq < SML_SSS) sq = SML_SSS;
                double cgm1 = g->s_lr / sq;
                double nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1;
                double vnoo = nxa * c->uoo + nya * c->voo + nza * c->woo;
                double vnin = nxa * uin + nya * vin + nza * win;
                double mach = sqrt(uin * uin + vin * vin + win * win) / cin;
                double rb, ub, vb, wb, pb;
                if (mach < 1.0) {
                    double rp = vnin + 2.0 * cin / (gama - 1.0);
                    double rn = vnoo - 2.0 * c->coo / (gama - 1.0);
                    double vnb = 0.5 * (rp + rn);
                    double ab = 0.25 * (gama - 1.0) * (rp - rn);
                    double sb, uref, vref, wref, vnref;
                    if (vnb > 0.0) { sb = s_in; uref = uin; vref = vin; wref = win; vnref = vnin; }
                    else { sb = s_oo; uref = c->uoo; vref = c->voo; wref = c->woo; vnref = vnoo; }
                    rb = pow(ab * ab / gama / sb, 1.0 / (gama - 1.0));
                    ub = uref + nxa * (vnb - vnref);
                    vb = vref + nya * (vnb - vnref);
                    wb = wref + nza * (vnb - vnref);
                    pb = sb * pow(rb, gama);
                } else if (vnin > 0.0) {
                    rb = rin; ub = uin; vb = vin; wb = win; pb = pin;
                } else {
                    rb = c->roo; ub = c->uoo; vb = c->voo; wb = c->woo; pb = c->poo;
                }
                SETP(b, i, j, k, rb, ub, vb, wb, pb);
                if (TURB(c)) {
                    /* sub_tgh.f90:713-787 BC_farfield_turbulence_tgh: inflow takes the
                     * freestream value, outflow the interior neighbour's, on the face
                     * node and the first ghost */
                    long g0 = G(b, i, j, k);
                    double tem_lr = (double)l3[g->s_nd - 1], ut;
                    if (g->s_nd == 1) ut = tem_lr * (b->kcx[g0] * b->u[g0] + b->kcy[g0] * b->v[g0] + b->kcz[g0] * b->w[g0]);
                    else if (g->s_nd == 2) ut = tem_lr * (b->etx[g0] * b->u[g0] + b->ety[g0] * b->v[g0] + b->etz[g0] * b->w[g0]);
                    else ut = tem_lr * (b->ctx[g0] * b->u[g0] + b->cty[g0] * b->v[g0] + b->ctz[g0] * b->w[g0]);
                    for (int m = 0; m <= 1; m++) {
                        long gm = G(b, i + m * l3[0], j + m * l3[1], k + m * l3[2]);
                        for (int v = 0; v < nturb(c); v++)
                            tvar(b, v)[gm] = ut < 0.0 ? tinf(c, v) : tvar(b, v)[G(b, it, jt, kt)];
                    }
                }
                SETP(b, i + l3[0], j + l3[1], k + l3[2], rb, ub, vb, wb, pb);
                SETP(b, i + 2 * l3[0], j + 2 * l3[1], k + 2 * l3[2], rb, ub, vb, wb, pb);
                SETP(b, i + 3 * l3[0], j + 3 * l3[1], k + 3 * l3[2], -rb, ub, vb, wb, pb);
            }
}

/* BC.f90:609-665 boundary_5 */
static void bc_fixed(o_case *c, oblock *b, oregion *g)
{
    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
                for (int n = 0; n <= 3; n++) {
                    SETP(b, i + n * g->s_lr3d[0], j + n * g->s_lr3d[1], k + n * g->s_lr3d[2],
                         c->roo, c->uoo, c->voo, c->woo, c->poo);
                    /* turblence.f90:2321-2338 BC_farfield_turbulence on the face node and first ghost */
                    if (TURB(c) && n <= 1)
                        for (int v = 0; v < nturb(c); v++)
                            tvar(b, v)[G(b, i + n * g->s_lr3d[0], j + n * g->s_lr3d[1], k + n * g->s_lr3d[2])] = tinf(c, v);
                }
}

/* BC.f90:667-749 boundary_6 */
static void bc_outflow(o_case *c, oblock *b, oregion *g)
{
    (void)c;
    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++) {
                const int *l3 = g->s_lr3d;
                int it = i - l3[0], jt = j - l3[1], kt = k - l3[2];
                double rm = R_(b,it,jt,kt), um = U_(b,it,jt,kt), vm = V_(b,it,jt,kt), wm = W_(b,it,jt,kt), pm = P_(b,it,jt,kt);
                for (int n = 0; n <= 3; n++) {
                    SETP(b, i + n * l3[0], j + n * l3[1], k + n * l3[2], rm, um, vm, wm, pm);
                    /* turblence.f90:2340-2355 BC_outflow_turbulence on the face node and first ghost */
                    if (TURB(c) && n <= 1)
                        for (int v = 0; v < nturb(c); v++)
                            tvar(b, v)[G(b, i + n * l3[0], j + n * l3[1], k + n * l3[2])] = tvar(b, v)[G(b, it, jt, kt)];
                }
            }
}

/* BC.f90:1972-2008 dif_average (+ BC_face_dif_turbulence, turblence.f90:2252-2265) */
static void dif_average(o_case *c, oblock *b, oregion *g)
{
    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++) {
                const int *l3 = g->s_lr3d;
                long gs = G(b, i + l3[0], j + l3[1], k + l3[2]), gt = G(b, i - l3[0], j - l3[1], k - l3[2]);
                long g0 = G(b, i, j, k);
                b->r[g0] = 0.5 * (b->r[gs] + b->r[gt]);
                b->u[g0] = 0.5 * (b->u[gs] + b->u[gt]);
                b->v[g0] = 0.5 * (b->v[gs] + b->v[gt]);
                b->w[g0] = 0.5 * (b->w[gs] + b->w[gt]);
                b->p[g0] = 0.5 * (b->p[gs] + b->p[gt]);
                for (int v = 0; v < nturb(c); v++) tvar(b, v)[g0] = 0.5 * (tvar(b, v)[gt] + tvar(b, v)[gs]);
            }
}

/* BC.f90:1412-1486 inviscid_wall */
static void bc_inviscid_wall(o_case *c, oblock *b, oregion *g)
{
    const int *l3 = g->s_lr3d;
    for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                double nx, ny, nz;
                getnxyz(b, i, j, k, l3[0], l3[1], l3[2], &nx, &ny, &nz);
                for (int l = 1; l <= 3; l++) {
                    int is = i + l * l3[0], js = j + l * l3[1], ks = k + l * l3[2];
                    int it = i - l * l3[0], jt = j - l * l3[1], kt = k - l * l3[2];
                    double sss1 = 2.0 * (U_(b,it,jt,kt) * nx + V_(b,it,jt,kt) * ny + W_(b,it,jt,kt) * nz);
                    double rr = R_(b,it,jt,kt);
                    SETP(b, is, js, ks, l == 3 ? -rr : rr, U_(b,it,jt,kt) - sss1 * nx,
                         V_(b,it,jt,kt) - sss1 * ny, W_(b,it,jt,kt) - sss1 * nz, P_(b,it,jt,kt));
                }
            }
    dif_average(c, b, g);
}

/* BC.f90:1244-1409 viscid_wall */
static void bc_viscous_wall(o_case *c, oblock *b, oregion *g)
{
    const int *l3 = g->s_lr3d;
    double moocp = c->prm.moo * c->prm.moo * c->prm.gama;
    for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                int is = i + l3[0], js = j + l3[1], ks = k + l3[2];
                int it = i - l3[0], jt = j - l3[1], kt = k - l3[2];
                SETP(b, is, js, ks, R_(b,it,jt,kt), -U_(b,it,jt,kt), -V_(b,it,jt,kt), -W_(b,it,jt,kt), P_(b,it,jt,kt));
                /* turblence.f90:2267-2306 BC_wall_turbulence: odd reflection of nu~ / k in the
                 * first ghost; SST omega there: 1600 mu/(rho Re d^2) - omega(first interior),
                 * with mu from the face node's temperature and the face node's old density */
                if (TURB(c)) b->qke[G(b, is, js, ks)] = -b->qke[G(b, it, jt, kt)];
                if (TURB(c) && c->prm.nlamtur == 2) {
                    long g0 = G(b, i, j, k), gt = G(b, it, jt, kt);
                    double tm = b->t[g0];
                    double vsl = tm * sqrt(tm) * (1.0 + c->visc) / (tm + c->visc);
                    double d = b->dist[gt];
                    b->qkw[G(b, is, js, ks)] = 1600.0 * vsl / b->r[g0] / c->prm.reynolds / (d * d) - b->qkw[gt];
                }
                double pres, tres;
                int i2 = it - l3[0], j2 = jt - l3[1], k2 = kt - l3[2];
                if (c->prm.twall > 0.0) {
                    pres = (4.0 * P_(b,it,jt,kt) - P_(b,i2,j2,k2)) / 3.0;
                    if (pres < PMIN_LIMIT) pres = P_(b,it,jt,kt);
                    tres = c->prm.twall / c->prm.tref;
                } else {
                    pres = (4.0 * P_(b,it,jt,kt) - P_(b,i2,j2,k2)) / 3.0;
                    tres = (4.0 * T_(b,it,jt,kt) - T_(b,i2,j2,k2)) / 3.0;
                    if (pres < PMIN_LIMIT) pres = P_(b,it,jt,kt);
                    if (tres < TMIN_LIMIT) tres = T_(b,it,jt,kt);
                }
                double dres = moocp * pres / tres;
                SETP(b, i, j, k, dres, 0.0, 0.0, 0.0, pres);
                /* BC_face_dif_turbulence: face value = mean of first ghost and first interior node */
                for (int v = 0; v < nturb(c); v++)
                    tvar(b, v)[G(b, i, j, k)] = 0.5 * (tvar(b, v)[G(b, it, jt, kt)] + tvar(b, v)[G(b, is, js, ks)]);
                is += l3[0]; js += l3[1]; ks += l3[2]; it -= l3[0]; jt -= l3[1]; kt -= l3[2];
                SETP(b, is, js, ks, R_(b,it,jt,kt), -U_(b,it,jt,kt), -V_(b,it,jt,kt), -W_(b,it,jt,kt), P_(b,it,jt,kt));
                is += l3[0]; js += l3[1]; ks += l3[2]; it -= l3[0]; jt -= l3[1]; kt -= l3[2];
                SETP(b, is, js, ks, -R_(b,it,jt,kt), -U_(b,it,jt,kt), -V_(b,it,jt,kt), -W_(b,it,jt,kt), P_(b,it,jt,kt));
            }
}

/* BC.f90:227-457 boundary_3 (method = 1, 3-D branch) */
static void bc_symmetry(o_case *c, oblock *b, oregion *g)
{
    const int *l3 = g->s_lr3d;
    int ni = b->ni, nj = b->nj, nk = b->nk;
    int dims[3] = {ni, nj, nk};
    if (dims[g->s_nd - 1] <= NIJK2D) {
        /* BC.f90:247-322: 2-D direction, reflect w */
        for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
            for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
                for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                    int it = i - l3[0], jt = j - l3[1], kt = k - l3[2];
                    it = it < 1 ? 1 : (it > ni ? ni : it);
                    jt = jt < 1 ? 1 : (jt > nj ? nj : jt);
                    kt = kt < 1 ? 1 : (kt > nk ? nk : kt);
                    double rm = R_(b,it,jt,kt), um = U_(b,it,jt,kt), vm = V_(b,it,jt,kt);
                    double wm = -W_(b,it,jt,kt), pm = P_(b,it,jt,kt);
                    SETP(b, i + l3[0], j + l3[1], k + l3[2], rm, um, vm, wm, pm);
                    for (int v = 0; v < nturb(c); v++)
                        tvar(b, v)[G(b, i + l3[0], j + l3[1], k + l3[2])] = tvar(b, v)[G(b, it, jt, kt)];
                    SETP(b, i + 2 * l3[0], j + 2 * l3[1], k + 2 * l3[2], rm, um, vm, 0.0, pm);
                    SETP(b, i + 3 * l3[0], j + 3 * l3[1], k + 3 * l3[2], rm, um, vm, 0.0, pm);
                }
        dif_average(c, b, g);
        return;
    }
    int is = (g->s_st[0] + g->s_ed[0]) / 2, js = (g->s_st[1] + g->s_ed[1]) / 2, ks = (g->s_st[2] + g->s_ed[2]) / 2;
    double nx, ny, nz;
    getnxyz(b, is, js, ks, l3[0], l3[1], l3[2], &nx, &ny, &nz);
    double sss1 = sqrt(nx * nx + ny * ny + nz * nz);
    nx = nx / sss1; ny = ny / sss1; nz = nz / sss1;
    for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
        for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
            for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
                for (int l = 1; l <= 3; l++) {
                    int it = i - l * l3[0], jt = j - l * l3[1], kt = k - l * l3[2];
                    double s = 2.0 * (nx * U_(b,it,jt,kt) + ny * V_(b,it,jt,kt) + nz * W_(b,it,jt,kt));
                    SETP(b, i + l * l3[0], j + l * l3[1], k + l * l3[2], R_(b,it,jt,kt),
                         U_(b,it,jt,kt) - s * nx, V_(b,it,jt,kt) - s * ny, W_(b,it,jt,kt) - s * nz, P_(b,it,jt,kt));
                    /* turblence.f90:2308-2319 BC_symmetry_turbulence on the first ghost */
                    if (TURB(c) && l == 1)
                        for (int v = 0; v < nturb(c); v++)
                            tvar(b, v)[G(b, i + l3[0], j + l3[1], k + l3[2])] = tvar(b, v)[G(b, it, jt, kt)];
                }
    dif_average(c, b, g);
}

/* sub_tgh.f90:123-186 boundary_all -> BC.f90:169-210 boundary_sequence */
int o_boundary_all(o_case *c)
{
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int ib = 0; ib < b->nregions; ib++) {
            oregion *g = &b->reg[b->bcindex[ib]];
            switch (g->bctype) {
            case 2: if (c->prm.nvis == 1) bc_viscous_wall(c, b, g); else bc_inviscid_wall(c, b, g); break;
            case 3: bc_symmetry(c, b, g); break;
            case 4: bc_farfield(c, b, g); break;
            case 5: bc_fixed(c, b, g); break;
            case 6: bc_outflow(c, b, g); break;
            default:
                if (g->bctype < 0) bc_interface(c, b, g);
                else return -1;
            }
        }
    }
    return 0;
}

/* ========================================================= step routines */
/* sub_tgh.f90:1-32 initial_q_dq_c_t -> RHS.f90:171-240 get_c_only / RHS.f90:307-409
 * get_c_t_only, GasModel.f90:219-235 compute_visl_ns, RHS.f90:75-166 get_q_only,
 * PreProcess.f90:1189-1205 set_dq_to_0 */
int o_initial_q_dq_c_t(o_case *c)
{
    double gama = c->prm.gama, moo2 = c->prm.moo * c->prm.moo;
    int nvis = c->prm.nvis;
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        int ni = b->ni, nj = b->nj, nk = b->nk;
        /* interior */
        #pragma omp parallel for collapse(2) num_threads(c->nthreads) if (c->nthreads > 1)
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
            long q = G(b, i, j, k);
            double a2 = gama * b->p[q] / b->r[q];
            b->c[q] = sqrt(a2);
            if (nvis == 1) b->t[q] = moo2 * a2;
        }
        /* ghost layers: c on layer 1; t on layers 1..3 (layer 3 repeats layer 2's a2) */
        for (int dir = 0; dir < 3; dir++) {
            int n1 = dir == 0 ? nj : ni, n2 = dir == 1 ? nk : (dir == 0 ? nk : nj);
            int nd = dir == 0 ? ni : (dir == 1 ? nj : nk);
            for (int b2 = 1; b2 <= n2; b2++) for (int b1 = 1; b1 <= n1; b1++) for (int side = 0; side < 2; side++) {
                int l1 = side == 0 ? 0 : nd + 1, l2 = side == 0 ? -1 : nd + 2, l3 = side == 0 ? -2 : nd + 3;
                int idx1[3], idx2[3], idx3[3];
                for (int m = 0; m < 3; m++) { idx1[m] = idx2[m] = idx3[m] = 0; }
                if (dir == 0) { idx1[0] = l1; idx2[0] = l2; idx3[0] = l3; idx1[1] = idx2[1] = idx3[1] = b1; idx1[2] = idx2[2] = idx3[2] = b2; }
                if (dir == 1) { idx1[1] = l1; idx2[1] = l2; idx3[1] = l3; idx1[0] = idx2[0] = idx3[0] = b1; idx1[2] = idx2[2] = idx3[2] = b2; }
                if (dir == 2) { idx1[2] = l1; idx2[2] = l2; idx3[2] = l3; idx1[0] = idx2[0] = idx3[0] = b1; idx1[1] = idx2[1] = idx3[1] = b2; }
                long q1 = G(b, idx1[0], idx1[1], idx1[2]);
                double a2 = gama * b->p[q1] / b->r[q1];
                b->c[q1] = sqrt(a2);
                if (nvis == 1) {
                    b->t[q1] = moo2 * a2;
                    long q2 = G(b, idx2[0], idx2[1], idx2[2]);
                    a2 = gama * b->p[q2] / b->r[q2];
                    b->t[q2] = moo2 * a2;
                    long q3 = G(b, idx3[0], idx3[1], idx3[2]);
                    /* RHS.f90:351-353: the i faces reuse layer 2's a2 for layer 3;
                     * RHS.f90:374-377,398-401: j and k faces use |r| of layer 3 */
                    if (dir != 0) a2 = gama * b->p[q3] / fabs(b->r[q3]);
                    b->t[q3] = moo2 * a2;
                }
            }
        }
        if (nvis == 1) {
            /* GasModel.f90:219-235, cmethod = 0 -> 1..n */
            #pragma omp parallel for collapse(2) num_threads(c->nthreads) if (c->nthreads > 1)
            for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
                long q = G(b, i, j, k);
                double tm = b->t[q];
                b->visl[q] = tm * sqrt(tm) * (1.0 + c->visc) / (tm + c->visc);
            }
        }
        /* get_q_only: interior with prim_to_q, ghost layer 1 with prim_to_q_bc */
        #pragma omp parallel for collapse(2) num_threads(c->nthreads) if (c->nthreads > 1)
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
            long q = G(b, i, j, k);
            double prim[5] = {b->r[q], b->u[q], b->v[q], b->w[q], b->p[q]};
            double em = getem(prim, gama);
            b->q[5*q+0] = prim[0]; b->q[5*q+1] = prim[0] * prim[1]; b->q[5*q+2] = prim[0] * prim[2];
            b->q[5*q+3] = prim[0] * prim[3]; b->q[5*q+4] = em;
        }
        for (int dir = 0; dir < 3; dir++) {
            int nd = dir == 0 ? ni : (dir == 1 ? nj : nk);
            int na = dir == 0 ? nj : ni, nbb = dir == 2 ? nj : nk;
            for (int b2 = 1; b2 <= nbb; b2++) for (int b1 = 1; b1 <= na; b1++) for (int side = 0; side < 2; side++) {
                int l = side == 0 ? 0 : nd + 1;
                int i = dir == 0 ? l : b1, j = dir == 1 ? l : (dir == 0 ? b1 : b2), k = dir == 2 ? l : b2;
                long q = G(b, i, j, k);
                double prim[5] = {b->r[q], b->u[q], b->v[q], b->w[q], b->p[q]};
                b->q[5*q+0] = prim[0];
                if (prim[0] < 0) prim[0] = -prim[0];
                double em = getem(prim, gama);
                b->q[5*q+1] = prim[0] * prim[1]; b->q[5*q+2] = prim[0] * prim[2];
                b->q[5*q+3] = prim[0] * prim[3]; b->q[5*q+4] = em;
            }
        }
        /* set_dq_to_0 over -1..n+1 (ghost layers 2..3 are never written either) */
        memset(b->dq, 0, sizeof(double) * 5 * b->g3);
    }
    return 0;
}

/* Convect.f90:1338-1461 spectinv + LHS.f90:117-138 spectrum_tgh + LHS.f90:1343-1377 localdt0 */
static void spect_point(const oblock *b, int i, int j, int k, int ii, int jj, int kk)
{
    oblock *bb = (oblock *)b;
    long q = G(b, i, j, k);
    double um = b->u[q], vm = b->v[q], wm = b->w[q], cm = b->c[q];
    double kx = F3(b,kcx,ii,jj,kk), ky = F3(b,kcy,ii,jj,kk), kz = F3(b,kcz,ii,jj,kk);
    double ex = F3(b,etx,ii,jj,kk), ey = F3(b,ety,ii,jj,kk), ez = F3(b,etz,ii,jj,kk);
    double cx = F3(b,ctx,ii,jj,kk), cy = F3(b,cty,ii,jj,kk), cz = F3(b,ctz,ii,jj,kk);
    double cita = kx * um + ky * vm + kz * wm + 0.0;
    double citb = ex * um + ey * vm + ez * wm + 0.0;
    double citc = cx * um + cy * vm + cz * wm + 0.0;
    double cgma = sqrt(kx * kx + ky * ky + kz * kz);
    double cgmb = sqrt(ex * ex + ey * ey + ez * ez);
    double cgmc = sqrt(cx * cx + cy * cy + cz * cz);
    bb->sra[q] = fabs(cita) + cm * cgma;
    bb->srb[q] = fabs(citb) + cm * cgmb;
    bb->src[q] = fabs(citc) + cm * cgmc;
}

int o_timestep(o_case *c)
{
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        int ni = b->ni, nj = b->nj, nk = b->nk;
        #pragma omp parallel for collapse(2) num_threads(c->nthreads) if (c->nthreads > 1)
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++)
            spect_point(b, i, j, k, i, j, k);
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) { spect_point(b, 0, j, k, 1, j, k); spect_point(b, ni + 1, j, k, ni, j, k); }
        for (int k = 1; k <= nk; k++) for (int i = 1; i <= ni; i++) { spect_point(b, i, 0, k, i, 1, k); spect_point(b, i, nj + 1, k, i, nj, k); }
        for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) { spect_point(b, i, j, 0, i, j, 1); spect_point(b, i, j, nk + 1, i, j, nk); }
        if (c->prm.nvis == 1 && c->prm.csrv > 0.0) {
            extern int o_spectvis(o_case *c, int nb);
            o_spectvis(c, nb);
        }
        /* localdt0 */
        #pragma omp parallel for collapse(2) num_threads(c->nthreads) if (c->nthreads > 1)
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) for (int i = 1; i <= ni; i++) {
            long q = G(b, i, j, k);
            double ra = b->sra[q], rb = b->srb[q], rc = b->src[q];
            double rv = b->srva[q] + b->srvb[q] + b->srvc[q];
            double rabc = ra + rb + rc + rv;
            if (c->prm.timedt_rate > SMALL) {
                double a = c->prm.cfl / rabc, bq = c->prm.timedt_rate / b->vol[q];
                b->dtdt[q] = a < bq ? a : bq;
            } else {
                b->dtdt[q] = c->prm.cfl / rabc;
            }
        }
    }
    return 0;
}

/* The metric of direction dir (0 kc, 1 et, 2 ct) at G index q */
static void metric_of(const oblock *b, int dir, long q, double *o)
{
    if (dir == 0) { o[0] = b->kcx[q]; o[1] = b->kcy[q]; o[2] = b->kcz[q]; }
    if (dir == 1) { o[0] = b->etx[q]; o[1] = b->ety[q]; o[2] = b->etz[q]; }
    if (dir == 2) { o[0] = b->ctx[q]; o[1] = b->cty[q]; o[2] = b->ctz[q]; }
}

/*
 * Convect.f90:180-297 (and the same for j, k): the four ghost metrics of a
 * pencil of direction dir at either end, trx1(1:3, -3..0) and (1:3, n+1..n+4)
 * (index i+3).  The region that owns the pencil's end node is the last one in
 * set_bc_index order whose box holds it on that face (fill_bc_flag,
 * PreProcess.f90:820-857).  An interface gives the target block's metrics of
 * its own normal direction, 1..4 nodes inside its face from the image node,
 * signed by the target side (t_lr on the low end, -t_lr on the high end); any
 * other region repeats the end node's metric.  kt and vol of the ghosts are
 * not read by any scheme (kt = 0 on a static grid; vol is not used).
 */
static void ghost_metrics(const o_case *c, const oblock *b, int dir, int ia, int ib, double *trx1)
{
    int nd = dir == 0 ? b->ni : (dir == 1 ? b->nj : b->nk);
    for (int side = 0; side < 2; side++) {
        int l = side == 0 ? 1 : nd;
        int i = dir == 0 ? l : ia, j = dir == 1 ? l : (dir == 0 ? ia : ib), k = dir == 2 ? l : ib;
        const oregion *own = NULL;
        for (int n = 0; n < b->nregions; n++) {
            const oregion *g = &b->reg[b->bcindex[n]];
            if (g->s_nd != dir + 1 || g->s_lr != (side == 0 ? -1 : 1)) continue;
            if (i < g->s_st[0] || i > g->s_ed[0] || j < g->s_st[1] || j > g->s_ed[1] ||
                k < g->s_st[2] || k > g->s_ed[2]) continue;
            own = g;
        }
        for (int i1 = 1; i1 <= 4; i1++) {
            int slot = side == 0 ? 1 - i1 : nd + i1;
            double *o = &trx1[5 * (slot + 3)];
            if (own && own->bctype < 0) {
                const oblock *bt = &c->blk[own->nbt - 1];
                long r = rbox(own, i, j, k);
                int it = own->image[r] - i1 * own->t_lr3d[0];
                int jt = own->jmage[r] - i1 * own->t_lr3d[1];
                int kt = own->kmage[r] - i1 * own->t_lr3d[2];
                double mt[3];
                metric_of(bt, own->t_nd - 1, G(bt, it, jt, kt), mt);
                double sg = side == 0 ? own->t_lr : -own->t_lr;
                for (int m = 0; m < 3; m++) o[m] = sg * mt[m];
            } else {
                for (int m = 0; m < 3; m++) o[m] = trx1[5 * (l + 3) + m];
            }
            o[3] = 0.0;
            o[4] = 0.0;
        }
    }
}

static inline double dmin(double a, double b) { return b < a ? b : a; }

/*
 * Convect.f90:78-132: the pressure ratios min(p_l/p_r, p_r/p_l) between
 * neighbours along each direction, rp[d] over -1..n+1 (index (i+1) +
 * (ni+3)((j+1) + (nj+3)(k+1))), the two ghost entries at a physical face
 * (ghost density below small) replaced by the first interior one.  The
 * reference's loops never write direction 2 and 3 at i = -1, 1 and 3 at
 * j = -1, 1 and 2 at k = -1, yet its rpmin reads them at the faces 0: there it
 * reads whatever the heap held.  Here they hold +inf, so that the minimum is
 * over the ratios that exist -- the reference's result whenever that memory
 * held no smaller value.
 */
static double *pressure_ratios(const oblock *b)
{
    int ni = b->ni, nj = b->nj, nk = b->nk;
    long sx = ni + 3, sy = (long)(ni + 3) * (nj + 3), sz = sy * (nk + 3);
    double *rp = (double *)malloc(sizeof(double) * 3 * sz);
    for (long q = 0; q < 3 * sz; q++) rp[q] = HUGE_VAL;
#define RP(d,i,j,k) rp[(d) * sz + ((i) + 1) + sx * ((j) + 1) + sy * ((k) + 1)]
    for (int k = 0; k <= nk + 1; k++)
        for (int j = 0; j <= nj + 1; j++) {
            for (int i = -1; i <= ni + 1; i++) {
                double pl = b->p[G(b, i, j, k)], pr = b->p[G(b, i + 1, j, k)];
                RP(0, i, j, k) = dmin(pl / pr, pr / pl);
            }
            if (b->r[G(b, -2, j, k)] < SMALL) { RP(0, 0, j, k) = RP(0, 1, j, k); RP(0, -1, j, k) = RP(0, 1, j, k); }
            if (b->r[G(b, ni + 3, j, k)] < SMALL) {
                RP(0, ni, j, k) = RP(0, ni - 1, j, k);
                RP(0, ni + 1, j, k) = RP(0, ni - 1, j, k);
            }
        }
    for (int k = 0; k <= nk + 1; k++)
        for (int i = 0; i <= ni + 1; i++) {
            for (int j = -1; j <= nj + 1; j++) {
                double pl = b->p[G(b, i, j, k)], pr = b->p[G(b, i, j + 1, k)];
                RP(1, i, j, k) = dmin(pl / pr, pr / pl);
            }
            if (b->r[G(b, i, -2, k)] < SMALL) { RP(1, i, 0, k) = RP(1, i, 1, k); RP(1, i, -1, k) = RP(1, i, 1, k); }
            if (b->r[G(b, i, nj + 3, k)] < SMALL) {
                RP(1, i, nj, k) = RP(1, i, nj - 1, k);
                RP(1, i, nj + 1, k) = RP(1, i, nj - 1, k);
            }
        }
    for (int j = 0; j <= nj + 1; j++)
        for (int i = 0; i <= ni + 1; i++) {
            for (int k = -1; k <= nk + 1; k++) {
                double pl = b->p[G(b, i, j, k)], pr = b->p[G(b, i, j, k + 1)];
                RP(2, i, j, k) = dmin(pl / pr, pr / pl);
            }
            if (b->r[G(b, i, j, -2)] < SMALL) { RP(2, i, j, 0) = RP(2, i, j, 1); RP(2, i, j, -1) = RP(2, i, j, 1); }
            if (b->r[G(b, i, j, nk + 3)] < SMALL) {
                RP(2, i, j, nk) = RP(2, i, j, nk - 1);
                RP(2, i, j, nk + 1) = RP(2, i, j, nk - 1);
            }
        }
    return rp;
}

/* Convect.f90:169-177 (i), 335-343 (j), 499-507 (k): the face's own ratio,
 * then the least of the four ratios of each other direction around it */
static void rpmin_line(const oblock *b, const double *rp, int dir, int ia, int ib, double *rpmin)
{
    int ni = b->ni, nj = b->nj, nk = b->nk;
    long sx = ni + 3, sy = (long)(ni + 3) * (nj + 3), sz = sy * (nk + 3);
    int nd = dir == 0 ? ni : (dir == 1 ? nj : nk);
    for (int f = 0; f <= nd; f++) {
        double v;
        if (dir == 0) {
            int i = f, j = ia, k = ib;
            v = RP(0, i, j, k);
            v = dmin(dmin(dmin(dmin(RP(1, i, j, k), RP(1, i, j + 1, k)), RP(1, i - 1, j, k)), RP(1, i - 1, j + 1, k)), v);
            v = dmin(dmin(dmin(dmin(RP(2, i, j, k), RP(2, i, j, k + 1)), RP(2, i - 1, j, k)), RP(2, i - 1, j, k + 1)), v);
        } else if (dir == 1) {
            int i = ia, j = f, k = ib;
            v = RP(1, i, j, k);
            v = dmin(dmin(dmin(dmin(RP(0, i, j, k), RP(0, i + 1, j, k)), RP(0, i, j - 1, k)), RP(0, i + 1, j - 1, k)), v);
            v = dmin(dmin(dmin(dmin(RP(2, i, j, k), RP(2, i, j, k + 1)), RP(2, i, j - 1, k)), RP(2, i, j - 1, k + 1)), v);
        } else {
            int i = ia, j = ib, k = f;
            v = RP(2, i, j, k);
            v = dmin(dmin(dmin(dmin(RP(0, i, j, k), RP(0, i + 1, j, k)), RP(0, i, j, k - 1)), RP(0, i + 1, j, k - 1)), v);
            v = dmin(dmin(dmin(dmin(RP(1, i, j, k), RP(1, i, j + 1, k)), RP(1, i, j, k - 1)), RP(1, i, j + 1, k - 1)), v);
        }
        rpmin[f] = v;
    }
#undef RP
}

/* RHS.f90:35-72 r_h_s_invis -> Convect.f90:2-649 invcode/inviscd3d.  The
 * pencils of one direction write disjoint nodes; with several threads each
 * thread takes whole pencils with line buffers of its own. */
int o_rhs_invis(o_case *c)
{
    int scheme = c->prm.nscheme, flux = c->prm.nflux;
    double gama = c->prm.gama, efix = c->prm.efix;
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        int ni = b->ni, nj = b->nj, nk = b->nk;
        int nmax = ni > nj ? ni : nj; nmax = nmax > nk ? nmax : nk;
        double *rp = flux == 4 ? pressure_ratios(b) : NULL;
        for (int dir = 0; dir < 3; dir++) {
            int nd = dir == 0 ? ni : (dir == 1 ? nj : nk);
            if (nd <= NIJK2D) continue;
            int na = dir == 0 ? nj : ni;          /* inner loop extent */
            int nbo = dir == 2 ? nj : nk;         /* outer loop extent */
            long nfall = 0;
            int err = 0;
            #pragma omp parallel num_threads(c->nthreads) if (c->nthreads > 1) reduction(+:nfall) reduction(|:err)
            {
                double *ql = (double *)calloc(6 * (nmax + 6) * 2, sizeof(double));
                double *tr = (double *)calloc(5 * (nmax + 8), sizeof(double));   /* nodes -3..n+4 */
                double *rpm = (double *)calloc(nmax + 1, sizeof(double));
                double *fc = (double *)calloc(5 * nmax, sizeof(double));
                #pragma omp for collapse(2) schedule(static)
                for (int ob = 1; ob <= nbo; ob++) for (int ib = 1; ib <= na; ib++) {
                    for (int l = -2; l <= nd + 3; l++) {
                        int i = dir == 0 ? l : ib, j = dir == 1 ? l : (dir == 0 ? ib : ob), k = dir == 2 ? l : ob;
                        long q = G(b, i, j, k);
                        double *o = &ql[6 * (l + 2)];
                        o[0] = b->r[q]; o[1] = b->u[q]; o[2] = b->v[q]; o[3] = b->w[q]; o[4] = b->p[q]; o[5] = gama;
                    }
                    for (int l = 1; l <= nd; l++) {
                        int i = dir == 0 ? l : ib, j = dir == 1 ? l : (dir == 0 ? ib : ob), k = dir == 2 ? l : ob;
                        long q = G(b, i, j, k);
                        double *o = &tr[5 * (l + 3)];
                        metric_of(b, dir, q, o);
                        o[3] = 0.0; o[4] = b->vol[q];
                    }
                    if (scheme == 45 || scheme == 47) ghost_metrics(c, b, dir, ib, ob, tr);
                    if (flux == 4) rpmin_line(b, rp, dir, ib, ob, rpm);
                    int ne = o_wcns_line_ex(scheme, flux, nd, nmax, efix, gama, tr, ql, rpm, fc);
                    if (ne < 0) { err = 1; continue; }
                    nfall += ne;
                    for (int l = 1; l <= nd; l++) {
                        int i = dir == 0 ? l : ib, j = dir == 1 ? l : (dir == 0 ? ib : ob), k = dir == 2 ? l : ob;
                        long q = G(b, i, j, k);
                        for (int m = 0; m < 5; m++) b->dq[5 * q + m] = b->dq[5 * q + m] + fc[5 * (l - 1) + m];
                    }
                }
                free(ql); free(tr); free(rpm); free(fc);
            }
            c->n_fallback += nfall;
            if (err) { free(rp); return -2; }
        }
        free(rp);
    }
    return 0;
}

/* parallel.f90:3020-3164 communicate_dq_npp (single process: boundary_match_dq_2pm, _3pp) */
int o_communicate_dq(o_case *c)
{
    /* gather pre-averaging dq/vol of every multi-point (parallel.f90:3028-3036) */
    int tot = 0;
    for (int g = 0; g < c->ngroups; g++) tot += c->groups[g].n;
    double *npp = (double *)malloc(sizeof(double) * 5 * (tot > 0 ? tot : 1));
    int o = 0;
    for (int g = 0; g < c->ngroups; g++)
        for (int s = 0; s < c->groups[g].n; s++, o++) {
            opnt p = c->groups[g].p[s];
            oblock *b = &c->blk[p.nb - 1];
            double vol_p = F3(b,vol,p.i,p.j,p.k);
            for (int m = 1; m <= 5; m++) npp[5 * o + m - 1] = DQ(b,m,p.i,p.j,p.k) / vol_p;
        }
    /* boundary_match_dq_2pm: ordered, in place */
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int r = 0; r < b->nregions; r++) {
            oregion *g = &b->reg[r];
            if (g->bctype >= 0) continue;
            oblock *bt = &c->blk[g->nbt - 1];
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
                for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
                    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                        long ob = rbox(g, i, j, k);
                        int it = g->image[ob], jt = g->jmage[ob], kt = g->kmage[ob];
                        double vol_p = F3(b,vol,i,j,k) / F3(bt,vol,it,jt,kt);
                        for (int m = 1; m <= 5; m++)
                            DQ(b,m,i,j,k) = 0.5 * (DQ(bt,m,it,jt,kt) * vol_p + DQ(b,m,i,j,k));
                    }
        }
    }
    /* boundary_match_dq_3pp */
    o = 0;
    for (int g = 0; g < c->ngroups; g++) {
        int npp_n = c->groups[g].n;
        double dqq[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < npp_n; s++)
            for (int m = 0; m < 5; m++) dqq[m] = dqq[m] + npp[5 * (o + s) + m];
        for (int m = 0; m < 5; m++) dqq[m] = dqq[m] / (npp_n * 1.0);
        for (int s = 0; s < npp_n; s++) {
            opnt p = c->groups[g].p[s];
            oblock *b = &c->blk[p.nb - 1];
            double vol_p = F3(b,vol,p.i,p.j,p.k);
            for (int m = 1; m <= 5; m++) DQ(b,m,p.i,p.j,p.k) = dqq[m - 1] * vol_p;
        }
        o += npp_n;
    }
    free(npp);
    return 0;
}

/* parallel.f90:3166-3326 communicate_pv_npp */
int o_communicate_pv(o_case *c)
{
    int tot = 0;
    for (int g = 0; g < c->ngroups; g++) tot += c->groups[g].n;
    double *npp = (double *)malloc(sizeof(double) * 6 * (tot > 0 ? tot : 1));
    int o = 0;
    for (int g = 0; g < c->ngroups; g++)
        for (int s = 0; s < c->groups[g].n; s++, o++) {
            opnt p = c->groups[g].p[s];
            oblock *b = &c->blk[p.nb - 1];
            long q = G(b, p.i, p.j, p.k);
            double vol_p = 1.0;
            npp[6*o+0] = b->r[q] / vol_p; npp[6*o+1] = b->u[q] / vol_p; npp[6*o+2] = b->v[q] / vol_p;
            npp[6*o+3] = b->w[q] / vol_p; npp[6*o+4] = b->p[q] / vol_p; npp[6*o+5] = b->t[q] / vol_p;
        }
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int r = 0; r < b->nregions; r++) {
            oregion *g = &b->reg[r];
            if (g->bctype >= 0) continue;
            oblock *bt = &c->blk[g->nbt - 1];
            for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
                for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
                    for (int i = g->s_st[0]; i <= g->s_ed[0]; i++) {
                        long ob = rbox(g, i, j, k);
                        long qt = G(bt, g->image[ob], g->jmage[ob], g->kmage[ob]), qs = G(b, i, j, k);
                        b->r[qs] = 0.5 * (bt->r[qt] + b->r[qs]);
                        b->u[qs] = 0.5 * (bt->u[qt] + b->u[qs]);
                        b->v[qs] = 0.5 * (bt->v[qt] + b->v[qs]);
                        b->w[qs] = 0.5 * (bt->w[qt] + b->w[qs]);
                        b->p[qs] = 0.5 * (bt->p[qt] + b->p[qs]);
                        b->t[qs] = 0.5 * (bt->t[qt] + b->t[qs]);
                    }
        }
    }
    o = 0;
    for (int g = 0; g < c->ngroups; g++) {
        int n = c->groups[g].n;
        double d[6] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < n; s++) for (int m = 0; m < 6; m++) d[m] = d[m] + npp[6 * (o + s) + m];
        for (int m = 0; m < 6; m++) d[m] = d[m] / (n * 1.0);
        for (int s = 0; s < n; s++) {
            opnt p = c->groups[g].p[s];
            oblock *b = &c->blk[p.nb - 1];
            long q = G(b, p.i, p.j, p.k);
            b->r[q] = d[0] * 1.0; b->u[q] = d[1] * 1.0; b->v[q] = d[2] * 1.0;
            b->w[q] = d[3] * 1.0; b->p[q] = d[4] * 1.0; b->t[q] = d[5] * 1.0;
        }
        o += n;
    }
    free(npp);
    return 0;
}

/* ============================================================== implicit */
/* LHS.f90:602-645 getcoed (nl = nm = 5, no dual time) */
static double getcoed(const oblock *b, int nvis, long q)
{
    double rad_ns = b->sra[q] + b->srb[q] + b->src[q];
    if (nvis >= 1) rad_ns = rad_ns + b->srva[q] + b->srvb[q] + b->srvc[q];
    double ct = 1.0 / b->dtdt[q];
    return 1.0 / (ct + 1.0 * 1.0 * rad_ns);
}

static inline void getprim(const oblock *b, int i, int j, int k, double *prim)
{
    long q = G(b, i, j, k);
    prim[0] = b->r[q]; prim[1] = b->u[q]; prim[2] = b->v[q]; prim[3] = b->w[q]; prim[4] = b->p[q];
}

/* Sub1.f90:676-717 getrkec_mml (clamps 0 -> 1 and n+1 -> n) */
static inline void getrkec(const oblock *b, int i, int j, int k, int dir, double *g)
{
    if (i == 0) i = 1;
    if (i == b->ni + 1) i = b->ni;
    if (j == 0) j = 1;
    if (j == b->nj + 1) j = b->nj;
    if (k == 0) k = 1;
    if (k == b->nk + 1) k = b->nk;
    long q = G(b, i, j, k);
    if (dir == 0) { g[0] = b->kcx[q]; g[1] = b->kcy[q]; g[2] = b->kcz[q]; }
    if (dir == 1) { g[0] = b->etx[q]; g[1] = b->ety[q]; g[2] = b->etz[q]; }
    if (dir == 2) { g[0] = b->ctx[q]; g[1] = b->cty[q]; g[2] = b->ctz[q]; }
    g[3] = 0.0;
}

/* LHS.f90:647-678 getrarbrc (method = 1) */
static inline void getrarbrc(const oblock *b, int i, int j, int k, int d, double *ra, double *rb, double *rc)
{
    int im = i + d, jm = j + d, km = k + d;
    im = im > b->ni ? b->ni : im; jm = jm > b->nj ? b->nj : jm; km = km > b->nk ? b->nk : km;
    im = im < 1 ? 1 : im; jm = jm < 1 ? 1 : jm; km = km < 1 ? 1 : km;
    *ra = b->sra[G(b, im, j, k)];
    *rb = b->srb[G(b, i, jm, k)];
    *rc = b->src[G(b, i, j, km)];
}

/* LHS.f90:381-478 lusgs_l */
static void lusgs_l(o_case *c, oblock *b)
{
    double gama = c->prm.gama;
    int nvis = c->prm.nvis;
    for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
        long q = G(b, i, j, k);
        double coed = getcoed(b, nvis, q);
        double dqi[5], dqj[5], dqk[5], pi[5], pj[5], pk[5], gi[4], gj[4], gk[4], de[5], df[5], dg[5];
        for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i-1,j,k); dqj[m-1] = DQ(b,m,i,j-1,k); dqk[m-1] = DQ(b,m,i,j,k-1); }
        getprim(b, i - 1, j, k, pi); getprim(b, i, j - 1, k, pj); getprim(b, i, j, k - 1, pk);
        getrkec(b, i - 1, j, k, 0, gi); getrkec(b, i, j - 1, k, 1, gj); getrkec(b, i, j, k - 1, 2, gk);
        double ra, rb, rc;
        getrarbrc(b, i, j, k, -1, &ra, &rb, &rc);
        o_mxdq_std(pi, gi, dqi, de, ra, 1, gama);
        o_mxdq_std(pj, gj, dqj, df, rb, 1, gama);
        o_mxdq_std(pk, gk, dqk, dg, rc, 1, gama);
        double rhs0[5];
        for (int m = 0; m < 5; m++) rhs0[m] = 1.0 * (de[m] + df[m] + dg[m]);
        if (nvis > 0) {
            double rva = b->srva[G(b,i-1,j,k)], rvb = b->srvb[G(b,i,j-1,k)], rvc = b->srvc[G(b,i,j,k-1)];
            for (int m = 1; m <= 5; m++)
                rhs0[m-1] = rhs0[m-1] + 0.5 * (rva * DQ(b,m,i-1,j,k) + rvb * DQ(b,m,i,j-1,k) + rvc * DQ(b,m,i,j,k-1));
        }
        for (int m = 1; m <= 5; m++) DQ(b,m,i,j,k) = (-DQ(b,m,i,j,k) + rhs0[m-1]) * coed;
        (void)q;
    }
}

/* LHS.f90:480-577 lusgs_u */
static void lusgs_u(o_case *c, oblock *b)
{
    double gama = c->prm.gama;
    int nvis = c->prm.nvis;
    for (int k = b->nk; k >= 1; k--) for (int j = b->nj; j >= 1; j--) for (int i = b->ni; i >= 1; i--) {
        long q = G(b, i, j, k);
        double coed = getcoed(b, nvis, q);
        double dqi[5], dqj[5], dqk[5], pi[5], pj[5], pk[5], gi[4], gj[4], gk[4], de[5], df[5], dg[5];
        for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i+1,j,k); dqj[m-1] = DQ(b,m,i,j+1,k); dqk[m-1] = DQ(b,m,i,j,k+1); }
        getprim(b, i + 1, j, k, pi); getprim(b, i, j + 1, k, pj); getprim(b, i, j, k + 1, pk);
        getrkec(b, i + 1, j, k, 0, gi); getrkec(b, i, j + 1, k, 1, gj); getrkec(b, i, j, k + 1, 2, gk);
        double ra, rb, rc;
        getrarbrc(b, i, j, k, 1, &ra, &rb, &rc);
        o_mxdq_std(pi, gi, dqi, de, ra, -1, gama);
        o_mxdq_std(pj, gj, dqj, df, rb, -1, gama);
        o_mxdq_std(pk, gk, dqk, dg, rc, -1, gama);
        double rhs0[5];
        for (int m = 0; m < 5; m++) rhs0[m] = 1.0 * (de[m] + df[m] + dg[m]);
        if (nvis > 0) {
            double rva = b->srva[G(b,i+1,j,k)], rvb = b->srvb[G(b,i,j+1,k)], rvc = b->srvc[G(b,i,j,k+1)];
            for (int m = 1; m <= 5; m++)
                rhs0[m-1] = rhs0[m-1] - 0.5 * (rva * DQ(b,m,i+1,j,k) + rvb * DQ(b,m,i,j+1,k) + rvc * DQ(b,m,i,j,k+1));
        }
        for (int m = 1; m <= 5; m++) DQ(b,m,i,j,k) = DQ(b,m,i,j,k) - rhs0[m-1] * coed;
        (void)q;
    }
}

/* LHS.f90:1777-1866 gs_pr_l and LHS.f90:1956-2046 gs_pr_u (one body, sweep direction sgn) */
static void gs_pr_node(o_case *c, oblock *b, int i, int j, int k)
{
    double gama = c->prm.gama;
    long q = G(b, i, j, k);
    double coed = getcoed(b, c->prm.nvis, q);
    double dqi[5], dqj[5], dqk[5], pi[5], pj[5], pk[5], gi[4], gj[4], gk[4];
    double del[5], dfl[5], dgl[5], der[5], dfr[5], dgr[5], ra, rb, rc;
    for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i-1,j,k); dqj[m-1] = DQ(b,m,i,j-1,k); dqk[m-1] = DQ(b,m,i,j,k-1); }
    if (i == 1) memset(dqi, 0, sizeof dqi);
    if (j == 1) memset(dqj, 0, sizeof dqj);
    if (k == 1) memset(dqk, 0, sizeof dqk);
    getprim(b, i - 1, j, k, pi); getprim(b, i, j - 1, k, pj); getprim(b, i, j, k - 1, pk);
    getrkec(b, i - 1, j, k, 0, gi); getrkec(b, i, j - 1, k, 1, gj); getrkec(b, i, j, k - 1, 2, gk);
    getrarbrc(b, i, j, k, -1, &ra, &rb, &rc);
    o_mxdq_std(pi, gi, dqi, del, ra, 1, gama);
    o_mxdq_std(pj, gj, dqj, dfl, rb, 1, gama);
    o_mxdq_std(pk, gk, dqk, dgl, rc, 1, gama);
    for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i+1,j,k); dqj[m-1] = DQ(b,m,i,j+1,k); dqk[m-1] = DQ(b,m,i,j,k+1); }
    if (i == b->ni) memset(dqi, 0, sizeof dqi);
    if (j == b->nj) memset(dqj, 0, sizeof dqj);
    if (k == b->nk) memset(dqk, 0, sizeof dqk);
    getprim(b, i + 1, j, k, pi); getprim(b, i, j + 1, k, pj); getprim(b, i, j, k + 1, pk);
    getrkec(b, i + 1, j, k, 0, gi); getrkec(b, i, j + 1, k, 1, gj); getrkec(b, i, j, k + 1, 2, gk);
    getrarbrc(b, i, j, k, 1, &ra, &rb, &rc);
    o_mxdq_std(pi, gi, dqi, der, ra, -1, gama);
    o_mxdq_std(pj, gj, dqj, dfr, rb, -1, gama);
    o_mxdq_std(pk, gk, dqk, dgr, rc, -1, gama);
    for (int m = 1; m <= 5; m++) {
        double rhs0 = 1.0 * (del[m-1] + dfl[m-1] + dgl[m-1] - der[m-1] - dfr[m-1] - dgr[m-1]);
        DQ(b,m,i,j,k) = (rhs0 - b->rhs_nb[5 * q + m - 1]) * coed;
    }
}

/* CIC.f90:2-17 set_nonconnect_boundary_dq_0 -> CIC.f90:482-506 boundary_dq_to_0 */
static void dq_to_zero_physical(oblock *b)
{
    for (int r = 0; r < b->nregions; r++) {
        oregion *g = &b->reg[r];
        if (g->bctype <= 0) continue;
        for (int i = g->s_st[0]; i <= g->s_ed[0]; i++)
            for (int j = g->s_st[1]; j <= g->s_ed[1]; j++)
                for (int k = g->s_st[2]; k <= g->s_ed[2]; k++)
                    for (int m = 1; m <= 5; m++) DQ(b,m,i,j,k) = 0.0;
    }
}

/* Sub2.f90:603-796 update_limited */
static int update_limited(o_case *c, oblock *b)
{
    double gama = c->prm.gama, gamam1 = gama - 1.0;
    int n_count = 0;
    for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
        long q = G(b, i, j, k);
        double qtem[5];
        for (int m = 0; m < 5; m++) qtem[m] = b->q[5 * q + m] + b->dq[5 * q + m];
        double rm = qtem[0], rm1 = 1.0 / rm;
        double um = qtem[1] * rm1, vm = qtem[2] * rm1, wm = qtem[3] * rm1, em = qtem[4];
        double pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
        double dp = -0.3 + fabs((pm - b->p[q]) / b->p[q]);
        double falpha = 1. / (1. + 2. * (0.0 > dp ? 0.0 : dp));
        for (int m = 0; m < 5; m++) { qtem[m] = b->q[5 * q + m] + b->dq[5 * q + m] * falpha; b->q[5 * q + m] = qtem[m]; }
        rm = qtem[0]; rm1 = 1.0 / rm;
        um = qtem[1] * rm1; vm = qtem[2] * rm1; wm = qtem[3] * rm1; em = qtem[4];
        pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
        if (pm <= PMIN_LIMIT || rm <= RMIN_LIMIT || pm >= PMAX_LIMIT || rm >= RMAX_LIMIT) {
            n_count++;
            rm = 0.0; um = 0.0; vm = 0.0; wm = 0.0; pm = 0.0;
            int np = 0;
            for (int ii = -1; ii <= 1; ii++) for (int jj = -1; jj <= 1; jj++) for (int kk = -1; kk <= 1; kk++) {
                if (abs(ii) + abs(jj) + abs(kk) == 1) {
                    np++;
                    long g2 = G(b, i + ii, j + jj, k + kk);
                    rm = rm + b->r[g2]; um = um + b->u[g2]; vm = vm + b->v[g2]; wm = wm + b->w[g2]; pm = pm + b->p[g2];
                }
            }
            double rnp = (double)np;
            rm = rm / rnp; um = um / rnp; vm = vm / rnp; wm = wm / rnp; pm = pm / rnp;
            em = pm / (gama - 1.0) + 0.5 * rm * (um * um + vm * vm + wm * wm);
            b->q[5*q+0] = rm; b->q[5*q+1] = rm * um; b->q[5*q+2] = rm * vm; b->q[5*q+3] = rm * wm; b->q[5*q+4] = em;
        }
        b->r[q] = rm; b->u[q] = um; b->v[q] = vm; b->w[q] = wm; b->p[q] = pm;
    }
    c->n_repair += n_count;
    if (n_count > b->ni * b->nj * b->nk / 64 || n_count >= 500) return -3;
    return 0;
}

/* lusgs_l / lusgs_u node bodies for the hyperplane order (the same
 * arithmetic as the lexicographic loops above, one node at a time) */
static void lusgs_l_node(o_case *c, oblock *b, int i, int j, int k)
{
    double gama = c->prm.gama;
    int nvis = c->prm.nvis;
    long q = G(b, i, j, k);
    double coed = getcoed(b, nvis, q);
    double dqi[5], dqj[5], dqk[5], pi[5], pj[5], pk[5], gi[4], gj[4], gk[4], de[5], df[5], dg[5];
    for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i-1,j,k); dqj[m-1] = DQ(b,m,i,j-1,k); dqk[m-1] = DQ(b,m,i,j,k-1); }
    getprim(b, i - 1, j, k, pi); getprim(b, i, j - 1, k, pj); getprim(b, i, j, k - 1, pk);
    getrkec(b, i - 1, j, k, 0, gi); getrkec(b, i, j - 1, k, 1, gj); getrkec(b, i, j, k - 1, 2, gk);
    double ra, rb, rc;
    getrarbrc(b, i, j, k, -1, &ra, &rb, &rc);
    o_mxdq_std(pi, gi, dqi, de, ra, 1, gama);
    o_mxdq_std(pj, gj, dqj, df, rb, 1, gama);
    o_mxdq_std(pk, gk, dqk, dg, rc, 1, gama);
    double rhs0[5];
    for (int m = 0; m < 5; m++) rhs0[m] = 1.0 * (de[m] + df[m] + dg[m]);
    if (nvis > 0) {
        double rva = b->srva[G(b,i-1,j,k)], rvb = b->srvb[G(b,i,j-1,k)], rvc = b->srvc[G(b,i,j,k-1)];
        for (int m = 1; m <= 5; m++)
            rhs0[m-1] = rhs0[m-1] + 0.5 * (rva * DQ(b,m,i-1,j,k) + rvb * DQ(b,m,i,j-1,k) + rvc * DQ(b,m,i,j,k-1));
    }
    for (int m = 1; m <= 5; m++) DQ(b,m,i,j,k) = (-DQ(b,m,i,j,k) + rhs0[m-1]) * coed;
}

static void lusgs_u_node(o_case *c, oblock *b, int i, int j, int k)
{
    double gama = c->prm.gama;
    int nvis = c->prm.nvis;
    long q = G(b, i, j, k);
    double coed = getcoed(b, nvis, q);
    double dqi[5], dqj[5], dqk[5], pi[5], pj[5], pk[5], gi[4], gj[4], gk[4], de[5], df[5], dg[5];
    for (int m = 1; m <= 5; m++) { dqi[m-1] = DQ(b,m,i+1,j,k); dqj[m-1] = DQ(b,m,i,j+1,k); dqk[m-1] = DQ(b,m,i,j,k+1); }
    getprim(b, i + 1, j, k, pi); getprim(b, i, j + 1, k, pj); getprim(b, i, j, k + 1, pk);
    getrkec(b, i + 1, j, k, 0, gi); getrkec(b, i, j + 1, k, 1, gj); getrkec(b, i, j, k + 1, 2, gk);
    double ra, rb, rc;
    getrarbrc(b, i, j, k, 1, &ra, &rb, &rc);
    o_mxdq_std(pi, gi, dqi, de, ra, -1, gama);
    o_mxdq_std(pj, gj, dqj, df, rb, -1, gama);
    o_mxdq_std(pk, gk, dqk, dg, rc, -1, gama);
    double rhs0[5];
    for (int m = 0; m < 5; m++) rhs0[m] = 1.0 * (de[m] + df[m] + dg[m]);
    if (nvis > 0) {
        double rva = b->srva[G(b,i+1,j,k)], rvb = b->srvb[G(b,i,j+1,k)], rvc = b->srvc[G(b,i,j,k+1)];
        for (int m = 1; m <= 5; m++)
            rhs0[m-1] = rhs0[m-1] - 0.5 * (rva * DQ(b,m,i+1,j,k) + rvb * DQ(b,m,i,j+1,k) + rvc * DQ(b,m,i,j,k+1));
    }
    for (int m = 1; m <= 5; m++) DQ(b,m,i,j,k) = DQ(b,m,i,j,k) - rhs0[m-1] * coed;
}

/* One sweep in hyperplane order: planes s = i+j+k ascending (fwd) or
 * descending; the nodes of one plane read only nodes of the neighbouring
 * planes, so they run in parallel.  kind 0 lusgs_l, 1 lusgs_u, 2 gs_pr. */
static void sweep_planes(o_case *c, oblock *b, int kind, int fwd)
{
    int ni = b->ni, nj = b->nj, nk = b->nk;
    for (int t = 0; t <= ni + nj + nk - 3; t++) {
        int s = fwd ? 3 + t : ni + nj + nk - t;
        #pragma omp parallel for collapse(2) schedule(static) num_threads(c->nthreads)
        for (int k = 1; k <= nk; k++) for (int j = 1; j <= nj; j++) {
            int i = s - j - k;
            if (i < 1 || i > ni) continue;
            if (kind == 0) lusgs_l_node(c, b, i, j, k);
            else if (kind == 1) lusgs_u_node(c, b, i, j, k);
            else gs_pr_node(c, b, i, j, k);
        }
    }
}

/* LHS.f90:1-34 l_h_s_tgh -> LHS.f90:1559-1587 GS_PR_Single */
int o_lhs(o_case *c)
{
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        dq_to_zero_physical(b);
        for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++) {
            long q = G(b, i, j, k);
            for (int m = 0; m < 5; m++) b->rhs_nb[5 * q + m] = b->dq[5 * q + m];
        }
        if (c->nthreads > 1) {
            sweep_planes(c, b, 0, 1);
            sweep_planes(c, b, 1, 0);
            sweep_planes(c, b, 2, 1);
            sweep_planes(c, b, 2, 0);
        } else {
            lusgs_l(c, b);
            lusgs_u(c, b);
            for (int k = 1; k <= b->nk; k++) for (int j = 1; j <= b->nj; j++) for (int i = 1; i <= b->ni; i++)
                gs_pr_node(c, b, i, j, k);
            for (int k = b->nk; k >= 1; k--) for (int j = b->nj; j >= 1; j--) for (int i = b->ni; i >= 1; i--)
                gs_pr_node(c, b, i, j, k);
        }
        dq_to_zero_physical(b);
        int e = update_limited(c, b);
        if (e) return e;
    }
    return 0;
}

/* PostProcess.f90:1195-1347 residual_curve (ndualtst = 0) */
int o_residual(o_case *c, double *rms, double *resmax, int loc[5])
{
    double dres_add = 0.0, resmax_total = 0.0;
    long np = 0;
    int nl = 5;
    loc[0] = 2; loc[1] = 2; loc[2] = 2; loc[3] = 1; loc[4] = 1;
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        double res = 0.0, rmax = 0.0;
        int nres[4] = {2, 2, 2, 1};
        for (int k = 2; k <= b->nk - 1; k++) for (int j = 2; j <= b->nj - 1; j++) for (int i = 2; i <= b->ni - 1; i++) {
            long q = G(b, i, j, k);
            double dtime1 = 1.0 / (b->dtdt[q] * b->vol[q]);
            for (int m = 1; m <= nl; m++) {
                double dq0 = fabs(b->dq[5 * q + m - 1]);
                double dresm = dq0 * dtime1;
                if (dresm > rmax) { rmax = dresm; nres[0] = i; nres[1] = j; nres[2] = k; nres[3] = m; }
                res = res + dresm * dresm;
            }
        }
        dres_add = dres_add + res;
        np += (long)(b->ni - 2) * (b->nj - 2) * (b->nk - 2) * nl;
        if (resmax_total < rmax) {
            resmax_total = rmax;
            loc[0] = nres[0]; loc[1] = nres[1]; loc[2] = nres[2]; loc[3] = nres[3]; loc[4] = nb + 1;
        }
    }
    *rms = sqrt(dres_add / np);
    *resmax = resmax_total;
    return 0;
}

int o_rhs_vis(o_case *c);

/* Main.f90:79-109 one pseudo-time step */
int o_step(o_case *c, int nsteps, double *rms_hist)
{
    for (int s = 0; s < nsteps; s++) {
        int e;
        if ((e = o_boundary_all(c))) return e;
        if ((e = o_initial_q_dq_c_t(c))) return e;
        if ((e = o_timestep(c))) return e;
        if (c->prm.nvis == 1) {
            if ((e = o_rhs_vis(c))) return e;
            if ((e = o_communicate_dq(c))) return e;
        }
        if ((e = o_rhs_invis(c))) return e;
        if ((e = o_communicate_dq(c))) return e;
        if ((e = o_lhs(c))) return e;
        if ((e = o_communicate_pv(c))) return e;
        double rms, rmax;
        int loc[5];
        o_residual(c, &rms, &rmax, loc);
        if (rms_hist) rms_hist[s] = rms;
    }
    return 0;
}

/* views for wcns_oracle_vis.c (field pointers of one block, namelist values) */
/* oblock_view: wcns_oracle_priv.h */

int o_view_block(o_case *c, int nb, oblock_view *v)
{
    if (nb < 0 || nb >= c->nblocks) return -1;
    oblock *b = &c->blk[nb];
    v->ni = b->ni; v->nj = b->nj; v->nk = b->nk; v->g3 = b->g3;
    v->x = b->x; v->y = b->y; v->z = b->z;
    v->r = b->r; v->u = b->u; v->v = b->v; v->w = b->w; v->p = b->p; v->t = b->t; v->c = b->c;
    v->q = b->q; v->dq = b->dq; v->rhs_nb = b->rhs_nb;
    v->kcx = b->kcx; v->kcy = b->kcy; v->kcz = b->kcz; v->etx = b->etx; v->ety = b->ety; v->etz = b->etz;
    v->ctx = b->ctx; v->cty = b->cty; v->ctz = b->ctz; v->vol = b->vol;
    v->sra = b->sra; v->srb = b->srb; v->src = b->src; v->srva = b->srva; v->srvb = b->srvb; v->srvc = b->srvc;
    v->dtdt = b->dtdt; v->visl = b->visl; v->vist = b->vist;
    v->qke = b->qke; v->dist = b->dist; v->qkw = b->qkw;
    return 0;
}

double o_view_muoo(const o_case *c) { return c->muoo; }
void o_view_turb_inflow(const o_case *c, double *koo, double *omgoo) { *koo = c->koo; *omgoo = c->omgoo; }

void o_view_params(const o_case *c, o_params *p, double *visc)
{
    *p = c->prm;
    *visc = c->visc;
}


/* ======================================================= wall force */
/* PostProcess.f90:1684-1796 getforce (4-triangle Heron quadrature of a wall quad) */
static double aver3(double a, double b, double c) { return (a + b + c) / 3.0; }
static double halen(double a, double b, double c)
{
    double p = 0.5 * (a + b + c);
    double S = p * (p - a) * (p - b) * (p - c);
    if (S < 0) return 0.0;
    return sqrt(S);
}
static double lenth(double x1, double y1, double z1, double x2, double y2, double z2)
{
    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    return sqrt(dx * dx + dy * dy + dz * dz);
}
static void getforce(const double *P /*12: p1..p4 xyz*/, const double *X /*12*/, const double *ref, double *d /*6*/)
{
    double p5[3], x5[3];
    for (int a = 0; a < 3; a++) {
        p5[a] = 0.25 * (P[a] + P[3 + a] + P[6 + a] + P[9 + a]);
        x5[a] = 0.25 * (X[a] + X[3 + a] + X[6 + a] + X[9 + a]);
    }
#define XX(n) X[3 * (n)], X[3 * (n) + 1], X[3 * (n) + 2]
    double r12 = lenth(XX(0), XX(1)), r23 = lenth(XX(1), XX(2)), r34 = lenth(XX(2), XX(3)), r41 = lenth(XX(3), XX(0));
    double r15 = lenth(XX(0), x5[0], x5[1], x5[2]), r25 = lenth(XX(1), x5[0], x5[1], x5[2]);
    double r35 = lenth(XX(2), x5[0], x5[1], x5[2]), r45 = lenth(XX(3), x5[0], x5[1], x5[2]);
#undef XX
    double sarea[4] = {halen(r12, r15, r25), halen(r23, r25, r35), halen(r34, r35, r45), halen(r41, r45, r15)};
    const int ia[4] = {0, 1, 2, 3}, ib[4] = {1, 2, 3, 0};
    for (int a = 0; a < 6; a++) d[a] = 0.0;
    for (int t = 0; t < 4; t++) {
        double pc[3], xc[3];
        for (int a = 0; a < 3; a++) {
            pc[a] = aver3(P[3 * ia[t] + a], P[3 * ib[t] + a], p5[a]);
            xc[a] = aver3(X[3 * ia[t] + a], X[3 * ib[t] + a], x5[a]);
        }
        double fx = pc[0] * sarea[t], fy = pc[1] * sarea[t], fz = pc[2] * sarea[t];
        double mx = (xc[1] - ref[1]) * fz - (xc[2] - ref[2]) * fy;
        double my = (xc[2] - ref[2]) * fx - (xc[0] - ref[0]) * fz;
        double mz = (xc[0] - ref[0]) * fy - (xc[1] - ref[1]) * fx;
        d[0] = d[0] + fx; d[1] = d[1] + fy; d[2] = d[2] + fz;
        d[3] = d[3] + mx; d[4] = d[4] + my; d[5] = d[5] + mz;
    }
}

/* Sub1.f90:383-480 getuvwtder: second-order one-sided at block faces, central inside */
static void uvwtder(const oblock *b, int i, int j, int k, int dir, double *d /*4*/)
{
    int n = dir == 0 ? b->ni : (dir == 1 ? b->nj : b->nk);
    int l = dir == 0 ? i : (dir == 1 ? j : k);
    const double aa = 1.5, bb = -2.0, cc = 0.5;
#define AT(a, o) ((dir) == 0 ? (a)[G(b, (o), j, k)] : ((dir) == 1 ? (a)[G(b, i, (o), k)] : (a)[G(b, i, j, (o))]))
    const double *fs[4] = {b->u, b->v, b->w, b->t};
    for (int m = 0; m < 4; m++) {
        const double *a = fs[m];
        if (l == 1) d[m] = -(aa * AT(a, l) + bb * AT(a, 2) + cc * AT(a, 3));
        else if (l == n) d[m] = aa * AT(a, l) + bb * AT(a, l - 1) + cc * AT(a, l - 2);
        else d[m] = 0.5 * (AT(a, l + 1) - AT(a, l - 1));
    }
#undef AT
}

int o_force(o_case *c, double out[9])
{
    double cfx = 0, cfy = 0, cfz = 0, cmx = 0, cmy = 0, cmz = 0;
    double ref[3] = {c->prm.xref, c->prm.yref, c->prm.zref};
    for (int nb = 0; nb < c->nblocks; nb++) {
        oblock *b = &c->blk[nb];
        for (int r = 0; r < b->nregions; r++) {
            oregion *g = &b->reg[r];
            if (!(g->bctype == 2 || g->bctype / 10 == 2)) continue;
            int bi = g->s_ed[0] - g->s_st[0] + 1, bj = g->s_ed[1] - g->s_st[1] + 1, bk = g->s_ed[2] - g->s_st[2] + 1;
            double *pn3 = (double *)malloc(sizeof(double) * 3 * bi * bj * bk);
#define PN(a, i, j, k) pn3[3 * (((i) - g->s_st[0]) + bi * (((j) - g->s_st[1]) + bj * ((k) - g->s_st[2]))) + (a)]
            for (int is = g->s_st[0]; is <= g->s_ed[0]; is++)
                for (int js = g->s_st[1]; js <= g->s_ed[1]; js++)
                    for (int ks = g->s_st[2]; ks <= g->s_ed[2]; ks++) {
                        double nx, ny, nz;
                        getnxyz(b, is, js, ks, g->s_lr3d[0], g->s_lr3d[1], g->s_lr3d[2], &nx, &ny, &nz);
                        long q = G(b, is, js, ks);
                        double pn = 2.0 * (b->p[q] - c->poo);
                        PN(0, is, js, ks) = g->s_lr * nx * pn;
                        PN(1, is, js, ks) = g->s_lr * ny * pn;
                        PN(2, is, js, ks) = g->s_lr * nz * pn;
                        if (c->prm.nvis == 1) {
                            double dk[4], de[4], dc[4];
                            uvwtder(b, is, js, ks, 0, dk); uvwtder(b, is, js, ks, 1, de); uvwtder(b, is, js, ks, 2, dc);
                            double kx = b->kcx[q], ky = b->kcy[q], kz = b->kcz[q], ex = b->etx[q], ey = b->ety[q];
                            double ez = b->etz[q], cx = b->ctx[q], cy = b->cty[q], cz = b->ctz[q], vj = b->vol[q];
                            /* Sub1.f90:251-270 getduvwdxyz */
                            double dudx = (kx * dk[0] + ex * de[0] + cx * dc[0]) / vj, dudy = (ky * dk[0] + ey * de[0] + cy * dc[0]) / vj;
                            double dudz = (kz * dk[0] + ez * de[0] + cz * dc[0]) / vj;
                            double dvdx = (kx * dk[1] + ex * de[1] + cx * dc[1]) / vj, dvdy = (ky * dk[1] + ey * de[1] + cy * dc[1]) / vj;
                            double dvdz = (kz * dk[1] + ez * de[1] + cz * dc[1]) / vj;
                            double dwdx = (kx * dk[2] + ex * de[2] + cx * dc[2]) / vj, dwdy = (ky * dk[2] + ey * de[2] + cy * dc[2]) / vj;
                            double dwdz = (kz * dk[2] + ez * de[2] + cz * dc[2]) / vj;
                            /* Sub1.f90:181-198 stress, c23 = 0.66666666667 as written */
                            double vis = b->visl[q], c23 = 0.66666666667;
                            double txx = vis * c23 * (2.0 * dudx - dvdy - dwdz);
                            double tyy = vis * c23 * (2.0 * dvdy - dwdz - dudx);
                            double tzz = vis * c23 * (2.0 * dwdz - dudx - dvdy);
                            double txy = vis * (dudy + dvdx), txz = vis * (dudz + dwdx), tyz = vis * (dvdz + dwdy);
                            double tyx = txy, tzx = txz, tzy = tyz;
                            double pv = 2.0 * g->s_lr;
                            PN(0, is, js, ks) = PN(0, is, js, ks) - pv * (nx * txx + ny * tyx + nz * tzx) * c->re;
                            PN(1, is, js, ks) = PN(1, is, js, ks) - pv * (nx * txy + ny * tyy + nz * tzy) * c->re;
                            PN(2, is, js, ks) = PN(2, is, js, ks) - pv * (nx * txz + ny * tyz + nz * tzz) * c->re;
                        }
                    }
            int id = 1 - abs(g->s_lr3d[0]), jd = 1 - abs(g->s_lr3d[1]), kd = 1 - abs(g->s_lr3d[2]);
            for (int is = g->s_st[0]; is <= g->s_ed[0] - id; is++)
                for (int js = g->s_st[1]; js <= g->s_ed[1] - jd; js++)
                    for (int ks = g->s_st[2]; ks <= g->s_ed[2] - kd; ks++) {
                        int q4[4][3];
                        q4[0][0] = is; q4[0][1] = js; q4[0][2] = ks;
                        if (id == 0) { int t[3][3] = {{is, js + 1, ks}, {is, js + 1, ks + 1}, {is, js, ks + 1}}; memcpy(&q4[1], t, sizeof t); }
                        if (jd == 0) { int t[3][3] = {{is + 1, js, ks}, {is + 1, js, ks + 1}, {is, js, ks + 1}}; memcpy(&q4[1], t, sizeof t); }
                        if (kd == 0) { int t[3][3] = {{is + 1, js, ks}, {is + 1, js + 1, ks}, {is, js + 1, ks}}; memcpy(&q4[1], t, sizeof t); }
                        double P[12], X[12], d[6];
                        for (int n = 0; n < 4; n++) {
                            for (int a = 0; a < 3; a++) P[3 * n + a] = PN(a, q4[n][0], q4[n][1], q4[n][2]);
                            long xi = (q4[n][0] - 1) + (long)b->ni * ((q4[n][1] - 1) + (long)b->nj * (q4[n][2] - 1));
                            X[3 * n] = b->x[xi]; X[3 * n + 1] = b->y[xi]; X[3 * n + 2] = b->z[xi];
                        }
                        getforce(P, X, ref, d);
                        cfx = cfx + d[0]; cfy = cfy + d[1]; cfz = cfz + d[2];
                        cmx = cmx + d[3]; cmy = cmy + d[4]; cmz = cmz + d[5];
                    }
#undef PN
            free(pn3);
        }
    }
    double scf = 1.0 / c->prm.sref, vcm = 1.0 / (c->prm.sref * c->prm.lfref);
    int w = 2 - c->prm.nwholefield;
    cfx = w * cfx * scf; cfy = w * cfy * scf; cfz = w * cfz * scf;
    cmx = w * cmx * vcm; cmy = w * cmy * vcm; cmz = w * cmz * vcm;
    double sina = sin(c->attack_r), cosa = cos(c->attack_r), sinb = sin(c->sideslip_r), cosb = cos(c->sideslip_r);
    double cl = cfy * cosa - cfx * sina;
    double cd = cosb * (cfy * sina + cfx * cosa) + sinb * cfz;
    double xcp = (cfy >= 0 ? 1.0 : -1.0) * cmz / (fabs(cfy) + SMALL);
    out[0] = cfx; out[1] = cfy; out[2] = cfz; out[3] = cmx; out[4] = cmy; out[5] = cmz;
    out[6] = cd; out[7] = cl; out[8] = xcp;
    return 0;
}
