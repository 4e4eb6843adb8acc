// wcns_device.hpp -- device-side data layout and per-node arithmetic shared by
// the gfx950 kernels.  Every formula keeps the reference's operation order
// (Fortran evaluates a-b+c as (a-b)+c) and is compiled with
// -ffp-contract=off, so that the FP64 path rounds exactly like the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wcns_ad.hpp"

namespace wcns {

// src/Globalvar.f90:215-229
constexpr double SMALL = 1.0e-38;
constexpr double SML_SSS = 1.0e-30;
constexpr double PMIN_LIMIT = 1.0e-8;
constexpr double PMAX_LIMIT = 1.0e+3;
constexpr double RMIN_LIMIT = 1.0e-8;
constexpr double RMAX_LIMIT = 1.0e+3;
constexpr double TMIN_LIMIT = 1.0e-8;

// One block's place in the field pools.  Node (i,j,k), i in -2..ni+3, lives at
// off + (i+2) + sx*(j+2) + sy*(k+2).
struct DBlock {
    int ni, nj, nk, pad;
    long sx, sy, off;
};

__host__ __device__ inline long gidx(const DBlock &b, int i, int j, int k)
{
    return b.off + (long)(i + 2) + b.sx * (long)(j + 2) + b.sy * (long)(k + 2);
}

// Scalars every kernel may need (freestream, namelist values).
struct DParams {
    double gama, efix, cfl, timedt_rate, moo, csrv, re, reynolds, visc, prl, prt;
    double roo, uoo, voo, woo, poo, coo, twall_nd;
    int nvis, nscheme, nflux, twall_fixed;
    // turbulence model: turb = 0 laminar, 1 Spalart-Allmaras, 2 SST k-omega;
    // freestream nu~ (SA), k, omega and mu_t (SST), namelist turbulence values
    double muoo, koo, omgoo, vtoo, kmaxlim, timedt_turb, xk, xb;
    int turb;
    int inv_waves;   // register budget of the inviscid walk: 0 by launch size, 1 or 2 waves per SIMD
};

// Base pointers of the field pools.  Conservative 5-vectors (rhs, dq) are
// variable-major: component m of node g at ptr[m*total + g].
struct DFields {
    double *r, *u, *v, *w, *p;           // primitive, current buffer
    double *rn, *un, *vn, *wn, *pn;      // primitive, update target
    double *t;
    const double *kcx, *kcy, *kcz, *etx, *ety, *etz, *ctx, *cty, *ctz, *vol;
    double *rhs, *dq;
    double *dtdt, *sra, *srb, *src, *srva, *srvb, *srvc, *visl;
    double *coed;                       // LU-SGS diagonal 1/(1/dt + sum of radii), per node
    float *dq32;                        // FP32 sweep operand (mixed mode)
    unsigned long long *counters;       // 0 fallbacks, 1 repairs, 2 roe, 4 repair abort
    double *qke, *vist, *dist;          // SA nu~ / SST k, mu_t, wall distance (null when laminar)
    double *qkw;                        // SST omega (null otherwise)
    // schemes 45 / 47: four ghost metrics at each end of every pencil
    // (12 doubles a pencil end; gmoff: each block's first); flux 4: the
    // pressure ratios of the three directions (3 x total)
    const double *gmet;
    const long *gmoff;
    double *rpres;
    long total;
};

// Sub1.f90:750-780 getnxyz_mml: unit normal of the face whose index offset is l3
__device__ inline void getnxyz(const DFields &F, long g, const int *l3, double &nx, double &ny, double &nz)
{
    double x = 0, y = 0, z = 0;
    if (l3[0]) { x = x + F.kcx[g]; y = y + F.kcy[g]; z = z + F.kcz[g]; }
    if (l3[1]) { x = x + F.etx[g]; y = y + F.ety[g]; z = z + F.etz[g]; }
    if (l3[2]) { x = x + F.ctx[g]; y = y + F.cty[g]; z = z + F.ctz[g]; }
    double s = sqrt(x * x + y * y + z * z);
    s = s < SML_SSS ? SML_SSS : s;
    nx = x / s; ny = y / s; nz = z / s;
}

// ------------------------------------------------------------- gas model
// The residual arithmetic below is generic over the scalar type T: double for
// the solver, D1 (wcns_ad.hpp) for the tangent and adjoint kernels.

// GasModel.f90:358-373 getem
template <typename T>
__device__ inline T getem(T rm, T um, T vm, T wm, T pm, double gama)
{
    return pm / (gama - 1.0) + 0.5 * rm * (um * um + vm * vm + wm * wm);
}

// flux.f90:17-77 flux_SW_pn (kt = 0: static grid, PreProcess.f90:1649-1657)
template <typename T>
__device__ inline void flux_sw_pn(const T *prim, T nx, T ny, T nz, T *f, double efix, double gama, double npn)
{
    T rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    T u2 = 0.5 * (um * um + vm * vm + wm * wm);
    T hint = gama / (gama - 1.0) * pm / rm;
    T eint = hint - pm / rm;
    T em = eint + u2;
    T cm = sqrt(gama * pm / rm);
    T ct = nx * um + ny * vm + nz * wm + 0.0;
    T cgm = sqrt(nx * nx + ny * ny + nz * nz);
    cgm = cgm < SML_SSS ? T(SML_SSS) : cgm;
    T l1 = ct, l4 = ct + cm * cgm, l5 = ct - cm * cgm;
    T eps = efix * efix * cgm * cgm;
    l1 = 0.5 * (l1 + npn * sqrt(l1 * l1 + eps));
    l4 = 0.5 * (l4 + npn * sqrt(l4 * l4 + eps));
    l5 = 0.5 * (l5 + npn * sqrt(l5 * l5 + eps));
    T cgm1 = 1.0 / cgm;
    T nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1;
    T cta = (ct - 0.0) * cgm1;
    T c2r = cm * cm / gama;
    T x1 = c2r * (2.0 * l1 - l4 - l5) / (2.0 * cm * cm);
    T x2 = c2r * (l4 - l5) / (2.0 * cm);
    T hm = em + pm / rm;
    f[0] = (l1 - x1) * rm;
    f[1] = (l1 * um - x1 * um + nxa * x2) * rm;
    f[2] = (l1 * vm - x1 * vm + nya * x2) * rm;
    f[3] = (l1 * wm - x1 * wm + nza * x2) * rm;
    f[4] = (l1 * em - x1 * hm + cta * x2) * rm;
}

// flux.f90:2-15 flux_SW
template <typename T>
__device__ inline void flux_sw(const T *ql, const T *qr, T nx, T ny, T nz, T *f, double efix, double gama)
{
    T fl[5], fr[5];
    flux_sw_pn(ql, nx, ny, nz, fl, efix, gama, 1.0);
    flux_sw_pn(qr, nx, ny, nz, fr, efix, gama, -1.0);
#pragma unroll
    for (int m = 0; m < 5; m++) f[m] = fl[m] + fr[m];
}

// Convect.f90:1245-1272 flux (kt = 0)
template <typename T>
__device__ inline void flux_phys(const T *prim, double gama, T nx, T ny, T nz, T *f)
{
    T rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    T em = getem(rm, um, vm, wm, pm, gama);
    T cit = nx * um + ny * vm + nz * wm + 0.0;
    f[0] = rm * cit;
    f[1] = f[0] * um + nx * pm;
    f[2] = f[0] * vm + ny * pm;
    f[3] = f[0] * wm + nz * pm;
    f[4] = (em + pm) * cit - pm * 0.0;
}

// flux.f90:116-244 flux_Roe.  Returns false when c^2 <= 0 (the reference stops).
template <typename T>
__device__ inline bool flux_roe(const T *qlin, const T *qrin, T nx, T ny, T nz, T *f, double efix, double gama)
{
    T ql[5], qr[5], fl[5], fr[5], dq[5], qroe[5], df[5];
#pragma unroll
    for (int m = 0; m < 5; m++) { ql[m] = qlin[m]; qr[m] = qrin[m]; }
    flux_phys(ql, gama, nx, ny, nz, fl);
    flux_phys(qr, gama, nx, ny, nz, fr);
#pragma unroll
    for (int m = 0; m < 5; m++) dq[m] = qr[m] - ql[m];
    const double gama1 = gama - 1.0;
    const double gama2 = gama / gama1;
    ql[4] = gama2 * ql[4] / ql[0] + 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
    qr[4] = gama2 * qr[4] / qr[0] + 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    qroe[0] = sqrt(ql[0] * qr[0]);
    T rrl = sqrt(qr[0] / ql[0]);
    T rrl_1 = rrl + 1.0;
#pragma unroll
    for (int m = 1; m < 5; m++) qroe[m] = (ql[m] + qr[m] * rrl) / rrl_1;
    T rm = qroe[0], um = qroe[1], vm = qroe[2], wm = qroe[3], hm = qroe[4];
    T c2 = (hm - 0.5 * (um * um + vm * vm + wm * wm)) * gama1;
    bool ok = c2 > 0.0;
    T c = sqrt(c2);
    T cta = um * nx + vm * ny + wm * nz + 0.0;
    T cgm = sqrt(nx * nx + ny * ny + nz * nz);
    cgm = cgm < SML_SSS ? T(SML_SSS) : cgm;
    T cgm1 = 1.0 / cgm;
    nx = nx * cgm1;
    ny = ny * cgm1;
    nz = nz * cgm1;
    T ccgm = c * cgm;
    T l1 = fabs(cta - ccgm), l4 = fabs(cta), l5 = fabs(cta + ccgm);
    T l12 = l1 * l1, l42 = l4 * l4, l52 = l5 * l5;
    T deti = efix * cgm;
    T det2 = deti * deti;
    l1 = sqrt(l12 + det2);
    l4 = sqrt(l42 + det2);
    l5 = sqrt(l52 + det2);
    T dcta = dq[1] * nx + dq[2] * ny + dq[3] * nz;
    T cta1 = cta * cgm1;
    T rodcta = qroe[0] * dcta;
    T dp_c2 = dq[4] / c2;
    T rodctac = rodcta / c;
    T a1 = l4 * (dq[0] - dp_c2);
    T a2 = l5 * (dp_c2 + rodctac) * 0.5;
    T a3 = l1 * (dp_c2 - rodctac) * 0.5;
    T a4 = a1 + a2 + a3;
    T a5 = c * (a2 - a3);
    T a6 = l4 * (rm * dq[1] - nx * rodcta);
    T a7 = l4 * (rm * dq[2] - ny * rodcta);
    T a8 = l4 * (rm * dq[3] - nz * rodcta);
    df[0] = a4;
    df[1] = um * a4 + nx * a5 + a6;
    df[2] = vm * a4 + ny * a5 + a7;
    df[3] = wm * a4 + nz * a5 + a8;
    df[4] = hm * a4 + cta1 * a5 + um * a6 + vm * a7 + wm * a8 - c2 * a1 / gama1;
#pragma unroll
    for (int m = 0; m < 5; m++) f[m] = 0.5 * (fr[m] + fl[m] - df[m]);
    return ok;
}

// flux.f90:410-501 flux_vanleer_pn (kt = 0), npn = +1 (F+) or -1 (F-).
// Returns false where gama p / rho < 0 (the reference stops, flux.f90:437-441).
template <typename T>
__device__ inline bool flux_vl_pn(const T *prim, T nx, T ny, T nz, T *f, double gama, int npn)
{
    T rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    T u2 = 0.5 * (um * um + vm * vm + wm * wm);
    T hint = gama / (gama - 1.0) * pm / rm;
    T hm = hint + u2;
    T c2 = gama * pm / rm;
    bool ok = !(c2 < 0.0);
    T cm = sqrt(c2);
    T ct = nx * um + ny * vm + nz * wm + 0.0;
    T cgm = sqrt(nx * nx + ny * ny + nz * nz) + SML_SSS;
    T cgm1 = 1.0 / cgm;
    T nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1, cta = ct * cgm1;
    T bmvl = cta / cm;
    if (fabs(bmvl) < 1.0) {
        T ctcm, f0;
        if (npn == 1) {
            ctcm = (-cta + 2.0 * cm) / gama;
            f0 = 0.25 * cgm * rm * cm * ((bmvl + 1.0) * (bmvl + 1.0));
        } else {
            ctcm = (-cta - 2.0 * cm) / gama;
            f0 = -0.25 * cgm * rm * cm * ((bmvl - 1.0) * (bmvl - 1.0));
        }
        f[0] = f0;
        f[1] = f0 * (um + nxa * ctcm);
        f[2] = f0 * (vm + nya * ctcm);
        f[3] = f0 * (wm + nza * ctcm);
        f[4] = f0 * hm;
    } else if ((npn == 1 && bmvl >= 1.0) || (npn == -1 && bmvl <= -1.0)) {
        T rmct = rm * cta;
        f[0] = cgm * rmct;
        f[1] = cgm * (rmct * um + pm * nxa);
        f[2] = cgm * (rmct * vm + pm * nya);
        f[3] = cgm * (rmct * wm + pm * nza);
        f[4] = cgm * rmct * hm;
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) f[m] = T(0.0);
    }
    return ok;
}

// flux.f90:395-408 flux_vanleer: F+(qL) + F-(qR)
template <typename T>
__device__ inline bool flux_vl(const T *ql, const T *qr, T nx, T ny, T nz, T *f, double gama)
{
    T fl[5], fr[5];
    bool ok = flux_vl_pn(ql, nx, ny, nz, fl, gama, 1);
    ok = flux_vl_pn(qr, nx, ny, nz, fr, gama, -1) && ok;
#pragma unroll
    for (int m = 0; m < 5; m++) f[m] = fl[m] + fr[m];
    return ok;
}

// The face flux of nflux = 1 (Steger-Warming), 2 (van Leer) or 3 (Roe),
// as invcode dispatches it (Convect.f90:20-37).  False: c^2 < 0 (Roe, van Leer).
template <int FLUX, typename T>
__device__ inline bool face_flux(const T *ql, const T *qr, T nx, T ny, T nz, T *f, double efix, double gama)
{
    if (FLUX == 3) return flux_roe(ql, qr, nx, ny, nz, f, efix, gama);
    if (FLUX == 2) return flux_vl(ql, qr, nx, ny, nz, f, gama);
    flux_sw(ql, qr, nx, ny, nz, f, efix, gama);
    return true;
}

// ------------------------------------------- characteristic schemes, RoeMx
// WCNSE5.f90:4858-4912 pave: Roe average of two primitive states.
__device__ inline void pave(const double *ql, const double *qr, double *qa, double gama)
{
    const double gam1 = gama - 1.0, gam2 = gama / gam1;
    const double vl2 = ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3];
    const double vr2 = qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3];
    const double hl = gam2 * ql[4] / ql[0] + 0.5 * vl2;
    const double hr = gam2 * qr[4] / qr[0] + 0.5 * vr2;
    double rls = sqrt(ql[0]), rrs = sqrt(qr[0]);
    const double ra = rls * rrs;
    const double oors = 1.0 / (rls + rrs);
    rls = rls * oors;
    rrs = rrs * oors;
    const double ua = rls * ql[1] + rrs * qr[1], va = rls * ql[2] + rrs * qr[2], wa = rls * ql[3] + rrs * qr[3];
    const double ha = rls * hl + rrs * hr;
    const double va2 = ua * ua + va * va + wa * wa;
    qa[0] = ra; qa[1] = ua; qa[2] = va; qa[3] = wa;
    qa[4] = (ha - 0.5 * va2) * ra / gam2;
}

// The unit normal of LxQ / RxQ (WCNSE5.f90:4971-4978).
__device__ inline void char_normal(const double *gykb, double &nx, double &ny, double &nz)
{
    double sn = sqrt(gykb[0] * gykb[0] + gykb[1] * gykb[1] + gykb[2] * gykb[2]);
    const double oon = 1.0 / (sn > SML_SSS ? sn : SML_SSS);
    nx = gykb[0] * oon; ny = gykb[1] * oon; nz = gykb[2] * oon;
}

// WCNSE5.f90:4960-5004 LxQ: the left eigenvectors of the normal gykb at the
// primitive state prim, applied to dq.
__device__ inline void lxq(const double *prim, const double *gykb, const double *dq, double *f, double gama)
{
    double nx, ny, nz;
    char_normal(gykb, nx, ny, nz);
    const double c2 = gama * prim[4] / prim[0];
    const double cm = sqrt(c2);
    const double dqn = nx * dq[1] + ny * dq[2] + nz * dq[3];
    const double wp1 = dq[0] - dq[4] / c2;
    const double wp2 = dq[4] / (prim[0] * cm);
    f[0] = nz * dq[2] - ny * dq[3] + nx * wp1;
    f[1] = -nz * dq[1] + nx * dq[3] + ny * wp1;
    f[2] = ny * dq[1] - nx * dq[2] + nz * wp1;
    f[3] = dqn + wp2;
    f[4] = -dqn + wp2;
}

// WCNSE5.f90:5006-5050 RxQ: the right eigenvectors, LxQ's inverse.
__device__ inline void rxq(const double *prim, const double *gykb, const double *dq, double *f, double gama)
{
    double nx, ny, nz;
    char_normal(gykb, nx, ny, nz);
    const double c2 = gama * prim[4] / prim[0];
    const double cm = sqrt(c2);
    const double dqn = nx * dq[0] + ny * dq[1] + nz * dq[2];
    const double wp1 = 0.5 * prim[0] * (dq[3] + dq[4]);
    const double wp2 = 0.5 * (dq[3] - dq[4]);
    f[0] = dqn + wp1 / cm;
    f[1] = -nz * dq[1] + ny * dq[2] + nx * wp2;
    f[2] = nz * dq[0] - nx * dq[2] + ny * wp2;
    f[3] = -ny * dq[0] + nx * dq[1] + nz * wp2;
    f[4] = wp1 * cm;
}

// flux.f90:246-393 flux_roemx: the shock-stable Roe flux of Kim and Kim
// (RoeM), its pressure-wiggle control weighted by the face's least pressure
// ratio rpmin (Convect.f90:78-200).
__device__ inline void flux_roemx(const double *ql, const double *qr, double sx, double sy, double sz, double rpmin,
                                  double *flr, double gama)
{
    double sn = sqrt(sx * sx + sy * sy + sz * sz);
    sn = sn < SML_SSS ? SML_SSS : sn;
    const double osn = 1.0 / sn;
    const double nx = sx * osn, ny = sy * osn, nz = sz * osn;
    const double rl = ql[0] + SMALL, ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4] + SMALL;
    const double rr = qr[0] + SMALL, ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4] + SMALL;
    const double vl2 = ul * ul + vl * vl + wl * wl;
    const double vr2 = ur * ur + vr * vr + wr * wr;
    const double ae = gama - 1.0;
    double el = pl / (rl * ae), er = pr / (rr * ae);
    el = el + 0.5 * vl2;
    const double hl = el + pl / rl;
    er = er + 0.5 * vr2;
    const double hr = er + pr / rr;
    const double vnl = nx * ul + ny * vl + nz * wl;
    const double vnr = nx * ur + ny * vr + nz * wr;
    double dq[5], qi[5], mxdq[5];
    dq[0] = rr - rl;
    dq[1] = rr * ur - rl * ul;
    dq[2] = rr * vr - rl * vl;
    dq[3] = rr * wr - rl * wl;
    dq[4] = rr * hr - rl * hl;
    double rls = sqrt(rl), rrs = sqrt(rr);
    qi[0] = rls * rrs;
    const double oors = 1.0 / (rls + rrs);
    rls = rls * oors;
    rrs = rrs * oors;
    qi[1] = rls * ul + rrs * ur;
    qi[2] = rls * vl + rrs * vr;
    qi[3] = rls * wl + rrs * wr;
    qi[4] = rls * hl + rrs * hr;
    const double gami = rls * gama + rrs * gama;
    const double vi2 = qi[1] * qi[1] + qi[2] * qi[2] + qi[3] * qi[3];
    const double ci2 = (gami - 1.0) * (qi[4] - 0.5 * vi2);
    const double ci = sqrt(ci2);
    const double vni = qi[1] * nx + qi[2] * ny + qi[3] * nz;
    double b1 = 0.0, b2 = 0.0;
    if (vni + ci > b1) b1 = vni + ci;
    if (vnr + ci > b1) b1 = vnr + ci;
    if (vni - ci < b2) b2 = vni - ci;
    if (vnl - ci < b2) b2 = vnl - ci;
    const double oob12 = 1.0 / (b1 - b2);
    b1 = b1 * oob12;
    const double b1b2 = b1 * b2;
    b2 = b2 * oob12;
    const double mia = fabs(vni / ci);
    const double fden = vi2 <= SMALL ? 1.0 : pow(mia, 1.0 - rpmin);
    const double rpmlr = pl / pr < pr / pl ? pl / pr : pr / pl;
    const double gcdc = mia <= SMALL ? 1.0 : pow(mia, 1.0 - rpmlr);
    const double dvn = vnr - vnl, dps = pr - pl;
    const double rmp = dq[0] - fden * dps / ci2;
    mxdq[0] = rmp;
    mxdq[1] = rmp * qi[1] + qi[0] * ((ur - ul) - nx * dvn);
    mxdq[2] = rmp * qi[2] + qi[0] * ((vr - vl) - ny * dvn);
    mxdq[3] = rmp * qi[3] + qi[0] * ((wr - wl) - nz * dvn);
    mxdq[4] = rmp * qi[4] + qi[0] * (hr - hl);
    const double rvnl = b1 * rl * vnl;
    const double rvnr = -b2 * rr * vnr;
    const double plr = b1 * pl - b2 * pr;
    flr[0] = rvnl + rvnr + b1b2 * dq[0];
    flr[1] = rvnl * ul + rvnr * ur + nx * plr + b1b2 * dq[1];
    flr[2] = rvnl * vl + rvnr * vr + ny * plr + b1b2 * dq[2];
    flr[3] = rvnl * wl + rvnr * wr + nz * plr + b1b2 * dq[3];
    flr[4] = rvnl * hl + rvnr * hr + b1b2 * dq[4];
    const double cmxd = -gcdc * b1b2 / (1.0 + mia);
#pragma unroll
    for (int m = 0; m < 5; m++) flr[m] = (flr[m] + cmxd * mxdq[m]) * sn;
}

// WCNSE5.f90:2138-2187 (one variable at one node): nonlinear weights from the
// five values qm2..qp2 around node i, returning the left state at face i+1/2
// (qwl(i)) and the right state at face i-1/2 (qwr(i-1)).
template <typename T>
__device__ inline void wcns_node(T qm2, T qm1, T q0, T qp1, T qp2, T &qwl, T &qwr)
{
    const double EPS = 1.0e-6;
    T g1 = 0.5 * (qm2 - 4.0 * qm1 + 3.0 * q0);
    T g2 = 0.5 * (qp1 - qm1);
    T g3 = 0.5 * (-3.0 * q0 + 4.0 * qp1 - qp2);
    T s1 = qm2 - 2.0 * qm1 + q0;
    T s2 = qm1 - 2.0 * q0 + qp1;
    T s3 = q0 - 2.0 * qp1 + qp2;
    T is1 = g1 * g1 + s1 * s1;
    T is2 = g2 * g2 + s2 * s2;
    T is3 = g3 * g3 + s3 * s3;
    T e1 = (EPS + is1) * (EPS + is1);
    T e2 = (EPS + is2) * (EPS + is2);
    T e3 = (EPS + is3) * (EPS + is3);
    T bl1 = (1.0 / 16.0) / e1, bl2 = (10.0 / 16.0) / e2, bl3 = (5.0 / 16.0) / e3;
    T br1 = (5.0 / 16.0) / e1, br2 = (10.0 / 16.0) / e2, br3 = (1.0 / 16.0) / e3;
    T sl = bl1 + bl2 + bl3;
    T sr = br1 + br2 + br3;
    T wl1 = bl1 / sl, wl2 = bl2 / sl, wl3 = bl3 / sl;
    T wr1 = br1 / sr, wr2 = br2 / sr, wr3 = br3 / sr;
    qwl = q0 + 0.125 * (wl1 * (s1 + 4.0 * g1) + wl2 * (s2 + 4.0 * g2) + wl3 * (s3 + 4.0 * g3));
    qwr = q0 + 0.125 * (wr1 * (s1 - 4.0 * g1) + wr2 * (s2 - 4.0 * g2) + wr3 * (s3 - 4.0 * g3));
}

// LHS.f90:756-828 mxdq_std (kt = 0), generic over the arithmetic type T of the
// implicit operator.  The routine is linear in dQ; it is split here into the
// dQ-independent coefficients (mxdq_coef: the node's eigenvalues, projections
// and normal) and their application to dQ (mxdq_apply), so that a sweep can
// form the coefficients before the neighbour's dQ is known.  The two parts
// perform exactly the reference's operations in its order.
template <typename T>
struct MxCoef {
    T l1, x1, x2, nx, ny, nz, ct, af, ae, c2, um, vm, wm, hm;
};

template <typename T>
__device__ inline void mxdq_coef(const T *prim, const T *gykb, T rad, T npn, T gama, MxCoef<T> &c)
{
    T rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    T hint = gama / (gama - T(1.0)) * pm / rm;
    T c2 = gama * pm / rm;
    T cm = sqrt(c2);
    T v2 = um * um + vm * vm + wm * wm;
    T hm = hint + T(0.5) * v2;
    T nx = gykb[0], ny = gykb[1], nz = gykb[2];
    T ct = nx * um + ny * vm + nz * wm + T(0.0);
    T cgm = sqrt(nx * nx + ny * ny + nz * nz);
    cgm = cgm < T(SML_SSS) ? T(SML_SSS) : cgm;
    T l1 = ct, l4 = ct + cm * cgm, l5 = ct - cm * cgm;
    l1 = T(0.5) * (l1 + npn * rad);
    l4 = T(0.5) * (l4 + npn * rad);
    l5 = T(0.5) * (l5 + npn * rad);
    c.x1 = (T(2.0) * l1 - l4 - l5) / (T(2.0) * c2);
    c.x2 = (l4 - l5) / (T(2.0) * cm);
    T cgm1 = T(1.0) / cgm;
    c.nx = nx * cgm1;
    c.ny = ny * cgm1;
    c.nz = nz * cgm1;
    c.ct = (ct - T(0.0)) * cgm1;
    c.ae = gama - T(1.0);
    c.af = T(0.5) * c.ae * v2;
    c.l1 = l1; c.c2 = c2; c.um = um; c.vm = vm; c.wm = wm; c.hm = hm;
}

template <typename T>
__device__ inline void mxdq_apply(const MxCoef<T> &c, const T *dq, T *f)
{
    T dc = c.ct * dq[0] - c.nx * dq[1] - c.ny * dq[2] - c.nz * dq[3];
    T dh = c.af * dq[0] - c.ae * (c.um * dq[1] + c.vm * dq[2] + c.wm * dq[3] - dq[4]);
    T c2dc = c.c2 * dc;
    f[0] = c.l1 * dq[0] - dh * c.x1 - dc * c.x2;
    f[1] = c.l1 * dq[1] + (c.nx * c2dc - c.um * dh) * c.x1 + (c.nx * dh - c.um * dc) * c.x2;
    f[2] = c.l1 * dq[2] + (c.ny * c2dc - c.vm * dh) * c.x1 + (c.ny * dh - c.vm * dc) * c.x2;
    f[3] = c.l1 * dq[3] + (c.nz * c2dc - c.wm * dh) * c.x1 + (c.nz * dh - c.wm * dc) * c.x2;
    f[4] = c.l1 * dq[4] + (c.ct * c2dc - c.hm * dh) * c.x1 + (c.ct * dh - c.hm * dc) * c.x2;
}

template <typename T>
__device__ inline void mxdq_std(const T *prim, const T *gykb, const T *dq, T *f, T rad, T npn, T gama)
{
    MxCoef<T> c;
    mxdq_coef<T>(prim, gykb, rad, npn, gama, c);
    mxdq_apply<T>(c, dq, f);
}

}  // namespace wcns
