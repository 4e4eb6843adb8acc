// wcns_rhs_inv.hip -- inviscid residual R += dF/dxi along i, j, k pencils
// (reference src/RHS.f90:35-72 -> src/Convect.f90:58-649 inviscd3d ->
// src/WCNSE5.f90:2-206 WCNS_E_5 / :2087-2759 WCNS_E_5_41 -> src/flux.f90 fluxes
// -> src/WCNSE5.f90:209-254 FLUX_DXYZ).
//
// Mapping: one thread marches one segment of one pencil, keeping the last
// five nodes (5 primitives each), four metric triples and six face fluxes in
// registers.  Adjacent threads own adjacent pencils, so for j- and k-pencils
// every load is coalesced along i.  All blocks go in one launch over a block
// work-list (BlockGrid).  A segment of SEG rows recomputes the
// (at most) 3+3 faces it shares with its neighbours; the segment that owns
// rows 1..3 (resp. n-2..n) also applies the WCNS_E_5_41 boundary rows.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

constexpr int SEG = PENCIL_SEG;

// The pencil walk below is generic over an IO policy that supplies the
// scalar type T, the node loads and what becomes of a finished residual row:
// InvIO (the solver: double, rows added to R), InvTanIO (tangent: D1 loads of
// value and direction, rows' derivatives added to dR) and InvVjpIO (adjoint:
// D1 loads seeded on one variable of one colour of nodes, rows' derivatives
// contracted with the adjoint weights and added to that node's gradient).
struct InvIO {
    using T = double;
    static constexpr int SEGLEN = SEG;
    static constexpr bool COUNT = true;
    const double *r, *u, *v, *w, *p;
    const double *mx, *my, *mz;    // metric triple of the pencil direction
    double *rhs;
    long total;
    __device__ void q(long g, int, double *o) const { o[0] = r[g]; o[1] = u[g]; o[2] = v[g]; o[3] = w[g]; o[4] = p[g]; }
    __device__ void m(long g, int, double *o) const { o[0] = mx[g]; o[1] = my[g]; o[2] = mz[g]; }
    __device__ void row(long g, int, const double *df) const
    {
#pragma unroll
        for (int c = 0; c < 5; c++) rhs[c * total + g] = rhs[c * total + g] + df[c];
    }
};

// InvIO for the two-wave register budget: a type of its own, so that the
// boundary functions it calls (not inlined) are compiled once per budget.
struct InvIO2w : InvIO {};

struct InvTanIO {
    using T = D1;
    static constexpr int SEGLEN = SEG;
    static constexpr bool COUNT = false;
    const double *r, *u, *v, *w, *p, *mx, *my, *mz;
    const double *dr, *du, *dv, *dw, *dp, *dmx, *dmy, *dmz;
    double *drhs;
    long total;
    __device__ void q(long g, int, D1 *o) const
    {
        o[0] = D1(r[g], dr[g]); o[1] = D1(u[g], du[g]); o[2] = D1(v[g], dv[g]);
        o[3] = D1(w[g], dw[g]); o[4] = D1(p[g], dp[g]);
    }
    __device__ void m(long g, int, D1 *o) const { o[0] = D1(mx[g], dmx[g]); o[1] = D1(my[g], dmy[g]); o[2] = D1(mz[g], dmz[g]); }
    __device__ void row(long g, int, const D1 *df) const
    {
#pragma unroll
        for (int c = 0; c < 5; c++) drhs[c * total + g] = drhs[c * total + g] + df[c].d;
    }
};

// seed variable sv: 0..4 the primitives r,u,v,w,p, 5..7 the metric triple;
// lam: the gradient array of that variable
struct InvVjpIO {
    using T = D1;
    static constexpr int SEGLEN = 1 << 30;     // one thread walks the whole pencil
    static constexpr bool COUNT = false;
    const double *r, *u, *v, *w, *p, *mx, *my, *mz;
    const double *wt;                          // adjoint weights of R (5 arrays)
    double *lam;
    long total, stride;
    int n, color, sv;
    __device__ double seed(int var, int l) const { return (var == sv && ad_mod(l) == color) ? 1.0 : 0.0; }
    __device__ void q(long g, int l, D1 *o) const
    {
        o[0] = D1(r[g], seed(0, l)); o[1] = D1(u[g], seed(1, l)); o[2] = D1(v[g], seed(2, l));
        o[3] = D1(w[g], seed(3, l)); o[4] = D1(p[g], seed(4, l));
    }
    __device__ void m(long g, int l, D1 *o) const
    {
        o[0] = D1(mx[g], seed(5, l)); o[1] = D1(my[g], seed(6, l)); o[2] = D1(mz[g], seed(7, l));
    }
    __device__ void row(long g, int l, const D1 *df) const
    {
        const int lj = ad_node(l, color);
        if (lj < -2 || lj > n + 3) return;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 5; c++) s = s + wt[c * total + g] * df[c].d;
        const long gj = g + (long)(lj - l) * stride;
        lam[gj] = lam[gj] + s;
    }
};

template <class IO>
struct Pencil {
    IO io;
    const double *r;               // density (boundary flags)
    long base;                     // index of node 0 of the pencil
    long stride;                   // index step along the pencil
    int n;
};

template <class IO>
__device__ inline void load_q(const Pencil<IO> &pc, int l, typename IO::T *q)
{
    pc.io.q(pc.base + (long)l * pc.stride, l, q);
}

template <class IO>
__device__ inline void load_m(const Pencil<IO> &pc, int l, typename IO::T *t)
{
    pc.io.m(pc.base + (long)l * pc.stride, l, t);
}

template <class IO>
__device__ inline void emit_row(const Pencil<IO> &pc, int l, const typename IO::T *df)
{
    pc.io.row(pc.base + (long)l * pc.stride, l, df);
}

// VALUE_HALF_NODE (WCNSE5.f90:799-852) of the metric triple at face f, n > 4.
template <class IO>
__device__ inline void half_metric(const Pencil<IO> &pc, int f, typename IO::T *nx)
{
    using T = typename IO::T;
    int n = pc.n;
    T a[3], b[3], c[3], d[3];
    if (f >= 2 && f <= n - 2) {
        load_m(pc, f, a); load_m(pc, f + 1, b); load_m(pc, f + 2, c); load_m(pc, f - 1, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (9.0 * (a[m] + b[m]) + -1.0 * (c[m] + d[m])) / 16.0;
    } else if (f == 1) {
        load_m(pc, 1, a); load_m(pc, 2, b); load_m(pc, 3, c); load_m(pc, 4, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (5.0 * a[m] + 15.0 * b[m] + -5.0 * c[m] + 1.0 * d[m]) / 16.0;
    } else if (f == n - 1) {
        load_m(pc, n, a); load_m(pc, n - 1, b); load_m(pc, n - 2, c); load_m(pc, n - 3, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (5.0 * a[m] + 15.0 * b[m] + -5.0 * c[m] + 1.0 * d[m]) / 16.0;
    } else if (f == 0) {
        load_m(pc, 1, a); load_m(pc, 2, b); load_m(pc, 3, c); load_m(pc, 4, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (35.0 * a[m] - 35.0 * b[m] + 21.0 * c[m] - 5.0 * d[m]) / 16.0;
    } else {
        load_m(pc, n, a); load_m(pc, n - 1, b); load_m(pc, n - 2, c); load_m(pc, n - 3, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (35.0 * a[m] - 35.0 * b[m] + 21.0 * c[m] - 5.0 * d[m]) / 16.0;
    }
}

// WCNS qwl(node) or qwr(node-1) straight from global memory (boundary rows).
template <class IO>
__device__ inline void wcns_at(const Pencil<IO> &pc, int node, typename IO::T *qwl, typename IO::T *qwr)
{
    typename IO::T q[5][5];
    for (int l = 0; l < 5; l++) load_q(pc, node - 2 + l, q[l]);
#pragma unroll
    for (int m = 0; m < 5; m++) wcns_node(q[0][m], q[1][m], q[2][m], q[3][m], q[4][m], qwl[m], qwr[m]);
}

// Fallback-checked face flux (WCNSE5.f90:2200-2262).  mode 0: ql<-q(max(f,1)),
// qr<-q(min(f+1,n)); mode 1: both <- (q(max(f,1))+q(f+1))/2 (WCNSE5.f90:2317-2330);
// mode 2: both <- (q(1)+q(f))/2 (WCNSE5.f90:2568-2582, min(i+1,1) == 1).
template <int FLUX, class IO>
__device__ inline void checked_flux(const Pencil<IO> &pc, int f, typename IO::T *ql, typename IO::T *qr,
                                    const typename IO::T *nx, typename IO::T *flr, double efix, double gama, int mode,
                                    int &nerr, int &bad)
{
    bool fl = ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT;
    bool fr = qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT;
    if (fl || fr) {
        typename IO::T qa[5], qb[5];
        if (mode == 0) {
            if (fl) { load_q(pc, f > 1 ? f : 1, qa); for (int m = 0; m < 5; m++) ql[m] = qa[m]; nerr++; }
            if (fr) { load_q(pc, f + 1 < pc.n ? f + 1 : pc.n, qb); for (int m = 0; m < 5; m++) qr[m] = qb[m]; nerr++; }
        } else if (mode == 1) {
            load_q(pc, f > 1 ? f : 1, qa); load_q(pc, f + 1, qb);
            if (fl) { for (int m = 0; m < 5; m++) ql[m] = (qa[m] + qb[m]) / 2.; nerr++; }
            if (fr) { for (int m = 0; m < 5; m++) qr[m] = (qa[m] + qb[m]) / 2.; nerr++; }
        } else {
            load_q(pc, 1, qa); load_q(pc, f, qb);
            if (fl) { for (int m = 0; m < 5; m++) ql[m] = (qa[m] + qb[m]) / 2.0; nerr++; }
            if (fr) { for (int m = 0; m < 5; m++) qr[m] = (qa[m] + qb[m]) / 2.0; nerr++; }
        }
    }
    if (!face_flux<FLUX>(ql, qr, nx[0], nx[1], nx[2], flr, efix, gama)) bad = 1;
}

// The WCNS_E_5_41 left boundary rows 1..3 (WCNSE5.f90:2269-2513).
template <int FLUX, class IO>
__device__ __noinline__ void left41(const Pencil<IO> &pc, typename IO::T *df1, typename IO::T *df2,
                                    typename IO::T *df3, double efix, double gama, bool s42, int &nerr, int &bad)
{
    using T = typename IO::T;
    T q1[5], q2[5], q3[5], q4[5], urold[5][5], tmp[5], nx[3], f[6][5], ql[5], qr[5];
    load_q(pc, 1, q1); load_q(pc, 2, q2); load_q(pc, 3, q3); load_q(pc, 4, q4);
    // u_r_old(1..4): qwr after the scheme-4 closures; qwr(1) = (5,15,-5,1)/16, qwr(2..4) WCNS
#pragma unroll
    for (int m = 0; m < 5; m++) urold[1][m] = (5.0 * q1[m] + 15.0 * q2[m] - 5.0 * q3[m] + q4[m]) / 16.0;
    for (int i = 2; i <= 4; i++) wcns_at(pc, i + 1, tmp, urold[i]);
    // (1) faces 0..3 -> df(1); scheme 42's states WCNSE5.f90:2967-2986
    for (int i = 0; i <= 3; i++) {
        for (int m = 0; m < 5; m++) {
            T v;
            if (s42) {
                if (i == 0) v = (677.0 * q1[m] - 421.0 * q2[m] + 99.0 * q3[m] + 13.0 * q4[m]) / 368.0;
                else if (i == 1) v = (5.0 * q1[m] + 15.0 * q2[m] - 5.0 * q3[m] + q4[m]) / 16.0;
                else if (i == 2) v = (-q1[m] + 9.0 * q2[m] + 9.0 * q3[m] - q4[m]) / 16.0;
                else v = (q1[m] - 5.0 * q2[m] + 15.0 * q3[m] + 5.0 * q4[m]) / 16.0;
            } else {
                if (i == 0) v = (3.0 * q1[m] - 1.0 * q2[m] - 0. * q3[m]) / 2.0;
                else if (i == 1) v = (19.0 * q1[m] + 25.0 * q2[m] - 2. * q3[m]) / 42.0;
                else if (i == 2) v = (4.0 * q1[m] + 3.0 * q2[m] + 9.0 * q3[m] + 2. * q4[m]) / 18.0;
                else v = (-2.0 * q1[m] + 3.0 * q2[m] + 3.0 * q3[m] + 2. * q4[m]) / 6.0;
            }
            ql[m] = v; qr[m] = v;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[i], efix, gama, 1, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) df1[m] = (-23.0 * f[0][m] + 21.0 * f[1][m] + 3.0 * f[2][m] - 1.0 * f[3][m]) / 24.0;
    // (2) faces 1..5 -> df(2)
    T s2l[6][5];
    for (int i = 1; i <= 5; i++) {
        for (int m = 0; m < 5; m++) {
            T vl, vr;
            if (i == 1) { vl = (3.0 * q1[m] + 6.0 * q2[m] - q3[m]) / 8.0; vr = vl; }
            else if (i == 2) {
                vl = (-30.0 * q1[m] + 145.0 * q2[m] + 29.0 * q3[m] - 8.0 * q4[m]) / 136.0;
                vr = (2.0 * q1[m] + 49.0 * q2[m] + 125.0 * q3[m] - 40.0 * q4[m]) / 136.0;
            } else if (i == 3) { vl = (-q2[m] + 6.0 * q3[m] + 3.0 * q4[m]) / 8.0; vr = vl; }
            else if (i == 4) { vl = (q2[m] + 2.0 * q3[m] + 5.0 * q4[m]) / 8.0; vr = vl; }
            else { vl = (q2[m] + q3[m] + 6.0 * q4[m]) / 8.0; vr = vl; }
            ql[m] = vl; qr[m] = vr; s2l[i][m] = vl;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[i], efix, gama, 0, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++)
        df2[m] = (-22.0 * f[1][m] + 17.0 * f[2][m] + 9.0 * f[3][m] - 5.0 * f[4][m] + f[5][m]) / 24.0;
    // (3) faces 1..4 -> df(3)
    for (int i = 1; i <= 4; i++) {
        for (int m = 0; m < 5; m++) {
            ql[m] = (i == 2) ? (-29.0 * q1[m] + 170.0 * q2[m] + 63.0 * q3[m] + 12.0 * q4[m]) / 216.0 : s2l[i][m];
            qr[m] = urold[i][m];
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[i], efix, gama, 0, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) df3[m] = (f[1][m] - 27.0 * f[2][m] + 27.0 * f[3][m] - f[4][m]) / 24.0;
}

// The WCNS_E_5_41 right boundary rows n-2..n (WCNSE5.f90:2518-2751).  f[k] holds face n-k.
template <int FLUX, class IO>
__device__ __noinline__ void right41(const Pencil<IO> &pc, typename IO::T *dfn, typename IO::T *dfn1,
                                     typename IO::T *dfn2, double efix, double gama, bool s42, int &nerr, int &bad)
{
    using T = typename IO::T;
    int n = pc.n;
    T qn[5], qn1[5], qn2[5], qn3[5], ulold[9][5], tmp[5], nx[3], f[6][5], ql[5], qr[5];
    load_q(pc, n, qn); load_q(pc, n - 1, qn1); load_q(pc, n - 2, qn2); load_q(pc, n - 3, qn3);
    // u_l_old(9-i) = qwl(n-i), i=1..4: qwl(n-1) = (5,15,-5,1)/16 closure, qwl(n-2..n-4) WCNS
#pragma unroll
    for (int m = 0; m < 5; m++) ulold[8][m] = (5.0 * qn[m] + 15.0 * qn1[m] - 5.0 * qn2[m] + qn3[m]) / 16.0;
    for (int i = 2; i <= 4; i++) wcns_at(pc, n - i, ulold[9 - i], tmp);
    // (4) faces n..n-3 -> df(n); scheme 42's states WCNSE5.f90:3207-3226
    for (int i = n; i >= n - 3; i--) {
        int o = n - i;
        for (int m = 0; m < 5; m++) {
            T v;
            if (s42) {
                if (o == 0) v = (677.0 * qn[m] - 421.0 * qn1[m] + 99.0 * qn2[m] + 13.0 * qn3[m]) / 368.0;
                else if (o == 1) v = (5.0 * qn[m] + 15.0 * qn1[m] - 5.0 * qn2[m] + qn3[m]) / 16.0;
                else if (o == 2) v = (-qn[m] + 9.0 * qn1[m] + 9.0 * qn2[m] - qn3[m]) / 16.0;
                else v = (qn[m] - 5.0 * qn1[m] + 15.0 * qn2[m] + 5.0 * qn3[m]) / 16.0;
            } else {
                if (o == 0) v = (3.0 * qn[m] - 1.0 * qn1[m] - 0. * qn2[m]) / 2.0;
                else if (o == 1) v = (19.0 * qn[m] + 25.0 * qn1[m] - 2. * qn2[m]) / 42.0;
                else if (o == 2) v = (4.0 * qn[m] + 3.0 * qn1[m] + 9.0 * qn2[m] + 2. * qn3[m]) / 18.0;
                else v = (-2.0 * qn[m] + 3.0 * qn1[m] + 3.0 * qn2[m] + 2. * qn3[m]) / 6.0;
            }
            ql[m] = v; qr[m] = v;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[o], efix, gama, 2, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) dfn[m] = -(-23.0 * f[0][m] + 21.0 * f[1][m] + 3.0 * f[2][m] - 1.0 * f[3][m]) / 24.0;
    // (5) faces n-1..n-5 -> df(n-1)
    T s5r[6][5];
    for (int i = n - 1; i >= n - 5; i--) {
        int o = n - i;
        for (int m = 0; m < 5; m++) {
            T vl, vr;
            if (o == 1) { vr = (3.0 * qn[m] + 6.0 * qn1[m] - qn2[m]) / 8.0; vl = vr; }
            else if (o == 2) {
                vr = (-30.0 * qn[m] + 145.0 * qn1[m] + 29.0 * qn2[m] - 8.0 * qn3[m]) / 136.0;
                vl = (2.0 * qn[m] + 49.0 * qn1[m] + 125.0 * qn2[m] - 40.0 * qn3[m]) / 136.0;
            } else if (o == 3) { vr = (-qn1[m] + 6.0 * qn2[m] + 3.0 * qn3[m]) / 8.0; vl = vr; }
            else if (o == 4) { vr = (qn1[m] + 2.0 * qn2[m] + 5.0 * qn3[m]) / 8.0; vl = vr; }
            else { vr = (qn1[m] + qn2[m] + 6.0 * qn3[m]) / 8.0; vl = vr; }
            ql[m] = vl; qr[m] = vr; s5r[o][m] = vr;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[o], efix, gama, 0, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++)
        dfn1[m] = -(-22.0 * f[1][m] + 17.0 * f[2][m] + 9.0 * f[3][m] - 5.0 * f[4][m] + f[5][m]) / 24.0;
    // (6) faces n-1..n-4 -> df(n-2)
    for (int i = n - 1; i >= n - 4; i--) {
        int o = n - i;
        for (int m = 0; m < 5; m++) {
            qr[m] = (o == 2) ? (-29.0 * qn[m] + 170.0 * qn1[m] + 63.0 * qn2[m] + 12.0 * qn3[m]) / 216.0 : s5r[o][m];
            ql[m] = ulold[9 - o][m];
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[o], efix, gama, 0, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) dfn2[m] = -(f[1][m] - 27.0 * f[2][m] + 27.0 * f[3][m] - f[4][m]) / 24.0;
}

// The scheme-4 boundary states (WCNSE5.f90:2192-2213): which 0,1,2 = qwl(0), qwl(1)=qwr(1), qwl(2)
// at a left physical boundary; 3,4,5 = qwr(n)=qwl(n), qwr(n-1)=qwl(n-1), qwr(n-2) at a right one.
template <class IO>
__device__ __noinline__ void closure_state(const Pencil<IO> &pc, int which, typename IO::T *out)
{
    typename IO::T q1[5], q2[5], q3[5], q4[5];
    int n = pc.n;
    if (which < 3) { load_q(pc, 1, q1); load_q(pc, 2, q2); load_q(pc, 3, q3); load_q(pc, 4, q4); }
    else { load_q(pc, n, q1); load_q(pc, n - 1, q2); load_q(pc, n - 2, q3); load_q(pc, n - 3, q4); }
    int c = which % 3;
    for (int m = 0; m < 5; m++) {
        if (c == 0) out[m] = (35.0 * q1[m] - 35.0 * q2[m] + 21.0 * q3[m] - 5.0 * q4[m]) / 16.0;
        else if (c == 1) out[m] = (5.0 * q1[m] + 15.0 * q2[m] - 5.0 * q3[m] + q4[m]) / 16.0;
        else out[m] = (-q1[m] + 9.0 * q2[m] + 9.0 * q3[m] - q4[m]) / 16.0;
    }
}

// One pencil's rows a..bb (DIR only selects nothing here: the pencil carries
// its own base, stride and length).  The product instance adds the rows to R.
template <int FLUX, class IO>
__device__ inline void inv_pencil(const Pencil<IO> &pc, int a, int bb, double gama, double efix, bool s41, bool s42,
                                  int &nerr, int &bad)
{
    using T = typename IO::T;
    const int n = pc.n;
    // physical-boundary flags: negative density in ghost layer 3 (WCNSE5.f90:2123-2133)
    const int ist = pc.r[pc.base + (long)(-2) * pc.stride] < SMALL ? 2 : 0;
    const int ied = pc.r[pc.base + (long)(n + 3) * pc.stride] < SMALL ? n - 2 : n;
    const bool lclos = ist > 1, rclos = ied < n;

    int F0 = a - 3 < 0 ? 0 : a - 3;
    int F1 = bb + 2 > n ? n : bb + 2;
    if (a <= 2 && F1 < 3) F1 = 3;
    if (bb >= n - 1 && F0 > n - 3) F0 = n - 3;

    T qw[5][5];       // nodes nd-2..nd+2
    T fw[6][5];       // faces fc-5..fc
    T qwl_prev[5];

    for (int l = 0; l < 4; l++) load_q(pc, F0 - 2 + l, qw[l + 1]);   // qw[1..4] = F0-2..F0+1
    for (int nd = F0; nd <= F1 + 1; nd++) {
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) qw[l][m] = qw[l + 1][m];
        load_q(pc, nd + 2, qw[4]);
        T qwl_cur[5], qwr_m1[5];
        bool need = (nd >= ist && nd <= ied + 1);
        if (need) {
#pragma unroll
            for (int m = 0; m < 5; m++)
                wcns_node(qw[0][m], qw[1][m], qw[2][m], qw[3][m], qw[4][m], qwl_cur[m], qwr_m1[m]);
        }
        // closure overrides of qwl(nd) and qwr(nd-1); the right side wins on overlap
        {
            int f = nd, c = -1;
            if (rclos && f == n) c = 3;
            else if (rclos && f == n - 1) c = 4;
            else if (lclos && f == 0) c = 0;
            else if (lclos && f == 1) c = 1;
            else if (lclos && f == 2) c = 2;
            if (c >= 0) closure_state(pc, c, qwl_cur);
            f = nd - 1;
            c = -1;
            if (rclos && f == n) c = 3;
            else if (rclos && f == n - 1) c = 4;
            else if (rclos && f == n - 2) c = 5;
            else if (lclos && f == 0) c = 0;
            else if (lclos && f == 1) c = 1;
            if (c >= 0) closure_state(pc, c, qwr_m1);
        }
        int fc = nd - 1;
        if (fc >= F0) {
            T nx[3];
            half_metric(pc, fc, nx);
            T ql[5], qr[5];
#pragma unroll
            for (int m = 0; m < 5; m++) { ql[m] = qwl_prev[m]; qr[m] = qwr_m1[m]; }
            bool fl = ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT;
            bool fr = qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT;
            // window slot 1 holds node nd-1 = fc, slot 2 node nd = fc+1 (no dynamic register indexing)
            if (fl) { nerr++; bool s2 = fc < 1; for (int m = 0; m < 5; m++) ql[m] = s2 ? qw[2][m] : qw[1][m]; }
            if (fr) { nerr++; bool s1 = fc + 1 > n; for (int m = 0; m < 5; m++) qr[m] = s1 ? qw[1][m] : qw[2][m]; }
#pragma unroll
            for (int l = 0; l < 5; l++)
#pragma unroll
                for (int m = 0; m < 5; m++) fw[l][m] = fw[l + 1][m];
            if (!face_flux<FLUX>(ql, qr, nx[0], nx[1], nx[2], fw[5], efix, gama)) bad = 1;
            // emit the rows that became computable (FLUX_DXYZ, WCNSE5.f90:229-251)
            T df[5];
            if (fc == 3) {
                if (a <= 1 && !(s41 && lclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = (-23.0 * fw[2][m] + 21.0 * fw[3][m] + 3.0 * fw[4][m] - fw[5][m]) / 24.0;
                    emit_row(pc, 1, df);
                }
                if (a <= 2 && bb >= 2 && !(s41 && lclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = (fw[2][m] - 27.0 * fw[3][m] + 27.0 * fw[4][m] - fw[5][m]) / 24.0;
                    emit_row(pc, 2, df);
                }
            }
            int row = fc - 2;
            if (row >= 3 && row <= n - 2 && row >= a && row <= bb &&
                !(s41 && lclos && row == 3) && !(s41 && rclos && row == n - 2)) {
#pragma unroll
                for (int m = 0; m < 5; m++)
                    df[m] = (2250.0 * (fw[3][m] - fw[2][m]) - 125.0 * (fw[4][m] - fw[1][m]) +
                             9.0 * (fw[5][m] - fw[0][m])) / 1920.0;
                emit_row(pc, row, df);
            }
            if (fc == n) {
                if (bb >= n - 1 && a <= n - 1 && !(s41 && rclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = -(fw[5][m] - 27.0 * fw[4][m] + 27.0 * fw[3][m] - fw[2][m]) / 24.0;
                    emit_row(pc, n - 1, df);
                }
                if (bb >= n && !(s41 && rclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = -(-23.0 * fw[5][m] + 21.0 * fw[4][m] + 3.0 * fw[3][m] - fw[2][m]) / 24.0;
                    emit_row(pc, n, df);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) qwl_prev[m] = qwl_cur[m];
    }

    if (s41 && lclos && a == 1) {
        T d1[5], d2[5], d3[5];
        left41<FLUX>(pc, d1, d2, d3, efix, gama, s42, nerr, bad);
        emit_row(pc, 1, d1);
        emit_row(pc, 2, d2);
        emit_row(pc, 3, d3);
    }
    if (s41 && rclos && bb == n) {
        T dn[5], dn1[5], dn2[5];
        right41<FLUX>(pc, dn, dn1, dn2, efix, gama, s42, nerr, bad);
        emit_row(pc, n, dn);
        emit_row(pc, n - 1, dn1);
        emit_row(pc, n - 2, dn2);
    }
}

// pencil (ia, ib) of direction DIR: node 0 and the index step
template <int DIR>
__device__ inline void pencil_at(const DBlock &b, int ia, int ib, long &base, long &stride, int &n)
{
    int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    base = gidx(b, i0, j0, k0);
    stride = DIR == 0 ? 1 : (DIR == 1 ? b.sx : b.sy);
    n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
}

template <int DIR>
__device__ inline void metric_triple(const DFields &F, const double *&mx, const double *&my, const double *&mz)
{
    mx = DIR == 0 ? F.kcx : (DIR == 1 ? F.etx : F.ctx);
    my = DIR == 0 ? F.kcy : (DIR == 1 ? F.ety : F.cty);
    mz = DIR == 0 ? F.kcz : (DIR == 1 ? F.etz : F.ctz);
}

// DIR0 = 3: the i direction run as j pencils of the transposed view.
template <int DIR0, int FLUX, class IO>
__device__ inline void rhs_inv_thread(const BlockGrid &G, const DFields &F, const DParams &P)
{
    constexpr int DIR = DIR0 == 3 ? 1 : DIR0;
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;              // fast transverse extent
    const int nb = DIR == 2 ? b.nj : b.nk;              // slow transverse extent
    const long npen = (long)na * nb;
    const int nseg = (n + SEG - 1) / SEG;
    if (t >= npen * nseg) return;
    int seg = (int)(t / npen);
    long pen = t - (long)seg * npen;
    int ia = (int)(pen % na) + 1;
    int ib = (int)(pen / na) + 1;

    Pencil<IO> pc;
    pc.io.r = F.r; pc.io.u = F.u; pc.io.v = F.v; pc.io.w = F.w; pc.io.p = F.p;
    metric_triple<DIR>(F, pc.io.mx, pc.io.my, pc.io.mz);
    pc.io.rhs = F.rhs;
    pc.io.total = F.total;
    pc.r = F.r;
    pencil_at<DIR>(b, ia, ib, pc.base, pc.stride, pc.n);

    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    int nerr = 0, bad = 0;
    inv_pencil<FLUX>(pc, a, bb, P.gama, P.efix, P.nscheme >= 41, P.nscheme == 42, nerr, bad);
    if (nerr) atomicAdd(&F.counters[0], (unsigned long long)nerr);
    if (bad) atomicAdd(&F.counters[2], 1ull);
}

// Two register budgets of the same walk.  k_rhs_inv keeps everything in
// registers (one wave per SIMD); k_rhs_inv_2w fits 256 registers, spilling a
// few, so that two waves share a SIMD -- which pays only when the launch has
// more than one wave per SIMD to give.
template <int DIR0, int FLUX>
__global__ void __launch_bounds__(128) k_rhs_inv(BlockGrid G, DFields F, DParams P)
{
    rhs_inv_thread<DIR0, FLUX, InvIO>(G, F, P);
}

template <int DIR0, int FLUX>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) k_rhs_inv_2w(BlockGrid G, DFields F,
                                                                                            DParams P)
{
    rhs_inv_thread<DIR0, FLUX, InvIO2w>(G, F, P);
}

// Tangent: dR += (dR_inv/dP) dP + (dR_inv/dM) dM along DIR for one block, the
// same pencil segments as the solver.
template <int DIR, int FLUX>
__global__ void __launch_bounds__(128) k_rhs_inv_jvp(DBlock b, DFields F, DParams P, AdTan D)
{
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nb = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nb;
    const int nseg = (n + SEG - 1) / SEG;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npen * nseg) return;
    int seg = (int)(t / npen);
    long pen = t - (long)seg * npen;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    Pencil<InvTanIO> pc;
    InvTanIO &io = pc.io;
    io.r = F.r; io.u = F.u; io.v = F.v; io.w = F.w; io.p = F.p;
    metric_triple<DIR>(F, io.mx, io.my, io.mz);
    const long T5 = F.total;
    io.dr = D.dp; io.du = D.dp + T5; io.dv = D.dp + 2 * T5; io.dw = D.dp + 3 * T5; io.dp = D.dp + 4 * T5;
    io.dmx = D.dm + (3 * DIR) * T5; io.dmy = D.dm + (3 * DIR + 1) * T5; io.dmz = D.dm + (3 * DIR + 2) * T5;
    io.drhs = D.dr;
    io.total = T5;
    pc.r = F.r;
    pencil_at<DIR>(b, ia, ib, pc.base, pc.stride, pc.n);
    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    int nerr = 0, bad = 0;
    inv_pencil<FLUX>(pc, a, bb, P.gama, P.efix, P.nscheme >= 41, P.nscheme == 42, nerr, bad);
}

// Adjoint: thread (pencil, colour, seed variable) walks the whole pencil with
// that variable seeded on every node of that colour and adds, for every node
// of the colour, sum over rows of W . dR_row/d(variable at the node) to the
// node's gradient (lamP: r,u,v,w,p; lamM: the pencil direction's metrics).
constexpr int INV_SEEDS = 8;
// Two waves per SIMD (256 registers, about 135 spilled): its launch is always
// large (colours x seeds x pencils), and that budget measured 44.0 -> 41.8 ms
// per adjoint call on the 128x96x64 SA block against about 450 registers at one.
template <int DIR, int FLUX>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) k_rhs_inv_vjp(DBlock b, DFields F,
                                                                                             DParams P, AdAdj A)
{
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nb = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nb;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npen * NCOLOR * INV_SEEDS) return;
    const long pen = t % npen;
    const int cs = (int)(t / npen);
    const int color = cs % NCOLOR, sv = cs / NCOLOR;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    Pencil<InvVjpIO> pc;
    InvVjpIO &io = pc.io;
    io.r = F.r; io.u = F.u; io.v = F.v; io.w = F.w; io.p = F.p;
    metric_triple<DIR>(F, io.mx, io.my, io.mz);
    io.wt = A.w;
    io.total = F.total;
    io.lam = sv < 5 ? A.gp + sv * F.total : A.gm + (3 * DIR + sv - 5) * F.total;
    io.color = color;
    io.sv = sv;
    pc.r = F.r;
    pencil_at<DIR>(b, ia, ib, pc.base, pc.stride, pc.n);
    io.stride = pc.stride;
    io.n = pc.n;
    int nerr = 0, bad = 0;
    inv_pencil<FLUX>(pc, 1, pc.n, P.gama, P.efix, P.nscheme >= 41, P.nscheme == 42, nerr, bad);
}

}  // namespace

// pencil segments of one block along dir; none when the direction has <= 4
// nodes (Convect.f90:150,314,479)
long rhs_inv_threads(int dir, const DBlock &b)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    if (n <= 4) return 0;
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    return npen * ((n + SEG - 1) / SEG);
}

template <int FLUX>
static void launch_inv(int dir, bool two_waves, dim3 grid, dim3 blk, hipStream_t s, const BlockGrid &G,
                       const DFields &F, const DParams &P)
{
    if (two_waves) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_2w<0, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_2w<1, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_2w<2, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 3) hipLaunchKernelGGL((k_rhs_inv_2w<3, FLUX>), grid, blk, 0, s, G, F, P);
    } else {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv<0, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv<1, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv<2, FLUX>), grid, blk, 0, s, G, F, P);
        if (dir == 3) hipLaunchKernelGGL((k_rhs_inv<3, FLUX>), grid, blk, 0, s, G, F, P);
    }
}

// SIMDs of the current device (four per CU)
static int simd_count()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || ncu < 1)
        ncu = 256;
    return 4 * ncu;
}

int launch_rhs_inv_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    dim3 grid((unsigned)G.nwg), blk(128);
    // two 64-wide waves per workgroup; the two-wave budget once they exceed
    // 1.5 waves per SIMD
    const bool two = P.inv_waves ? P.inv_waves == 2 : 2L * G.nwg * 2 > 3L * simd_count();
    if (P.nflux == 3) launch_inv<3>(dir, two, grid, blk, s, G, F, P);
    else if (P.nflux == 2) launch_inv<2>(dir, two, grid, blk, s, G, F, P);
    else launch_inv<1>(dir, two, grid, blk, s, G, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_rhs_inv_jvp(int dir, const DBlock &b, const DFields &F, const DParams &P, const AdTan &D, hipStream_t s)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    if (n <= 4) return 0;
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    long nt = npen * ((n + SEG - 1) / SEG);
    dim3 grid((unsigned)((nt + 127) / 128)), blk(128);
    if (P.nflux == 3) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_jvp<0, 3>), grid, blk, 0, s, b, F, P, D);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_jvp<1, 3>), grid, blk, 0, s, b, F, P, D);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_jvp<2, 3>), grid, blk, 0, s, b, F, P, D);
    } else if (P.nflux == 2) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_jvp<0, 2>), grid, blk, 0, s, b, F, P, D);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_jvp<1, 2>), grid, blk, 0, s, b, F, P, D);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_jvp<2, 2>), grid, blk, 0, s, b, F, P, D);
    } else {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_jvp<0, 1>), grid, blk, 0, s, b, F, P, D);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_jvp<1, 1>), grid, blk, 0, s, b, F, P, D);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_jvp<2, 1>), grid, blk, 0, s, b, F, P, D);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_rhs_inv_vjp(int dir, const DBlock &b, const DFields &F, const DParams &P, const AdAdj &A, hipStream_t s)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    if (n <= 4) return 0;
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    long nt = npen * NCOLOR * INV_SEEDS;
    dim3 grid((unsigned)((nt + 127) / 128)), blk(128);
    if (P.nflux == 3) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_vjp<0, 3>), grid, blk, 0, s, b, F, P, A);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_vjp<1, 3>), grid, blk, 0, s, b, F, P, A);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_vjp<2, 3>), grid, blk, 0, s, b, F, P, A);
    } else if (P.nflux == 2) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_vjp<0, 2>), grid, blk, 0, s, b, F, P, A);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_vjp<1, 2>), grid, blk, 0, s, b, F, P, A);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_vjp<2, 2>), grid, blk, 0, s, b, F, P, A);
    } else {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_vjp<0, 1>), grid, blk, 0, s, b, F, P, A);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_vjp<1, 1>), grid, blk, 0, s, b, F, P, A);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_vjp<2, 1>), grid, blk, 0, s, b, F, P, A);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
