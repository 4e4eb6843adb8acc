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

struct Pencil {
    const double *r, *u, *v, *w, *p;
    const double *mx, *my, *mz;    // metric triple of the pencil direction
    long base;                     // index of node 0 of the pencil
    long stride;                   // index step along the pencil
    int n;
};

__device__ inline void load_q(const Pencil &pc, int l, double *q)
{
    long g = pc.base + (long)l * pc.stride;
    q[0] = pc.r[g]; q[1] = pc.u[g]; q[2] = pc.v[g]; q[3] = pc.w[g]; q[4] = pc.p[g];
}

__device__ inline void load_m(const Pencil &pc, int l, double *t)
{
    long g = pc.base + (long)l * pc.stride;
    t[0] = pc.mx[g]; t[1] = pc.my[g]; t[2] = pc.mz[g];
}

// VALUE_HALF_NODE (WCNSE5.f90:799-852) of the metric triple at face f, n > 4.
__device__ inline void half_metric(const Pencil &pc, int f, double *nx)
{
    int n = pc.n;
    double a[3], b[3], c[3], d[3];
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
__device__ inline void wcns_at(const Pencil &pc, int node, double *qwl, double *qwr)
{
    double q[5][5];
    for (int l = 0; l < 5; l++) load_q(pc, node - 2 + l, q[l]);
#pragma unroll
    for (int m = 0; m < 5; m++) wcns_node(q[0][m], q[1][m], q[2][m], q[3][m], q[4][m], qwl[m], qwr[m]);
}

// Fallback-checked face flux (WCNSE5.f90:2200-2262).  mode 0: ql<-q(max(f,1)),
// qr<-q(min(f+1,n)); mode 1: both <- (q(max(f,1))+q(f+1))/2 (WCNSE5.f90:2317-2330);
// mode 2: both <- (q(1)+q(f))/2 (WCNSE5.f90:2568-2582, min(i+1,1) == 1).
template <int FLUX>
__device__ inline void checked_flux(const Pencil &pc, int f, double *ql, double *qr, const double *nx,
                                    double *flr, double efix, double gama, int mode, int &nerr, int &bad)
{
    bool fl = ql[0] <= RMIN_LIMIT || ql[4] <= PMIN_LIMIT;
    bool fr = qr[0] <= RMIN_LIMIT || qr[4] <= PMIN_LIMIT;
    if (fl || fr) {
        double qa[5], qb[5];
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
template <int FLUX>
__device__ __noinline__ void left41(const Pencil &pc, double *df1, double *df2, double *df3, double efix, double gama,
                       int &nerr, int &bad)
{
    double q1[5], q2[5], q3[5], q4[5], urold[5][5], tmp[5], nx[3], f[6][5], ql[5], qr[5];
    load_q(pc, 1, q1); load_q(pc, 2, q2); load_q(pc, 3, q3); load_q(pc, 4, q4);
    // u_r_old(1..4): qwr after the scheme-4 closures; qwr(1) = (5,15,-5,1)/16, qwr(2..4) WCNS
#pragma unroll
    for (int m = 0; m < 5; m++) urold[1][m] = (5.0 * q1[m] + 15.0 * q2[m] - 5.0 * q3[m] + q4[m]) / 16.0;
    for (int i = 2; i <= 4; i++) wcns_at(pc, i + 1, tmp, urold[i]);
    // (1) faces 0..3 -> df(1)
    for (int i = 0; i <= 3; i++) {
        for (int m = 0; m < 5; m++) {
            double v;
            if (i == 0) v = (3.0 * q1[m] - 1.0 * q2[m] - 0. * q3[m]) / 2.0;
            else if (i == 1) v = (19.0 * q1[m] + 25.0 * q2[m] - 2. * q3[m]) / 42.0;
            else if (i == 2) v = (4.0 * q1[m] + 3.0 * q2[m] + 9.0 * q3[m] + 2. * q4[m]) / 18.0;
            else v = (-2.0 * q1[m] + 3.0 * q2[m] + 3.0 * q3[m] + 2. * q4[m]) / 6.0;
            ql[m] = v; qr[m] = v;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[i], efix, gama, 1, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) df1[m] = (-23.0 * f[0][m] + 21.0 * f[1][m] + 3.0 * f[2][m] - 1.0 * f[3][m]) / 24.0;
    // (2) faces 1..5 -> df(2)
    double s2l[6][5];
    for (int i = 1; i <= 5; i++) {
        for (int m = 0; m < 5; m++) {
            double vl, vr;
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
template <int FLUX>
__device__ __noinline__ void right41(const Pencil &pc, double *dfn, double *dfn1, double *dfn2, double efix, double gama,
                        int &nerr, int &bad)
{
    int n = pc.n;
    double qn[5], qn1[5], qn2[5], qn3[5], ulold[9][5], tmp[5], nx[3], f[6][5], ql[5], qr[5];
    load_q(pc, n, qn); load_q(pc, n - 1, qn1); load_q(pc, n - 2, qn2); load_q(pc, n - 3, qn3);
    // u_l_old(9-i) = qwl(n-i), i=1..4: qwl(n-1) = (5,15,-5,1)/16 closure, qwl(n-2..n-4) WCNS
#pragma unroll
    for (int m = 0; m < 5; m++) ulold[8][m] = (5.0 * qn[m] + 15.0 * qn1[m] - 5.0 * qn2[m] + qn3[m]) / 16.0;
    for (int i = 2; i <= 4; i++) wcns_at(pc, n - i, ulold[9 - i], tmp);
    // (4) faces n..n-3 -> df(n)
    for (int i = n; i >= n - 3; i--) {
        int o = n - i;
        for (int m = 0; m < 5; m++) {
            double v;
            if (o == 0) v = (3.0 * qn[m] - 1.0 * qn1[m] - 0. * qn2[m]) / 2.0;
            else if (o == 1) v = (19.0 * qn[m] + 25.0 * qn1[m] - 2. * qn2[m]) / 42.0;
            else if (o == 2) v = (4.0 * qn[m] + 3.0 * qn1[m] + 9.0 * qn2[m] + 2. * qn3[m]) / 18.0;
            else v = (-2.0 * qn[m] + 3.0 * qn1[m] + 3.0 * qn2[m] + 2. * qn3[m]) / 6.0;
            ql[m] = v; qr[m] = v;
        }
        half_metric(pc, i, nx);
        checked_flux<FLUX>(pc, i, ql, qr, nx, f[o], efix, gama, 2, nerr, bad);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) dfn[m] = -(-23.0 * f[0][m] + 21.0 * f[1][m] + 3.0 * f[2][m] - 1.0 * f[3][m]) / 24.0;
    // (5) faces n-1..n-5 -> df(n-1)
    double s5r[6][5];
    for (int i = n - 1; i >= n - 5; i--) {
        int o = n - i;
        for (int m = 0; m < 5; m++) {
            double vl, vr;
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
__device__ __noinline__ void closure_state(const Pencil &pc, int which, double *out)
{
    double q1[5], q2[5], q3[5], q4[5];
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

// DIR0 = 3: the i direction run as j pencils of the transposed view
template <int DIR0, int FLUX>
__global__ void __launch_bounds__(128) k_rhs_inv(BlockGrid G, DFields F, DParams P)
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
    int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;

    Pencil pc;
    pc.r = F.r; pc.u = F.u; pc.v = F.v; pc.w = F.w; pc.p = F.p;
    pc.mx = DIR == 0 ? F.kcx : (DIR == 1 ? F.etx : F.ctx);
    pc.my = DIR == 0 ? F.kcy : (DIR == 1 ? F.ety : F.cty);
    pc.mz = DIR == 0 ? F.kcz : (DIR == 1 ? F.etz : F.ctz);
    pc.base = gidx(b, i0, j0, k0);
    pc.stride = DIR == 0 ? 1 : (DIR == 1 ? b.sx : b.sy);
    pc.n = n;

    const double gama = P.gama, efix = P.efix;
    const bool s41 = P.nscheme == 41;
    // physical-boundary flags: negative density in ghost layer 3 (WCNSE5.f90:2123-2133)
    const int ist = pc.r[pc.base + (long)(-2) * pc.stride] < SMALL ? 2 : 0;
    const int ied = pc.r[pc.base + (long)(n + 3) * pc.stride] < SMALL ? n - 2 : n;
    const bool lclos = ist > 1, rclos = ied < n;

    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    int F0 = a - 3 < 0 ? 0 : a - 3;
    int F1 = bb + 2 > n ? n : bb + 2;
    if (a <= 2 && F1 < 3) F1 = 3;
    if (bb >= n - 1 && F0 > n - 3) F0 = n - 3;

    double qw[5][5];       // nodes nd-2..nd+2
    double fw[6][5];       // faces fc-5..fc
    double qwl_prev[5];
    int nerr = 0, bad = 0;

    for (int l = 0; l < 4; l++) load_q(pc, F0 - 2 + l, qw[l + 1]);   // qw[1..4] = F0-2..F0+1
    for (int nd = F0; nd <= F1 + 1; nd++) {
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) qw[l][m] = qw[l + 1][m];
        load_q(pc, nd + 2, qw[4]);
        double qwl_cur[5], qwr_m1[5];
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
            double nx[3];
            half_metric(pc, fc, nx);
            double ql[5], qr[5];
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
            double df[5];
            if (fc == 3) {
                if (a <= 1 && !(s41 && lclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = (-23.0 * fw[2][m] + 21.0 * fw[3][m] + 3.0 * fw[4][m] - fw[5][m]) / 24.0;
                    long g = pc.base + pc.stride;
#pragma unroll
                    for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] + df[m];
                }
                if (a <= 2 && bb >= 2 && !(s41 && lclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = (fw[2][m] - 27.0 * fw[3][m] + 27.0 * fw[4][m] - fw[5][m]) / 24.0;
                    long g = pc.base + 2 * pc.stride;
#pragma unroll
                    for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] + df[m];
                }
            }
            int row = fc - 2;
            if (row >= 3 && row <= n - 2 && row >= a && row <= bb &&
                !(s41 && lclos && row == 3) && !(s41 && rclos && row == n - 2)) {
#pragma unroll
                for (int m = 0; m < 5; m++)
                    df[m] = (2250.0 * (fw[3][m] - fw[2][m]) - 125.0 * (fw[4][m] - fw[1][m]) +
                             9.0 * (fw[5][m] - fw[0][m])) / 1920.0;
                long g = pc.base + (long)row * pc.stride;
#pragma unroll
                for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] + df[m];
            }
            if (fc == n) {
                if (bb >= n - 1 && a <= n - 1 && !(s41 && rclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = -(fw[5][m] - 27.0 * fw[4][m] + 27.0 * fw[3][m] - fw[2][m]) / 24.0;
                    long g = pc.base + (long)(n - 1) * pc.stride;
#pragma unroll
                    for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] + df[m];
                }
                if (bb >= n && !(s41 && rclos)) {
#pragma unroll
                    for (int m = 0; m < 5; m++)
                        df[m] = -(-23.0 * fw[5][m] + 21.0 * fw[4][m] + 3.0 * fw[3][m] - fw[2][m]) / 24.0;
                    long g = pc.base + (long)n * pc.stride;
#pragma unroll
                    for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] + df[m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) qwl_prev[m] = qwl_cur[m];
    }

    if (s41 && lclos && a == 1) {
        double d1[5], d2[5], d3[5];
        left41<FLUX>(pc, d1, d2, d3, efix, gama, nerr, bad);
        long g1 = pc.base + pc.stride, g2 = pc.base + 2 * pc.stride, g3 = pc.base + 3 * pc.stride;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            F.rhs[m * F.total + g1] = F.rhs[m * F.total + g1] + d1[m];
            F.rhs[m * F.total + g2] = F.rhs[m * F.total + g2] + d2[m];
            F.rhs[m * F.total + g3] = F.rhs[m * F.total + g3] + d3[m];
        }
    }
    if (s41 && rclos && bb == n) {
        double dn[5], dn1[5], dn2[5];
        right41<FLUX>(pc, dn, dn1, dn2, efix, gama, nerr, bad);
        long g0 = pc.base + (long)n * pc.stride, g1 = pc.base + (long)(n - 1) * pc.stride;
        long g2 = pc.base + (long)(n - 2) * pc.stride;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            F.rhs[m * F.total + g0] = F.rhs[m * F.total + g0] + dn[m];
            F.rhs[m * F.total + g1] = F.rhs[m * F.total + g1] + dn1[m];
            F.rhs[m * F.total + g2] = F.rhs[m * F.total + g2] + dn2[m];
        }
    }
    if (nerr) atomicAdd(&F.counters[0], (unsigned long long)nerr);
    if (bad) atomicAdd(&F.counters[2], 1ull);
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

int launch_rhs_inv_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    dim3 grid((unsigned)G.nwg), blk(128);
    if (P.nflux == 3) {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv<0, 3>), grid, blk, 0, s, G, F, P);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv<1, 3>), grid, blk, 0, s, G, F, P);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv<2, 3>), grid, blk, 0, s, G, F, P);
        if (dir == 3) hipLaunchKernelGGL((k_rhs_inv<3, 3>), grid, blk, 0, s, G, F, P);
    } else {
        if (dir == 0) hipLaunchKernelGGL((k_rhs_inv<0, 1>), grid, blk, 0, s, G, F, P);
        if (dir == 1) hipLaunchKernelGGL((k_rhs_inv<1, 1>), grid, blk, 0, s, G, F, P);
        if (dir == 2) hipLaunchKernelGGL((k_rhs_inv<2, 1>), grid, blk, 0, s, G, F, P);
        if (dir == 3) hipLaunchKernelGGL((k_rhs_inv<3, 1>), grid, blk, 0, s, G, F, P);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
