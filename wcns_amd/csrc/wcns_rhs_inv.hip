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
    static constexpr bool PRE_ROW = true;    // rows are plain adds to rhs: read one iteration ahead
    const double *r, *u, *v, *w, *p;
    const double *mx, *my, *mz;    // metric triple of the pencil direction
    double *rhs;
    long total;
    // store: rhs is the transposed i-direction contribution, whose every
    // interior row one thread writes once (0 + df: the same bits as adding to
    // the zeroed buffer); otherwise the rows are added to rhs
    bool store;
    __device__ void q(long g, int, double *o) const { o[0] = r[g]; o[1] = u[g]; o[2] = v[g]; o[3] = w[g]; o[4] = p[g]; }
    __device__ void m(long g, int, double *o) const { o[0] = mx[g]; o[1] = my[g]; o[2] = mz[g]; }
    __device__ void row(long g, int, const double *df) const
    {
#pragma unroll
        for (int c = 0; c < 5; c++) rhs[c * total + g] = (store ? 0.0 : rhs[c * total + g]) + df[c];
    }
    __device__ void row_read(long g, double *o) const
    {
#pragma unroll
        for (int c = 0; c < 5; c++) o[c] = rhs[c * total + g];
    }
    __device__ void row_write(long g, const double *pre, const double *df) const
    {
#pragma unroll
        for (int c = 0; c < 5; c++) rhs[c * total + g] = (store ? 0.0 : pre[c]) + df[c];
    }
};

// InvIO for the two-wave register budget: a type of its own, so that the
// boundary functions it calls (not inlined) are compiled once per budget.
struct InvIO2w : InvIO {};

struct InvTanIO {
    using T = D1;
    static constexpr int SEGLEN = SEG;
    static constexpr bool COUNT = false;
    static constexpr bool PRE_ROW = false;
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
    static constexpr bool PRE_ROW = false;
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
    // schemes 45 / 47: the pencil's ghost metrics, nodes 0,-1,-2,-3 (gml) and
    // n+1..n+4 (gmr), three components each
    const double *gml, *gmr;
    // flux 4: pressure ratios of direction d at rp + d * tot; own direction
    // dir, the other two xd[] with their index steps xs[]
    const double *rp;
    long tot, xs[2];
    int dir, xd[2];
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

// The metric triple of node l, l in -3..n+4: the ghost records beyond the ends.
template <class IO>
__device__ inline void load_mx(const Pencil<IO> &pc, int l, typename IO::T *t)
{
    if (l >= 1 && l <= pc.n) { load_m(pc, l, t); return; }
    const double *g = l < 1 ? pc.gml + 3 * (-l) : pc.gmr + 3 * (l - pc.n - 1);
#pragma unroll
    for (int m = 0; m < 3; m++) t[m] = g[m];
}

// VALUE_HALF_NODE_45 (WCNSE5.f90:5052-5123) at face f in -2..n+2: faces 0..n
// as VALUE_HALF_NODE, the two beyond each end from the ghost metrics.
template <class IO>
__device__ inline void half_metric_x(const Pencil<IO> &pc, int f, typename IO::T *nx)
{
    using T = typename IO::T;
    const int n = pc.n;
    if (f >= 0 && f <= n) { half_metric(pc, f, nx); return; }
    T a[3], b[3], c[3], d[3];
    if (f == -1 || f == n + 1) {
        const int o = f == -1 ? 0 : n + 1, sg = f == -1 ? -1 : 1;
        load_mx(pc, o, a); load_mx(pc, o + sg, b); load_mx(pc, o + 2 * sg, c); load_mx(pc, o + 3 * sg, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (5.0 * a[m] + 15.0 * b[m] + -5.0 * c[m] + 1.0 * d[m]) / 16.0;
    } else if (f == -2) {
        load_mx(pc, -2, a); load_mx(pc, -1, b); load_mx(pc, 0, c); load_mx(pc, -3, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (9.0 * (a[m] + b[m]) + -1.0 * (c[m] + d[m])) / 16.0;
    } else {
        load_mx(pc, n + 2, a); load_mx(pc, n + 3, b); load_mx(pc, n + 4, c); load_mx(pc, n + 1, d);
#pragma unroll
        for (int m = 0; m < 3; m++) nx[m] = (9.0 * (a[m] + b[m]) + -1.0 * (c[m] + d[m])) / 16.0;
    }
}

__device__ inline double dmin(double a, double b) { return b < a ? b : a; }

// Convect.f90:169-177 / 335-343 / 499-507: the least pressure ratio around
// face f -- its own, then four of each other direction (the face's node and
// the one before it, at the cross index and the next).
template <class IO>
__device__ inline double face_rpmin(const Pencil<IO> &pc, int f)
{
    const long g = pc.base + (long)f * pc.stride;
    double v = pc.rp[(long)pc.dir * pc.tot + g];
#pragma unroll
    for (int x = 0; x < 2; x++) {
        const double *a = pc.rp + (long)pc.xd[x] * pc.tot;
        const long sx = pc.xs[x];
        v = dmin(dmin(dmin(dmin(a[g], a[g + sx]), a[g - pc.stride]), a[g - pc.stride + sx]), v);
    }
    return v;
}

// The face flux of invcode's dispatch (Convect.f90:18-47) at face f.
template <int FLUX, class IO>
__device__ inline bool flux_at(const Pencil<IO> &pc, int f, const typename IO::T *ql, const typename IO::T *qr,
                               const typename IO::T *nx, typename IO::T *flr, double efix, double gama)
{
    if constexpr (FLUX == 4) {
        flux_roemx(ql, qr, nx[0], nx[1], nx[2], face_rpmin(pc, f), flr, gama);
        return true;
    } else {
        return face_flux<FLUX>(ql, qr, nx[0], nx[1], nx[2], flr, efix, gama);
    }
}

// WCNS_E_5_45 (WCNSE5.f90:3492-3505): cq(i), the jump q(i+1) - q(i) in the
// characteristic variables of face i (Roe average of the |rho| states, face
// metric nxf).
__device__ inline void char45_jump(const double *qa, const double *qb, const double *nxf, double gama, double *cq)
{
    double al[5], ar[5], d[5], qave[5];
#pragma unroll
    for (int m = 0; m < 5; m++) { al[m] = qa[m]; ar[m] = qb[m]; }
    al[0] = fabs(al[0]);
    ar[0] = fabs(ar[0]);
#pragma unroll
    for (int m = 0; m < 5; m++) d[m] = ar[m] - al[m];
    pave(al, ar, qave, gama);
    lxq(qave, nxf, d, cq, gama);
}

// The nonlinear weights of one variable from its three (g, s) pairs
// (WCNSE5.f90:2149-2170, the same in every scheme).
__device__ inline void wcns_weights(double g1, double g2, double g3, double s1, double s2, double s3, double *wl,
                                    double *wr)
{
    const double EPS = 1.0e-6;
    double is1 = g1 * g1 + s1 * s1, is2 = g2 * g2 + s2 * s2, is3 = g3 * g3 + s3 * s3;
    double e1 = (EPS + is1) * (EPS + is1), e2 = (EPS + is2) * (EPS + is2), e3 = (EPS + is3) * (EPS + is3);
    double bl1 = (1.0 / 16.0) / e1, bl2 = (10.0 / 16.0) / e2, bl3 = (5.0 / 16.0) / e3;
    double br1 = (5.0 / 16.0) / e1, br2 = (10.0 / 16.0) / e2, br3 = (1.0 / 16.0) / e3;
    double sl = bl1 + bl2 + bl3, sr = br1 + br2 + br3;
    wl[0] = bl1 / sl; wl[1] = bl2 / sl; wl[2] = bl3 / sl;
    wr[0] = br1 / sr; wr[1] = br2 / sr; wr[2] = br3 / sr;
}

// WCNS_E_5_45 at node nd (WCNSE5.f90:3512-3580) from cq(nd-2..nd+1) and the
// nodes nd-2..nd+2: qwl(nd) through face nd's eigenvectors, qwr(nd-1)
// through face nd-1's.
__device__ inline void char45_node(const double (*cq)[5], const double (*qw)[5], const double *nxm1,
                                   const double *nx0, double gama, double *qwl, double *qwr)
{
    double al[5], ar[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double g1 = 0.5 * (3.0 * cq[1][m] - cq[0][m]);
        double g2 = 0.5 * (cq[2][m] + cq[1][m]);
        double g3 = 0.5 * (3.0 * cq[2][m] - cq[3][m]);
        double s1 = cq[1][m] - cq[0][m], s2 = cq[2][m] - cq[1][m], s3 = cq[3][m] - cq[2][m];
        double wl[3], wr[3];
        wcns_weights(g1, g2, g3, s1, s2, s3, wl, wr);
        al[m] = 0.125 * (wl[0] * (s1 + 4.0 * g1) + wl[1] * (s2 + 4.0 * g2) + wl[2] * (s3 + 4.0 * g3));
        ar[m] = 0.125 * (wr[0] * (s1 - 4.0 * g1) + wr[1] * (s2 - 4.0 * g2) + wr[2] * (s3 - 4.0 * g3));
    }
    double qave[5], t[5];
    pave(qw[2], qw[3], qave, gama);
    rxq(qave, nx0, al, t, gama);
#pragma unroll
    for (int m = 0; m < 5; m++) qwl[m] = qw[2][m] + t[m];
    pave(qw[1], qw[2], qave, gama);
    rxq(qave, nxm1, ar, t, gama);
#pragma unroll
    for (int m = 0; m < 5; m++) qwr[m] = qw[2][m] + t[m];
}

// WCNS_E_5_47 at node nd (WCNSE5.f90:4200-4278): differences of the
// primitives, taken with the node's own value into the characteristic
// variables of the node's state and metric mt, weighted there, and both
// states back through the same eigenvectors.
__device__ inline void char47_node(const double (*qw)[5], const double *mt, double gama, double *qwl, double *qwr)
{
    const double *q0 = qw[2];
    double lq[5], g[3][5], s[3][5], lg[3][5], ls[3][5];
    lxq(q0, mt, q0, lq, gama);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        g[0][m] = 0.5 * (qw[0][m] - 4.0 * qw[1][m] + 3.0 * qw[2][m]);
        g[1][m] = 0.5 * (qw[3][m] - qw[1][m]);
        g[2][m] = 0.5 * (-3.0 * qw[2][m] + 4.0 * qw[3][m] - qw[4][m]);
        s[0][m] = qw[0][m] - 2.0 * qw[1][m] + qw[2][m];
        s[1][m] = qw[1][m] - 2.0 * qw[2][m] + qw[3][m];
        s[2][m] = qw[2][m] - 2.0 * qw[3][m] + qw[4][m];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        lxq(q0, mt, g[k], lg[k], gama);
        lxq(q0, mt, s[k], ls[k], gama);
    }
    double a[5], b[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double wl[3], wr[3];
        wcns_weights(lg[0][m], lg[1][m], lg[2][m], ls[0][m], ls[1][m], ls[2][m], wl, wr);
        a[m] = lq[m] + 0.125 * (wl[0] * (ls[0][m] + 4.0 * lg[0][m]) + wl[1] * (ls[1][m] + 4.0 * lg[1][m]) +
                                wl[2] * (ls[2][m] + 4.0 * lg[2][m]));
        b[m] = lq[m] + 0.125 * (wr[0] * (ls[0][m] - 4.0 * lg[0][m]) + wr[1] * (ls[1][m] - 4.0 * lg[1][m]) +
                                wr[2] * (ls[2][m] - 4.0 * lg[2][m]));
    }
    rxq(q0, mt, a, qwl, gama);
    rxq(q0, mt, b, qwr, gama);
}

// The scheme's qwl(node) and qwr(node-1) straight from global memory
// (boundary rows): SCH 0 the variable-wise WCNS of schemes 4, 41, 42.
template <int SCH, class IO>
__device__ inline void wcns_at(const Pencil<IO> &pc, int node, typename IO::T *qwl, typename IO::T *qwr, double gama)
{
    typename IO::T q[5][5];
    for (int l = 0; l < 5; l++) load_q(pc, node - 2 + l, q[l]);
    if constexpr (SCH == 47) {
        double mt[3];
        load_mx(pc, node, mt);
        char47_node(q, mt, gama, qwl, qwr);
    } else if constexpr (SCH == 45) {
        double cq[4][5], nxf[3], nxm1[3], nx0[3];
        for (int l = 0; l < 4; l++) {
            half_metric_x(pc, node - 2 + l, nxf);
            char45_jump(q[l], q[l + 1], nxf, gama, cq[l]);
        }
        half_metric_x(pc, node - 1, nxm1);
        half_metric_x(pc, node, nx0);
        char45_node(cq, q, nxm1, nx0, gama, qwl, qwr);
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) wcns_node(q[0][m], q[1][m], q[2][m], q[3][m], q[4][m], qwl[m], qwr[m]);
    }
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
    if (!flux_at<FLUX>(pc, f, ql, qr, nx, flr, efix, gama)) bad = 1;
}

// The WCNS_E_5_41 left boundary rows 1..3 (WCNSE5.f90:2269-2513).
template <int SCH, int FLUX, class IO>
__device__ __noinline__ void left41(const Pencil<IO> &pc, typename IO::T *df1, typename IO::T *df2,
                                    typename IO::T *df3, double efix, double gama, bool s42, int &nerr, int &bad)
{
    using T = typename IO::T;
    T q1[5], q2[5], q3[5], q4[5], urold[5][5], tmp[5], nx[3], f[6][5], ql[5], qr[5];
    load_q(pc, 1, q1); load_q(pc, 2, q2); load_q(pc, 3, q3); load_q(pc, 4, q4);
    // u_r_old(1..4): qwr after the scheme-4 closures; qwr(1) = (5,15,-5,1)/16, qwr(2..4) WCNS
#pragma unroll
    for (int m = 0; m < 5; m++) urold[1][m] = (5.0 * q1[m] + 15.0 * q2[m] - 5.0 * q3[m] + q4[m]) / 16.0;
    for (int i = 2; i <= 4; i++) wcns_at<SCH>(pc, i + 1, tmp, urold[i], gama);
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
template <int SCH, int FLUX, class IO>
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
    for (int i = 2; i <= 4; i++) wcns_at<SCH>(pc, n - i, ulold[9 - i], tmp, gama);
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
template <int FLUX, class IO, int SCH = 0>
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
    // one iteration ahead: the next node, the next metric of the rolled
    // window mw = m(nd-2..nd+1), and (solver) the next interior residual row
    // (solver instances only: in the dual-number tangent and adjoint walks the
    // extra registers spill)
    constexpr bool PF = IO::PRE_ROW;
    T qnext[5], mw[4][3], mnext[3];
    if constexpr (PF) {
        load_q(pc, F0 + 2, qnext);
        for (int l = 0; l < 3; l++) load_m(pc, F0 - 2 + l, mw[l + 1]);
        load_m(pc, F0 + 1, mnext);
    }
    double rpre[PF ? 5 : 1];
    if constexpr (PF)
        if (F0 - 3 >= 1 && !pc.io.store) pc.io.row_read(pc.base + (long)(F0 - 3) * pc.stride, rpre);
    // scheme 45: cq(nd-2..nd+1), rolled with the node window
    double cqw[SCH == 45 ? 4 : 1][5];
    if constexpr (SCH == 45) {
        for (int l = 0; l < 3; l++) {
            double nxf[3];
            half_metric_x(pc, F0 - 2 + l, nxf);
            char45_jump(qw[1 + l], qw[2 + l], nxf, gama, cqw[1 + l]);
        }
    }
    for (int nd = F0; nd <= F1 + 1; nd++) {
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) qw[l][m] = qw[l + 1][m];
        if constexpr (PF) {
#pragma unroll
            for (int m = 0; m < 5; m++) qw[4][m] = qnext[m];
#pragma unroll
            for (int l = 0; l < 3; l++)
#pragma unroll
                for (int m = 0; m < 3; m++) mw[l][m] = mw[l + 1][m];
#pragma unroll
            for (int m = 0; m < 3; m++) mw[3][m] = mnext[m];
            if (nd <= F1) {
                load_q(pc, nd + 3, qnext);
                load_m(pc, nd + 2, mnext);
            }
        } else {
            load_q(pc, nd + 2, qw[4]);
        }
        if constexpr (SCH == 45) {
#pragma unroll
            for (int l = 0; l < 3; l++)
#pragma unroll
                for (int m = 0; m < 5; m++) cqw[l][m] = cqw[l + 1][m];
            double nxf[3];
            half_metric_x(pc, nd + 1, nxf);
            char45_jump(qw[3], qw[4], nxf, gama, cqw[3]);
        }
        T qwl_cur[5], qwr_m1[5];
        bool need = (nd >= ist && nd <= ied + 1);
        if (need) {
            if constexpr (SCH == 47) {
                double mt[3];
                load_mx(pc, nd, mt);
                char47_node(qw, mt, gama, qwl_cur, qwr_m1);
            } else if constexpr (SCH == 45) {
                double nxm1[3], nx0[3];
                half_metric_x(pc, nd - 1, nxm1);
                half_metric_x(pc, nd, nx0);
                char45_node(cqw, qw, nxm1, nx0, gama, qwl_cur, qwr_m1);
            } else {
#pragma unroll
                for (int m = 0; m < 5; m++)
                    wcns_node(qw[0][m], qw[1][m], qw[2][m], qw[3][m], qw[4][m], qwl_cur[m], qwr_m1[m]);
            }
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
            if (PF && fc >= 2 && fc <= n - 2) {
                // VALUE_HALF_NODE's interior formula on the rolled window
#pragma unroll
                for (int m = 0; m < 3; m++)
                    nx[m] = (9.0 * (mw[1][m] + mw[2][m]) + -1.0 * (mw[3][m] + mw[0][m])) / 16.0;
            } else {
                half_metric(pc, fc, nx);
            }
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
            if (!flux_at<FLUX>(pc, fc, ql, qr, nx, fw[5], efix, gama)) bad = 1;
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
                if constexpr (IO::PRE_ROW) pc.io.row_write(pc.base + (long)row * pc.stride, rpre, df);
                else emit_row(pc, row, df);
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
        // the interior row of the next iteration, nd-2
        if constexpr (IO::PRE_ROW)
            if (nd - 2 >= 1 && nd - 2 <= n && !pc.io.store) pc.io.row_read(pc.base + (long)(nd - 2) * pc.stride, rpre);
    }

    if (s41 && lclos && a == 1) {
        T d1[5], d2[5], d3[5];
        left41<SCH, FLUX>(pc, d1, d2, d3, efix, gama, s42, nerr, bad);
        emit_row(pc, 1, d1);
        emit_row(pc, 2, d2);
        emit_row(pc, 3, d3);
    }
    if (s41 && rclos && bb == n) {
        T dn[5], dn1[5], dn2[5];
        right41<SCH, FLUX>(pc, dn, dn1, dn2, efix, gama, s42, nerr, bad);
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

// The ghost-metric record of pencil (ia, ib) of direction d at end `side`
// (0 low, 1 high): per block, direction by direction, the low end's pencils
// then the high end's, 12 doubles each.
__device__ inline long gm_record(const DBlock &b, long base, int d, int side, int ia, int ib)
{
    const long a0 = (long)b.nj * b.nk, a1 = (long)b.ni * b.nk, a2 = (long)b.ni * b.nj;
    const long na = d == 0 ? b.nj : b.ni;
    long off = base;
    if (d >= 1) off += 2 * 12 * a0;
    if (d >= 2) off += 2 * 12 * a1;
    off += (long)side * 12 * (d == 0 ? a0 : (d == 1 ? a1 : a2));
    return off + 12 * ((long)(ib - 1) * na + (ia - 1));
}

__device__ inline void metric_triple_of(const DFields &F, int d, const double *&mx, const double *&my,
                                        const double *&mz)
{
    mx = d == 0 ? F.kcx : (d == 1 ? F.etx : F.ctx);
    my = d == 0 ? F.kcy : (d == 1 ? F.ety : F.cty);
    mz = d == 0 ? F.kcz : (d == 1 ? F.etz : F.ctz);
}

// DIR0 = 3: the i direction run as j pencils of the transposed view.
template <int DIR0, int FLUX, class IO, int SCH = 0>
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
    pc.io.store = DIR0 == 3;
    pc.r = F.r;
    pencil_at<DIR>(b, ia, ib, pc.base, pc.stride, pc.n);
    if constexpr (SCH != 0) {
        const long base = F.gmoff[G.wg[blockIdx.x].x];
        pc.gml = F.gmet + gm_record(b, base, DIR, 0, ia, ib);
        pc.gmr = F.gmet + gm_record(b, base, DIR, 1, ia, ib);
    }
    if constexpr (FLUX == 4) {
        pc.rp = F.rpres;
        pc.tot = F.total;
        pc.dir = DIR;
        pc.xd[0] = DIR == 0 ? 1 : 0;
        pc.xd[1] = DIR == 2 ? 1 : 2;
        pc.xs[0] = DIR == 0 ? b.sx : 1;
        pc.xs[1] = DIR == 2 ? b.sx : b.sy;
    }

    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    int nerr = 0, bad = 0;
    inv_pencil<FLUX, IO, SCH>(pc, a, bb, P.gama, P.efix, P.nscheme >= 41, P.nscheme >= 42, nerr, bad);
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

// Schemes 45 / 47 (SCH) and the RoeMx flux (FLUX 4), direct pencils only.
template <int DIR, int FLUX, int SCH>
__global__ void __launch_bounds__(128) k_rhs_inv_var(BlockGrid G, DFields F, DParams P)
{
    rhs_inv_thread<DIR, FLUX, InvIO, SCH>(G, F, P);
}

// Convect.f90:78-132: the pressure ratios of direction D along line (a, b)
// (the two other indices, each over 0..n+1), written over -1..n+1 with the
// ghost pair of a physical end (ghost density below small) taken from the
// first interior entry, in the reference's order.  Entries no line writes
// keep the +inf they were given once.
template <int D>
__global__ void __launch_bounds__(256) k_pressure_ratios(DBlock b, DFields F)
{
    const int n = D == 0 ? b.ni : (D == 1 ? b.nj : b.nk);
    const int na = (D == 0 ? b.nj : b.ni) + 2, nb = (D == 2 ? b.nj : b.nk) + 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)na * nb) return;
    const int la = (int)(t % na), lb = (int)(t / na);
    const int i0 = D == 0 ? 0 : la, j0 = D == 1 ? 0 : (D == 0 ? la : lb), k0 = D == 2 ? 0 : lb;
    const long g0 = gidx(b, i0, j0, k0);
    const long st = D == 0 ? 1 : (D == 1 ? b.sx : b.sy);
    double *rp = F.rpres + (long)D * F.total;
    for (int l = -1; l <= n + 1; l++) {
        const double pl = F.p[g0 + l * st], pr = F.p[g0 + (l + 1) * st];
        rp[g0 + l * st] = dmin(pl / pr, pr / pl);
    }
    if (F.r[g0 - 2 * st] < SMALL) {
        rp[g0] = rp[g0 + st];
        rp[g0 - st] = rp[g0 + st];
    }
    if (F.r[g0 + (n + 3) * st] < SMALL) {
        rp[g0 + n * st] = rp[g0 + (n - 1) * st];
        rp[g0 + (n + 1) * st] = rp[g0 + (n - 1) * st];
    }
}

// One region's share of the ghost metrics (Convect.f90:180-297): for every
// face node, the four ghosts of the pencil that ends there -- the target
// block's metrics of its own normal direction 1..4 nodes inside its face,
// signed by the target side, for an interface; the end node's own metric
// repeated for any other region.  Launched region by region in the BC
// order, so that the last region holding a node decides (fill_bc_flag,
// PreProcess.f90:820-857).
__global__ void __launch_bounds__(256) k_ghost_metrics(DRegion g, DBlock b, DBlock bt, DFields F, double *gm,
                                                       long base)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    const int d = g.s_nd - 1, side = g.s_lr < 0 ? 0 : 1;
    const int ia = d == 0 ? j : i, ib = d == 2 ? j : k;
    double *o = gm + gm_record(b, base, d, side, ia, ib);
    const double *mx, *my, *mz;
    if (g.bctype < 0) {
        int it, jt, kt;
        region_image(g, i, j, k, it, jt, kt);
        const int td = g.t_l3[0] ? 0 : (g.t_l3[1] ? 1 : 2);
        const double sg = side == 0 ? (double)g.t_l3[td] : -(double)g.t_l3[td];
        metric_triple_of(F, td, mx, my, mz);
        for (int L = 1; L <= 4; L++) {
            const long q = gidx(bt, it - L * g.t_l3[0], jt - L * g.t_l3[1], kt - L * g.t_l3[2]);
            o[3 * (L - 1)] = sg * mx[q];
            o[3 * (L - 1) + 1] = sg * my[q];
            o[3 * (L - 1) + 2] = sg * mz[q];
        }
    } else {
        metric_triple_of(F, d, mx, my, mz);
        const long q = gidx(b, i, j, k);
        for (int L = 1; L <= 4; L++) {
            o[3 * (L - 1)] = mx[q];
            o[3 * (L - 1) + 1] = my[q];
            o[3 * (L - 1) + 2] = mz[q];
        }
    }
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

template <int FLUX, int SCH>
static void launch_var(int dir, dim3 grid, dim3 blk, hipStream_t s, const BlockGrid &G, const DFields &F,
                       const DParams &P)
{
    if (dir == 0) hipLaunchKernelGGL((k_rhs_inv_var<0, FLUX, SCH>), grid, blk, 0, s, G, F, P);
    if (dir == 1) hipLaunchKernelGGL((k_rhs_inv_var<1, FLUX, SCH>), grid, blk, 0, s, G, F, P);
    if (dir == 2) hipLaunchKernelGGL((k_rhs_inv_var<2, FLUX, SCH>), grid, blk, 0, s, G, F, P);
}

template <int SCH>
static void launch_var_flux(int nflux, int dir, dim3 grid, dim3 blk, hipStream_t s, const BlockGrid &G,
                            const DFields &F, const DParams &P)
{
    if (nflux == 4) launch_var<4, SCH>(dir, grid, blk, s, G, F, P);
    else if (nflux == 3) launch_var<3, SCH>(dir, grid, blk, s, G, F, P);
    else if (nflux == 2) launch_var<2, SCH>(dir, grid, blk, s, G, F, P);
    else launch_var<1, SCH>(dir, grid, blk, s, G, F, P);
}

bool inv_variant(const DParams &P) { return P.nscheme == 45 || P.nscheme == 47 || P.nflux == 4; }

int launch_rhs_inv_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    dim3 grid((unsigned)G.nwg), blk(128);
    if (inv_variant(P)) {
        // the variants run on the direct pencils only (dir 0..2)
        if (dir > 2) return -3;
        if (P.nscheme == 45) launch_var_flux<45>(P.nflux, dir, grid, blk, s, G, F, P);
        else if (P.nscheme == 47) launch_var_flux<47>(P.nflux, dir, grid, blk, s, G, F, P);
        else launch_var<4, 0>(dir, grid, blk, s, G, F, P);
        return hipGetLastError() == hipSuccess ? 0 : -3;
    }
    // two 64-wide waves per workgroup; the two-wave budget once they exceed
    // 1.5 waves per SIMD
    const bool two = P.inv_waves ? P.inv_waves == 2 : 2L * G.nwg * 2 > 3L * simd_count();
    if (P.nflux == 3) launch_inv<3>(dir, two, grid, blk, s, G, F, P);
    else if (P.nflux == 2) launch_inv<2>(dir, two, grid, blk, s, G, F, P);
    else launch_inv<1>(dir, two, grid, blk, s, G, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_pressure_ratios(const DBlock &b, const DFields &F, hipStream_t s)
{
    for (int d = 0; d < 3; d++) {
        const long na = (d == 0 ? b.nj : b.ni) + 2, nb = (d == 2 ? b.nj : b.nk) + 2;
        dim3 grid((unsigned)((na * nb + 255) / 256)), blk(256);
        if (d == 0) hipLaunchKernelGGL(k_pressure_ratios<0>, grid, blk, 0, s, b, F);
        if (d == 1) hipLaunchKernelGGL(k_pressure_ratios<1>, grid, blk, 0, s, b, F);
        if (d == 2) hipLaunchKernelGGL(k_pressure_ratios<2>, grid, blk, 0, s, b, F);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_ghost_metrics(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F, double *gm, long base,
                         hipStream_t s)
{
    if (g.npts <= 0) return 0;
    hipLaunchKernelGGL(k_ghost_metrics, dim3((unsigned)((g.npts + 255) / 256)), dim3(256), 0, s, g, b, bt, F, gm,
                       base);
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
