// wcns_metrics.hip -- conservative (Visbal-Gaitonde) grid metrics on the
// device, GRID_DERIVATIVE_gcl with gcl = 1 (reference src/WCNSE5.f90:1478-1728):
// the same node->half interpolation (VALUE_HALF_NODE, :799-852) and flux
// difference (FLUX_DXYZ, :209-254) as the scheme, so that a uniform flow is
// preserved to round-off on any grid.  Runs once at setup.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

struct Line {
    const double *q;
    long base, stride;   // q(i) at base + i*stride, i = 1..n
    int n;
    __device__ double operator()(int i) const { return q[base + (long)i * stride]; }
};

// a unit vector along a line: probing the (linear) formulas below with it
// yields their coefficient of node `at`, which transposes them
struct UnitLine {
    int at, n;
    __device__ double operator()(int i) const { return i == at ? 1.0 : 0.0; }
};

// VALUE_HALF_NODE of one component at half node f (0..n)
template <class S>
__device__ double half_at(const S &Q, int f)
{
    int n = Q.n;
    if (n <= 4) {
        if (f >= 1 && f <= n - 1) return 0.5 * (Q(f) + Q(f + 1));
        if (f == 0) return (15.0 * Q(1) - 10.0 * Q(n < 2 ? n : 2) + 3.0 * Q(n < 3 ? n : 3)) / 8.0;
        return (15.0 * Q(n) - 10.0 * Q(n - 1 > 1 ? n - 1 : 1) + 3.0 * Q(n - 2 > 1 ? n - 2 : 1)) / 8.0;
    }
    if (f >= 2 && f <= n - 2) return (9.0 * (Q(f) + Q(f + 1)) + -1.0 * (Q(f + 2) + Q(f - 1))) / 16.0;
    if (f == 1) return (5.0 * Q(1) + 15.0 * Q(2) + -5.0 * Q(3) + 1.0 * Q(4)) / 16.0;
    if (f == n - 1) return (5.0 * Q(n) + 15.0 * Q(n - 1) + -5.0 * Q(n - 2) + 1.0 * Q(n - 3)) / 16.0;
    if (f == 0) return (35.0 * Q(1) - 35.0 * Q(2) + 21.0 * Q(3) - 5.0 * Q(4)) / 16.0;
    return (35.0 * Q(n) - 35.0 * Q(n - 1) + 21.0 * Q(n - 2) - 5.0 * Q(n - 3)) / 16.0;
}

// FLUX_DXYZ of the half-node values at row i (1..n)
template <class S>
__device__ double dxyz_at(const S &Q, int i)
{
    int n = Q.n;
    if (n <= 4) return half_at(Q, i) - half_at(Q, i - 1);
    if (i >= 3 && i <= n - 2)
        return (2250.0 * (half_at(Q, i) - half_at(Q, i - 1)) - 125.0 * (half_at(Q, i + 1) - half_at(Q, i - 2)) +
                9.0 * (half_at(Q, i + 2) - half_at(Q, i - 3))) / 1920.0;
    if (i == 2) return (half_at(Q, 0) - 27.0 * half_at(Q, 1) + 27.0 * half_at(Q, 2) - half_at(Q, 3)) / 24.0;
    if (i == n - 1)
        return -(half_at(Q, n) - 27.0 * half_at(Q, n - 1) + 27.0 * half_at(Q, n - 2) - half_at(Q, n - 3)) / 24.0;
    if (i == 1) return (-23.0 * half_at(Q, 0) + 21.0 * half_at(Q, 1) + 3.0 * half_at(Q, 2) - half_at(Q, 3)) / 24.0;
    return -(-23.0 * half_at(Q, n) + 21.0 * half_at(Q, n - 1) + 3.0 * half_at(Q, n - 2) - half_at(Q, n - 3)) / 24.0;
}

// node coordinates are in PLOT3D order (1-based (i,j,k) at (i-1)+ni*((j-1)+nj*(k-1)));
// the work array W holds 21 G3-shaped slots: 18 derivatives + x_kc, x_et, x_ct.
__device__ inline long xidx(const DBlock &b, int i, int j, int k)
{
    return (long)(i - 1) + (long)b.ni * ((long)(j - 1) + (long)b.nj * (long)(k - 1));
}

// Pass A: d(x,y,z)/d(dir) -> slots c*3+dir (c = x,y,z)
__global__ void k_gcl_a(DBlock b, const double *x, const double *y, const double *z, double *W, long slot, int dir)
{
    int na = dir == 0 ? b.nj : b.ni, nb = dir == 2 ? b.nj : b.nk;
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3L * na * nb) return;
    int c = (int)(t % 3);
    long pen = t / 3;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    int i0 = dir == 0 ? 0 : ia, j0 = dir == 1 ? 0 : (dir == 0 ? ia : ib), k0 = dir == 2 ? 0 : ib;
    Line Q;
    Q.q = c == 0 ? x : (c == 1 ? y : z);
    long st = dir == 0 ? 1 : (dir == 1 ? b.ni : (long)b.ni * b.nj);
    Q.stride = st;
    Q.base = (dir == 0 ? xidx(b, 1, j0, k0) : (dir == 1 ? xidx(b, i0, 1, k0) : xidx(b, i0, j0, 1))) - st;
    Q.n = n;
    double *out = W + (long)(c * 3 + dir) * slot;
    for (int l = 1; l <= n; l++) {
        int i = dir == 0 ? l : ia, j = dir == 1 ? l : (dir == 0 ? ia : ib), k = dir == 2 ? l : ib;
        out[gidx(b, i, j, k) - b.off] = dxyz_at(Q, l);
    }
}

// Pass B: keep x_kc,x_et,x_ct; metric arrays temporarily hold the products
__global__ void k_gcl_b(DBlock b, const double *x, const double *y, const double *z, double *W, long slot,
                        DFields F)
{
    long n = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k), w = g - b.off;
    double D[9];
    for (int s = 0; s < 9; s++) D[s] = W[s * slot + w];
    W[18 * slot + w] = D[0];
    W[19 * slot + w] = D[1];
    W[20 * slot + w] = D[2];
    long xi = xidx(b, i, j, k);
    double xm = x[xi], ym = y[xi], zm = z[xi];
    double *m[9] = {(double *)F.kcx, (double *)F.kcy, (double *)F.kcz, (double *)F.etx, (double *)F.ety,
                    (double *)F.etz, (double *)F.ctx, (double *)F.cty, (double *)F.ctz};
    m[0][g] = D[0] * ym; m[1][g] = D[1] * ym; m[2][g] = D[2] * ym;
    m[3][g] = D[3] * zm; m[4][g] = D[4] * zm; m[5][g] = D[5] * zm;
    m[6][g] = D[6] * xm; m[7][g] = D[7] * xm; m[8][g] = D[8] * xm;
}

// Pass C: derivatives of six products along dir -> slots 6*dir .. 6*dir+5
__global__ void k_gcl_c(DBlock b, double *W, long slot, DFields F, int dir)
{
    int na = dir == 0 ? b.nj : b.ni, nb = dir == 2 ? b.nj : b.nk;
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6L * na * nb) return;
    int c = (int)(t % 6);
    long pen = t / 6;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    int i0 = dir == 0 ? 0 : ia, j0 = dir == 1 ? 0 : (dir == 0 ? ia : ib), k0 = dir == 2 ? 0 : ib;
    // WCNSE5.f90:1597-1604, 1624-1631, 1651-1658
    const double *sel0[6] = {F.cty, F.ctz, F.kcy, F.kcz, F.ety, F.etz};
    const double *sel1[6] = {F.ctx, F.ctz, F.kcx, F.kcz, F.etx, F.etz};
    const double *sel2[6] = {F.ctx, F.cty, F.kcx, F.kcy, F.etx, F.ety};
    Line Q;
    Q.q = dir == 0 ? sel0[c] : (dir == 1 ? sel1[c] : sel2[c]);
    Q.stride = dir == 0 ? 1 : (dir == 1 ? b.sx : b.sy);
    Q.base = gidx(b, i0, j0, k0);
    Q.n = n;
    double *out = W + (long)(6 * dir + c) * slot;
    for (int l = 1; l <= n; l++) {
        int i = dir == 0 ? l : ia, j = dir == 1 ? l : (dir == 0 ? ia : ib), k = dir == 2 ? l : ib;
        out[gidx(b, i, j, k) - b.off] = dxyz_at(Q, l);
    }
}

// Pass D: WCNSE5.f90:1691-1715
__global__ void k_gcl_d(DBlock b, double *W, long slot, DFields F, int *nbad)
{
    long n = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k), w = g - b.off;
    auto D = [&](int s) { return W[(long)(s - 1) * slot + w]; };
    double kcx = D(18) - D(12), kcy = D(14) - D(8), kcz = D(16) - D(10);
    double etx = D(6) - D(17), ety = D(2) - D(13), etz = D(4) - D(15);
    double ctx = D(11) - D(5), cty = D(7) - D(1), ctz = D(9) - D(3);
    double vol = W[18 * slot + w] * kcx + W[19 * slot + w] * etx + W[20 * slot + w] * ctx;
    double *m[10] = {(double *)F.kcx, (double *)F.kcy, (double *)F.kcz, (double *)F.etx, (double *)F.ety,
                     (double *)F.etz, (double *)F.ctx, (double *)F.cty, (double *)F.ctz, (double *)F.vol};
    m[0][g] = kcx; m[1][g] = kcy; m[2][g] = kcz;
    m[3][g] = etx; m[4][g] = ety; m[5][g] = etz;
    m[6][g] = ctx; m[7][g] = cty; m[8][g] = ctz;
    m[9][g] = vol;
    if (!(vol > 0.0)) atomicAdd(nbad, 1);
}

// ------------------------------------------------ tangent and adjoint of the metrics
// The metrics are linear line operators L_dir (FLUX_DXYZ of VALUE_HALF_NODE,
// half_at/dxyz_at above) and products:
//   A: D[3c+dir] = L_dir(x_c)                     (c = x, y, z)
//   B: Pm[s] = D[s] * (y, y, y, z, z, z, x, x, x)[s]   (s in metric-pool order)
//   C: C[6dir+c] = L_dir(Pm[sel_dir(c)])
//   D: metrics = differences of C;  vol = D[0]*kcx + D[1]*etx + D[2]*ctx.
// The tangent follows the same passes on directions; the adjoint runs them
// backwards, transposing each L_dir by probing it with unit vectors.
// Work slots are G3-shaped arrays in the block-local frame (gidx with off 0).

// WCNSE5.f90:1597-1604, 1624-1631, 1651-1658: the products differenced along dir
// (metric-pool index: 0 kcx, 1 kcy, 2 kcz, 3 etx, 4 ety, 5 etz, 6 ctx, 7 cty, 8 ctz)
__device__ inline int gcl_sel(int dir, int c)
{
    const int s0[6] = {7, 8, 1, 2, 4, 5}, s1[6] = {6, 8, 0, 2, 3, 5}, s2[6] = {6, 7, 0, 1, 3, 4};
    return dir == 0 ? s0[c] : (dir == 1 ? s1[c] : s2[c]);
}
// metric m = C[plus[m]] - C[minus[m]] (WCNSE5.f90:1691-1699, 0-based C slots)
__device__ inline void gcl_pm(int m, int &plus, int &minus)
{
    const int pl[9] = {17, 13, 15, 5, 1, 3, 10, 6, 8}, mi[9] = {11, 7, 9, 16, 12, 14, 4, 0, 2};
    plus = pl[m];
    minus = mi[m];
}

__device__ inline DBlock local_frame(DBlock b)
{
    b.off = 0;
    return b;
}

// dst = L_dir(src) on the nodes 1..n; src in PLOT3D order (p3d) or a local slot
__global__ void k_gcl_line(DBlock b, const double *src, int p3d, double *dst, int dir)
{
    int na = dir == 0 ? b.nj : b.ni, nb = dir == 2 ? b.nj : b.nk;
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)na * nb * n) return;
    int l = (int)(t % n) + 1;
    long pen = t / n;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    const DBlock lb = local_frame(b);
    int i = dir == 0 ? l : ia, j = dir == 1 ? l : (dir == 0 ? ia : ib), k = dir == 2 ? l : ib;
    Line Q;
    Q.q = src;
    Q.n = n;
    if (p3d) {
        Q.stride = dir == 0 ? 1 : (dir == 1 ? b.ni : (long)b.ni * b.nj);
        Q.base = xidx(b, i, j, k) - (long)l * Q.stride;
    } else {
        Q.stride = dir == 0 ? 1 : (dir == 1 ? lb.sx : lb.sy);
        Q.base = gidx(lb, i, j, k) - (long)l * Q.stride;
    }
    dst[gidx(lb, i, j, k)] = dxyz_at(Q, l);
}

// out += L_dir^T lam on the nodes 1..n; lam a local slot, out PLOT3D (p3d) or a local slot
__global__ void k_gcl_line_t(DBlock b, const double *lam, double *out, int p3d, int dir)
{
    long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    const DBlock lb = local_frame(b);
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    int l = dir == 0 ? i : (dir == 1 ? j : k);
    long st = dir == 0 ? 1 : (dir == 1 ? lb.sx : lb.sy);
    long g0 = gidx(lb, i, j, k) - (long)l * st;
    UnitLine U;
    U.at = l;
    U.n = n;
    double acc = 0.0;
    int r0 = l - 5 < 1 ? 1 : l - 5, r1 = l + 5 > n ? n : l + 5;
    for (int r = r0; r <= r1; r++) {
        double c = dxyz_at(U, r);
        if (c != 0.0) acc = acc + c * lam[g0 + (long)r * st];
    }
    long o = p3d ? xidx(b, i, j, k) : gidx(lb, i, j, k);
    out[o] = out[o] + acc;
}

// tangent of pass B: dPm[s] = dD[s] * q + D[s] * dq
__global__ void k_gcl_tb(DBlock b, const double *X, const double *dX, const double *D, const double *dD, double *dPm,
                         long slot)
{
    long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long w = gidx(local_frame(b), i, j, k), xi = xidx(b, i, j, k);
    for (int s = 0; s < 9; s++) {
        int c = s < 3 ? 1 : (s < 6 ? 2 : 0);
        dPm[s * slot + w] = dD[s * slot + w] * X[c * nn + xi] + D[s * slot + w] * dX[c * nn + xi];
    }
}

// tangent of pass D into dm (metric-pool frame, 10 arrays of `total`)
__global__ void k_gcl_td(DBlock b, DFields F, const double *D, const double *dD, const double *dC, long slot,
                         double *dm, long total)
{
    long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long w = gidx(local_frame(b), i, j, k), g = gidx(b, i, j, k);
    double d[9];
    for (int m = 0; m < 9; m++) {
        int pl, mi;
        gcl_pm(m, pl, mi);
        d[m] = dC[pl * slot + w] - dC[mi * slot + w];
        dm[m * total + g] = d[m];
    }
    dm[9 * total + g] = (dD[0 * slot + w] * F.kcx[g] + D[0 * slot + w] * d[0]) +
                        (dD[1 * slot + w] * F.etx[g] + D[1 * slot + w] * d[3]) +
                        (dD[2 * slot + w] * F.ctx[g] + D[2 * slot + w] * d[6]);
}

// adjoint of pass D: lamC (18 slots) and the volume's share of lamD[0..2]
__global__ void k_gcl_rd(DBlock b, DFields F, const double *D, const double *gm, long total, double *lamC,
                         double *lamD, long slot)
{
    long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long w = gidx(local_frame(b), i, j, k), g = gidx(b, i, j, k);
    double gv = gm[9 * total + g];
    double lm[9];
    for (int m = 0; m < 9; m++) lm[m] = gm[m * total + g];
    lm[0] = lm[0] + gv * D[0 * slot + w];
    lm[3] = lm[3] + gv * D[1 * slot + w];
    lm[6] = lm[6] + gv * D[2 * slot + w];
    lamD[0 * slot + w] = lamD[0 * slot + w] + gv * F.kcx[g];
    lamD[1 * slot + w] = lamD[1 * slot + w] + gv * F.etx[g];
    lamD[2 * slot + w] = lamD[2 * slot + w] + gv * F.ctx[g];
    for (int m = 0; m < 9; m++) {
        int pl, mi;
        gcl_pm(m, pl, mi);
        lamC[pl * slot + w] = lm[m];
        lamC[mi * slot + w] = -lm[m];
    }
}

// adjoint of pass C: lamPm[s] += sum over (dir, c) with sel = s of L_dir^T lamC[6dir+c]
// (through k_gcl_line_t); adjoint of pass B:
__global__ void k_gcl_rb(DBlock b, const double *X, const double *D, const double *lamPm, double *lamD, double *gX,
                         long slot)
{
    long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long w = gidx(local_frame(b), i, j, k), xi = xidx(b, i, j, k);
    double gq[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < 9; s++) {
        int c = s < 3 ? 1 : (s < 6 ? 2 : 0);
        double lp = lamPm[s * slot + w];
        lamD[s * slot + w] = lamD[s * slot + w] + lp * X[c * nn + xi];
        gq[c] = gq[c] + lp * D[s * slot + w];
    }
    for (int c = 0; c < 3; c++) gX[c * nn + xi] = gX[c * nn + xi] + gq[c];
}

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

// work: 21 * (ni+6)(nj+6)(nk+6) doubles + one int (bad-volume count) at the end
int launch_gcl_metrics(const DBlock &b, const double *x, const double *y, const double *z, DFields F,
                       double *work, hipStream_t s)
{
    long slot = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    int *nbad = (int *)(work + 21 * slot);
    long n = (long)b.ni * b.nj * b.nk;
    for (int dir = 0; dir < 3; dir++) {
        long np = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
        hipLaunchKernelGGL(k_gcl_a, grid1(3 * np, 64), dim3(64), 0, s, b, x, y, z, work, slot, dir);
    }
    hipLaunchKernelGGL(k_gcl_b, grid1(n, 256), dim3(256), 0, s, b, x, y, z, work, slot, F);
    for (int dir = 0; dir < 3; dir++) {
        long np = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
        hipLaunchKernelGGL(k_gcl_c, grid1(6 * np, 64), dim3(64), 0, s, b, work, slot, F, dir);
    }
    hipLaunchKernelGGL(k_gcl_d, grid1(n, 256), dim3(256), 0, s, b, work, slot, F, nbad);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// work slots: D 9, then tangent dD 9, dPm 9, dC 18 / adjoint lamC 18, lamD 9, lamPm 9
int gcl_ad_slots() { return 45; }

namespace {
int gcl_sel_host(int dir, int c)
{
    static const int s0[6] = {7, 8, 1, 2, 4, 5}, s1[6] = {6, 8, 0, 2, 3, 5}, s2[6] = {6, 7, 0, 1, 3, 4};
    return dir == 0 ? s0[c] : (dir == 1 ? s1[c] : s2[c]);
}
inline long g3_slot(const DBlock &b) { return (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6); }
inline long pencil_nodes(const DBlock &b) { return (long)b.ni * b.nj * b.nk; }
// D = pass A of the primal (slots 0..8)
void gcl_pass_a(const DBlock &b, const double *X, double *D, hipStream_t s)
{
    long slot = g3_slot(b), nn = pencil_nodes(b);
    for (int c = 0; c < 3; c++)
        for (int dir = 0; dir < 3; dir++)
            hipLaunchKernelGGL(k_gcl_line, grid1(nn, 128), dim3(128), 0, s, b, X + c * nn, 1, D + (3 * c + dir) * slot,
                               dir);
}
}  // namespace

int launch_gcl_jvp(const DBlock &b, const DFields &F, const double *X, const double *dX, double *dm, double *work,
                   hipStream_t s)
{
    long slot = g3_slot(b), nn = pencil_nodes(b);
    double *D = work, *dD = work + 9 * slot, *dPm = work + 18 * slot, *dC = work + 27 * slot;
    if (hipMemsetAsync(work, 0, sizeof(double) * 45 * slot, s) != hipSuccess) return -3;
    gcl_pass_a(b, X, D, s);
    for (int c = 0; c < 3; c++)
        for (int dir = 0; dir < 3; dir++)
            hipLaunchKernelGGL(k_gcl_line, grid1(nn, 128), dim3(128), 0, s, b, dX + c * nn, 1,
                               dD + (3 * c + dir) * slot, dir);
    hipLaunchKernelGGL(k_gcl_tb, grid1(nn, 256), dim3(256), 0, s, b, X, dX, (const double *)D, (const double *)dD,
                       dPm, slot);
    for (int dir = 0; dir < 3; dir++)
        for (int c = 0; c < 6; c++) {
            int sel = gcl_sel_host(dir, c);
            hipLaunchKernelGGL(k_gcl_line, grid1(nn, 128), dim3(128), 0, s, b, (const double *)(dPm + sel * slot), 0,
                               dC + (6 * dir + c) * slot, dir);
        }
    hipLaunchKernelGGL(k_gcl_td, grid1(nn, 256), dim3(256), 0, s, b, F, (const double *)D, (const double *)dD,
                       (const double *)dC, slot, dm, F.total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_gcl_vjp(const DBlock &b, const DFields &F, const double *X, const double *gm, double *gX, double *work,
                   hipStream_t s)
{
    long slot = g3_slot(b), nn = pencil_nodes(b);
    double *D = work, *lamC = work + 9 * slot, *lamD = work + 27 * slot, *lamPm = work + 36 * slot;
    if (hipMemsetAsync(work, 0, sizeof(double) * 45 * slot, s) != hipSuccess) return -3;
    gcl_pass_a(b, X, D, s);
    hipLaunchKernelGGL(k_gcl_rd, grid1(nn, 256), dim3(256), 0, s, b, F, (const double *)D, gm, F.total, lamC, lamD,
                       slot);
    for (int dir = 0; dir < 3; dir++)
        for (int c = 0; c < 6; c++) {
            int sel = gcl_sel_host(dir, c);
            hipLaunchKernelGGL(k_gcl_line_t, grid1(nn, 256), dim3(256), 0, s, b, (const double *)(lamC + (6 * dir + c) * slot),
                               lamPm + sel * slot, 0, dir);
        }
    hipLaunchKernelGGL(k_gcl_rb, grid1(nn, 256), dim3(256), 0, s, b, X, (const double *)D, (const double *)lamPm,
                       lamD, gX, slot);
    for (int c = 0; c < 3; c++)
        for (int dir = 0; dir < 3; dir++)
            hipLaunchKernelGGL(k_gcl_line_t, grid1(nn, 256), dim3(256), 0, s, b,
                               (const double *)(lamD + (3 * c + dir) * slot), gX + c * nn, 1, dir);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
