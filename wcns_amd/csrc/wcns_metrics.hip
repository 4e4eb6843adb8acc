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

// VALUE_HALF_NODE of one component at half node f (0..n)
__device__ double half_at(const Line &Q, int f)
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
__device__ double dxyz_at(const Line &Q, int i)
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

}  // namespace wcns
