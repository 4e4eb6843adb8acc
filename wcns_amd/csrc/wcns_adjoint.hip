// wcns_adjoint.hip -- pointwise pieces of the residual's tangent and adjoint:
// the temperature and laminar viscosity of every node (get_c_t_only,
// RHS.f90:307-409; compute_visl_ns, GasModel.f90:219-235) and the packing of
// interior-node arrays.  The pencil parts live beside their solver kernels
// (wcns_rhs_inv.hip, wcns_rhs_vis.hip), the metrics' in wcns_metrics.hip and
// the wall force's in wcns_force.hip.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

// T = M^2 gamma p / rho on every node of the ghosted block, mu = Sutherland(T)
// on the interior (the only nodes whose mu the residual reads), in D1
// arithmetic: value and direction.
__device__ inline void tmu(const DParams &P, D1 r, D1 p, D1 &t, D1 &mu)
{
    double moo2 = P.moo * P.moo;
    D1 a2 = P.gama * p / r;
    t = moo2 * a2;
    mu = t * sqrt(t) * (1.0 + P.visc) / (t + P.visc);
}

__global__ void k_tmu_jvp(DBlock b, DFields F, DParams P, double *dt, double *dvisl, const double *dp)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e) return;
    const long g = b.off + t;
    const double rv = F.r[g];
    if (rv == 0.0) return;
    D1 tt, mu;
    tmu(P, D1(rv, dp[g]), D1(F.p[g], dp[4 * F.total + g]), tt, mu);
    dt[g] = tt.d;
    dvisl[g] = mu.d;
}

__global__ void k_tmu_vjp(DBlock b, DFields F, DParams P, const double *gt, const double *gvisl, double *gp)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e) return;
    const long g = b.off + t;
    const double rv = F.r[g];
    if (rv == 0.0) return;
    const int i = (int)(t % (b.ni + 6)) - 2, j = (int)((t / (b.ni + 6)) % (b.nj + 6)) - 2;
    const int k = (int)(t / ((long)(b.ni + 6) * (b.nj + 6))) - 2;
    const bool interior = i >= 1 && i <= b.ni && j >= 1 && j <= b.nj && k >= 1 && k <= b.nk;
    const double lt = gt[g], lm = interior ? gvisl[g] : 0.0;
    if (lt == 0.0 && lm == 0.0) return;
    D1 tt, mu;
    tmu(P, D1(rv, 1.0), D1(F.p[g]), tt, mu);
    const double dr = lt * tt.d + lm * mu.d;
    tmu(P, D1(rv), D1(F.p[g], 1.0), tt, mu);
    const double dpp = lt * tt.d + lm * mu.d;
    gp[g] = gp[g] + dr;
    gp[4 * F.total + g] = gp[4 * F.total + g] + dpp;
}

// compact (nv arrays of ni*nj*nk, i fastest) <-> pool arrays at the interior nodes
__global__ void k_interior(DBlock b, double *pool, long total, double *compact, int nv, int to_pool)
{
    const long nn = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    for (int m = 0; m < nv; m++) {
        if (to_pool) pool[m * total + g] = compact[m * nn + t];
        else compact[m * nn + t] = pool[m * total + g];
    }
}

}  // namespace

int launch_tmu_jvp(const DBlock &b, const DFields &F, const DParams &P, double *dt, double *dvisl, const double *dp,
                   hipStream_t s)
{
    long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    hipLaunchKernelGGL(k_tmu_jvp, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, b, F, P, dt, dvisl, dp);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_tmu_vjp(const DBlock &b, const DFields &F, const DParams &P, const double *gt, const double *gvisl,
                   double *gp, hipStream_t s)
{
    long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    hipLaunchKernelGGL(k_tmu_vjp, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, b, F, P, gt, gvisl, gp);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_interior(const DBlock &b, double *pool, long total, double *compact, int nv, int to_pool, hipStream_t s)
{
    long nn = (long)b.ni * b.nj * b.nk;
    hipLaunchKernelGGL(k_interior, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, b, pool, total, compact, nv,
                       to_pool);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
