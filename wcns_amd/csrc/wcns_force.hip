// wcns_force.hip -- aerodynamic force and moment on no-slip walls
// (PostProcess.f90:1415-1503 force, 1505-1680 force_dif, 1684-1796 getforce).
//
// Two passes per wall region:
//   k_force_pn    one thread per face node: surface traction (pressure plus,
//                 for Navier-Stokes, the viscous stress) times the outward
//                 unit normal, as force_dif's first loop nest;
//   k_force_quad  one thread per face quad: four-triangle quadrature of
//                 getforce, then a fixed-shape tree reduction per workgroup.
// The partial sums are added in workgroup order by one thread, so the result
// is deterministic run to run (its summation order is not the reference's
// sequential one: agreement is to rounding, not bit-exact).
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

constexpr int FORCE_THREADS = 256;

// Sub1.f90:383-480 getuvwtder: one-sided second order on block faces, central inside
__device__ inline double uvwt_der(const double *a, const DBlock &b, int i, int j, int k, int dir)
{
    const double aa = 1.5, bb = -2.0, cc = 0.5;
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    int l = dir == 0 ? i : (dir == 1 ? j : k);
    int di = dir == 0, dj = dir == 1, dk = dir == 2;
    auto at = [&](int o) { return a[gidx(b, i + di * (o - l), j + dj * (o - l), k + dk * (o - l))]; };
    if (l == 1) return -(aa * at(l) + bb * at(2) + cc * at(3));
    if (l == n) return aa * at(l) + bb * at(l - 1) + cc * at(l - 2);
    return 0.5 * (at(l + 1) - at(l - 1));
}

__global__ void k_force_pn(DRegion g, DBlock b, DFields F, DParams P, double *pn3, int npts)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    long q = gidx(b, i, j, k);
    double nx, ny, nz;
    getnxyz(F, q, g.l3, nx, ny, nz);
    double pn = 2.0 * (F.p[q] - P.poo);
    double fx = g.s_lr * nx * pn, fy = g.s_lr * ny * pn, fz = g.s_lr * nz * pn;
    if (P.nvis == 1) {
        const double *uvwt[4] = {F.u, F.v, F.w, F.t};
        double dk[3], de[3], dc[3];
        for (int m = 0; m < 3; m++) {
            dk[m] = uvwt_der(uvwt[m], b, i, j, k, 0);
            de[m] = uvwt_der(uvwt[m], b, i, j, k, 1);
            dc[m] = uvwt_der(uvwt[m], b, i, j, k, 2);
        }
        double kx = F.kcx[q], ky = F.kcy[q], kz = F.kcz[q];
        double ex = F.etx[q], ey = F.ety[q], ez = F.etz[q];
        double cx = F.ctx[q], cy = F.cty[q], cz = F.ctz[q], vj = F.vol[q];
        // Sub1.f90:251-270 getduvwdxyz
        double dudx = (kx * dk[0] + ex * de[0] + cx * dc[0]) / vj;
        double dudy = (ky * dk[0] + ey * de[0] + cy * dc[0]) / vj;
        double dudz = (kz * dk[0] + ez * de[0] + cz * dc[0]) / vj;
        double dvdx = (kx * dk[1] + ex * de[1] + cx * dc[1]) / vj;
        double dvdy = (ky * dk[1] + ey * de[1] + cy * dc[1]) / vj;
        double dvdz = (kz * dk[1] + ez * de[1] + cz * dc[1]) / vj;
        double dwdx = (kx * dk[2] + ex * de[2] + cx * dc[2]) / vj;
        double dwdy = (ky * dk[2] + ey * de[2] + cy * dc[2]) / vj;
        double dwdz = (kz * dk[2] + ez * de[2] + cz * dc[2]) / vj;
        // Sub1.f90:181-198 stress; the 2/3 constant is the reference's 11-digit literal
        const double c23 = 0.66666666667;
        double vis = F.visl[q];
        double txx = vis * c23 * (2.0 * dudx - dvdy - dwdz);
        double tyy = vis * c23 * (2.0 * dvdy - dwdz - dudx);
        double tzz = vis * c23 * (2.0 * dwdz - dudx - dvdy);
        double txy = vis * (dudy + dvdx), txz = vis * (dudz + dwdx), tyz = vis * (dvdz + dwdy);
        double pv = 2.0 * g.s_lr;
        fx = fx - pv * (nx * txx + ny * txy + nz * txz) * P.re;
        fy = fy - pv * (nx * txy + ny * tyy + nz * tyz) * P.re;
        fz = fz - pv * (nx * txz + ny * tyz + nz * tzz) * P.re;
    }
    pn3[t] = fx;
    pn3[npts + t] = fy;
    pn3[2 * npts + t] = fz;
}

__device__ inline double lenth(const double *a, const double *b)
{
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Sub1.f90:865-878 halen (Heron; a negative radicand gives zero area)
__device__ inline double halen(double a, double b, double c)
{
    double p = 0.5 * (a + b + c);
    double S = p * (p - a) * (p - b) * (p - c);
    return S < 0 ? 0.0 : sqrt(S);
}

__global__ void __launch_bounds__(FORCE_THREADS)
k_force_quad(DRegion g, const double *pn3, const double *xyz, int npts, double3 ref, double *partial)
{
    __shared__ double red[6][FORCE_THREADS];
    int bi = g.s_hi[0] - g.s_lo[0] + 1, bj = g.s_hi[1] - g.s_lo[1] + 1, bk = g.s_hi[2] - g.s_lo[2] + 1;
    int id = 1 - abs(g.l3[0]), jd = 1 - abs(g.l3[1]), kd = 1 - abs(g.l3[2]);
    int qi = bi - id, qj = bj - jd, qk = bk - kd;
    long nq = (long)qi * qj * qk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d[6] = {0, 0, 0, 0, 0, 0};
    if (t < nq) {
        int a = (int)(t % qi), c = (int)((t / qi) % qj), e = (int)(t / ((long)qi * qj));
        // corners 1..4 of force_dif (PostProcess.f90:1577-1613), in region-local offsets
        int o[4][3] = {{a, c, e}, {a, c, e}, {a, c, e}, {a, c, e}};
        if (id == 0) { o[1][1] += 1; o[2][1] += 1; o[2][2] += 1; o[3][2] += 1; }
        if (jd == 0) { o[1][0] += 1; o[2][0] += 1; o[2][2] += 1; o[3][2] += 1; }
        if (kd == 0) { o[1][0] += 1; o[2][0] += 1; o[2][1] += 1; o[3][1] += 1; }
        double P[4][3], X[5][3];
        for (int n = 0; n < 4; n++) {
            int pt = o[n][0] + bi * (o[n][1] + bj * o[n][2]);
            for (int m = 0; m < 3; m++) {
                P[n][m] = pn3[(long)m * npts + pt];
                X[n][m] = xyz[(long)m * npts + pt];
            }
        }
        double P5[3];
        for (int m = 0; m < 3; m++) {
            P5[m] = 0.25 * (P[0][m] + P[1][m] + P[2][m] + P[3][m]);
            X[4][m] = 0.25 * (X[0][m] + X[1][m] + X[2][m] + X[3][m]);
        }
        double r12 = lenth(X[0], X[1]), r23 = lenth(X[1], X[2]), r34 = lenth(X[2], X[3]), r41 = lenth(X[3], X[0]);
        double r15 = lenth(X[0], X[4]), r25 = lenth(X[1], X[4]), r35 = lenth(X[2], X[4]), r45 = lenth(X[3], X[4]);
        double s[4] = {halen(r12, r15, r25), halen(r23, r25, r35), halen(r34, r35, r45), halen(r41, r45, r15)};
        double rf[3] = {ref.x, ref.y, ref.z};
        for (int tr = 0; tr < 4; tr++) {
            int ia = tr, ib = (tr + 1) & 3;
            double pc[3], xc[3];
            for (int m = 0; m < 3; m++) {
                pc[m] = (P[ia][m] + P[ib][m] + P5[m]) / 3.0;
                xc[m] = (X[ia][m] + X[ib][m] + X[4][m]) / 3.0;
            }
            double fx = pc[0] * s[tr], fy = pc[1] * s[tr], fz = pc[2] * s[tr];
            d[0] = d[0] + fx; d[1] = d[1] + fy; d[2] = d[2] + fz;
            d[3] = d[3] + ((xc[1] - rf[1]) * fz - (xc[2] - rf[2]) * fy);
            d[4] = d[4] + ((xc[2] - rf[2]) * fx - (xc[0] - rf[0]) * fz);
            d[5] = d[5] + ((xc[0] - rf[0]) * fy - (xc[1] - rf[1]) * fx);
        }
    }
    for (int m = 0; m < 6; m++) red[m][threadIdx.x] = d[m];
    __syncthreads();
    for (int w = FORCE_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int m = 0; m < 6; m++) red[m][threadIdx.x] = red[m][threadIdx.x] + red[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_force_final(const double *partial, int nparts, double *out)
{
    if (threadIdx.x >= 6) return;
    double s = 0.0;
    for (int p = 0; p < nparts; p++) s = s + partial[p * 6 + threadIdx.x];
    out[threadIdx.x] = s;
}

}  // namespace

long force_quads(const DRegion &g)
{
    int bi = g.s_hi[0] - g.s_lo[0] + 1, bj = g.s_hi[1] - g.s_lo[1] + 1, bk = g.s_hi[2] - g.s_lo[2] + 1;
    return (long)(bi - (1 - abs(g.l3[0]))) * (bj - (1 - abs(g.l3[1]))) * (bk - (1 - abs(g.l3[2])));
}

int force_parts(const DRegion &g)
{
    return (int)((force_quads(g) + FORCE_THREADS - 1) / FORCE_THREADS);
}

int launch_force_region(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, const double *xyz,
                        double *pn3, double *partial, double *out6, const double ref[3], hipStream_t s)
{
    int npts = g.npts;
    if (npts <= 0) return 0;
    hipLaunchKernelGGL(k_force_pn, dim3((npts + 255) / 256), dim3(256), 0, s, g, b, F, P, pn3, npts);
    int np = force_parts(g);
    if (np > 0)
        hipLaunchKernelGGL(k_force_quad, dim3(np), dim3(FORCE_THREADS), 0, s, g, (const double *)pn3, xyz, npts,
                           make_double3(ref[0], ref[1], ref[2]), partial);
    hipLaunchKernelGGL(k_force_final, dim3(1), dim3(64), 0, s, (const double *)partial, np, out6);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace wcns
