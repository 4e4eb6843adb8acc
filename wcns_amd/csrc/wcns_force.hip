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

// tecout_bin_wall's values of one face node (PostProcess.f90:451-478):
// cp = 2 (p - poo); for Navier-Stokes also the viscous traction
// -2 s_lr (tau n) Re, its magnitude and the wall heat flux
// -s_lr (mu cp / Pr) (n . grad T) Re.  cpk is the cp of that heat flux
// (the reference reads a local there that it never sets; the caller passes
// the gas model's 1/((gama-1) moo^2), sub_tgh.f90:868).  out: six arrays
// of npts, cp, fnx, fny, fnz, cf, qw.
__global__ void k_wall_plot(DRegion g, DBlock b, DFields F, DParams P, double cpk, double *out, int npts)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    long q = gidx(b, i, j, k);
    out[t] = 2.0 * (F.p[q] - P.poo);
    if (P.nvis != 1) return;
    double nx, ny, nz;
    getnxyz(F, q, g.l3, nx, ny, nz);
    const double *uvwt[4] = {F.u, F.v, F.w, F.t};
    double dk[4], de[4], dc[4];
    for (int m = 0; m < 4; m++) {
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
    // Sub1.f90:181-198 stress (tyx = txy, tzx = txz, tzy = tyz)
    const double c23 = 0.66666666667;
    double vis = F.visl[q];
    double txx = vis * c23 * (2.0 * dudx - dvdy - dwdz);
    double tyy = vis * c23 * (2.0 * dvdy - dwdz - dudx);
    double tzz = vis * c23 * (2.0 * dwdz - dudx - dvdy);
    double txy = vis * (dudy + dvdx), txz = vis * (dudz + dwdx), tyz = vis * (dvdz + dwdy);
    double pn = 2.0 * g.s_lr;
    double fnx = -pn * (nx * txx + ny * txy + nz * txz) * P.re;
    double fny = -pn * (nx * txy + ny * tyy + nz * tyz) * P.re;
    double fnz = -pn * (nx * txz + ny * tyz + nz * tzz) * P.re;
    double dtdx = (kx * dk[3] + ex * de[3] + cx * dc[3]) / vj;
    double dtdy = (ky * dk[3] + ey * de[3] + cy * dc[3]) / vj;
    double dtdz = (kz * dk[3] + ez * de[3] + cz * dc[3]) / vj;
    double kcp = vis * cpk / P.prl;
    double qw = -(double)g.s_lr * kcp * (nx * dtdx + ny * dtdy + nz * dtdz) * P.re;
    out[npts + t] = fnx;
    out[2 * npts + t] = fny;
    out[3 * npts + t] = fnz;
    out[4 * npts + t] = sqrt(fnx * fnx + fny * fny + fnz * fnz);
    out[5 * npts + t] = qw;
}

// region-local corner offsets of quad (a, c, e), corners 1..4 of force_dif
// (PostProcess.f90:1577-1613)
__device__ inline void quad_corners(const DRegion &g, int a, int c, int e, int o[4][3])
{
    int id = 1 - abs(g.l3[0]), jd = 1 - abs(g.l3[1]), kd = 1 - abs(g.l3[2]);
    for (int n = 0; n < 4; n++) { o[n][0] = a; o[n][1] = c; o[n][2] = e; }
    if (id == 0) { o[1][1] += 1; o[2][1] += 1; o[2][2] += 1; o[3][2] += 1; }
    if (jd == 0) { o[1][0] += 1; o[2][0] += 1; o[2][2] += 1; o[3][2] += 1; }
    if (kd == 0) { o[1][0] += 1; o[2][0] += 1; o[2][1] += 1; o[3][1] += 1; }
}

template <class T>
__device__ inline T lenth(const T *a, const T *b)
{
    T dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Sub1.f90:865-878 halen (Heron; a negative radicand gives zero area)
template <class T>
__device__ inline T halen(T a, T b, T c)
{
    T p = 0.5 * (a + b + c);
    T S = p * (p - a) * (p - b) * (p - c);
    return S < 0 ? T(0.0) : sqrt(S);
}

// getforce (PostProcess.f90:1684-1796): four triangles around the quad's centre,
// traction at each centroid times the Heron area; d[0..2] force, d[3..5] moment
template <class T>
__device__ inline void quad_force(const T (&P)[4][3], const T (&Xc)[4][3], const double rf[3], T d[6])
{
    T X[5][3], P5[3];
    for (int n = 0; n < 4; n++)
        for (int m = 0; m < 3; m++) X[n][m] = Xc[n][m];
    for (int m = 0; m < 3; m++) {
        P5[m] = 0.25 * (P[0][m] + P[1][m] + P[2][m] + P[3][m]);
        X[4][m] = 0.25 * (X[0][m] + X[1][m] + X[2][m] + X[3][m]);
    }
    T r12 = lenth(X[0], X[1]), r23 = lenth(X[1], X[2]), r34 = lenth(X[2], X[3]), r41 = lenth(X[3], X[0]);
    T r15 = lenth(X[0], X[4]), r25 = lenth(X[1], X[4]), r35 = lenth(X[2], X[4]), r45 = lenth(X[3], X[4]);
    T s[4] = {halen(r12, r15, r25), halen(r23, r25, r35), halen(r34, r35, r45), halen(r41, r45, r15)};
    for (int tr = 0; tr < 4; tr++) {
        int ia = tr, ib = (tr + 1) & 3;
        T pc[3], xc[3];
        for (int m = 0; m < 3; m++) {
            pc[m] = (P[ia][m] + P[ib][m] + P5[m]) / 3.0;
            xc[m] = (X[ia][m] + X[ib][m] + X[4][m]) / 3.0;
        }
        T fx = pc[0] * s[tr], fy = pc[1] * s[tr], fz = pc[2] * s[tr];
        d[0] = d[0] + fx; d[1] = d[1] + fy; d[2] = d[2] + fz;
        d[3] = d[3] + ((xc[1] - rf[1]) * fz - (xc[2] - rf[2]) * fy);
        d[4] = d[4] + ((xc[2] - rf[2]) * fx - (xc[0] - rf[0]) * fz);
        d[5] = d[5] + ((xc[0] - rf[0]) * fy - (xc[1] - rf[1]) * fx);
    }
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
        int o[4][3];
        quad_corners(g, a, c, e, o);
        double P[4][3], X[4][3];
        for (int n = 0; n < 4; n++) {
            int pt = o[n][0] + bi * (o[n][1] + bj * o[n][2]);
            for (int m = 0; m < 3; m++) {
                P[n][m] = pn3[(long)m * npts + pt];
                X[n][m] = xyz[(long)m * npts + pt];
            }
        }
        const double rf[3] = {ref.x, ref.y, ref.z};
        quad_force(P, X, rf, d);
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


// ------------------------------------------- adjoint of the pressure force
// J = sum_c omega[c] * (region sums)[c] of the pressure traction only
// (PostProcess.f90:1560-1564); its gradient w.r.t. the traction and the
// coordinates of each region point gathers, over the up to four quads that
// hold the point, the quad's derivative along each of the point's six inputs.
struct Omega6 {
    double w[6];
};

__global__ void k_force_quad_vjp(DRegion g, const double *pn3, const double *xyz, int npts, double3 ref, Omega6 om,
                                 double *lpn3, double *lxyz)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    int bi = g.s_hi[0] - g.s_lo[0] + 1, bj = g.s_hi[1] - g.s_lo[1] + 1, bk = g.s_hi[2] - g.s_lo[2] + 1;
    int tg[3] = {1 - abs(g.l3[0]), 1 - abs(g.l3[1]), 1 - abs(g.l3[2])};   // 1 along the tangential dims
    int qn[3] = {bi - tg[0], bj - tg[1], bk - tg[2]};
    int pt[3] = {t % bi, (t / bi) % bj, t / (bi * bj)};
    const double rf[3] = {ref.x, ref.y, ref.z};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int cand = 0; cand < 8; cand++) {
        int q[3];
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            int off = (cand >> d) & 1;
            if (off && !tg[d]) { ok = false; break; }
            q[d] = pt[d] - off;
            if (q[d] < 0 || q[d] >= qn[d]) { ok = false; break; }
        }
        if (!ok) continue;
        int o[4][3];
        quad_corners(g, q[0], q[1], q[2], o);
        int hit = -1;
        for (int n = 0; n < 4; n++)
            if (o[n][0] == pt[0] && o[n][1] == pt[1] && o[n][2] == pt[2]) hit = n;
        if (hit < 0) continue;
        double P[4][3], X[4][3];
        for (int n = 0; n < 4; n++) {
            int ptn = o[n][0] + bi * (o[n][1] + bj * o[n][2]);
            for (int m = 0; m < 3; m++) {
                P[n][m] = pn3[(long)m * npts + ptn];
                X[n][m] = xyz[(long)m * npts + ptn];
            }
        }
        for (int sd = 0; sd < 6; sd++) {
            D1 PD[4][3], XD[4][3], d6[6];
            for (int n = 0; n < 4; n++)
                for (int m = 0; m < 3; m++) {
                    PD[n][m] = D1(P[n][m], (n == hit && sd == m) ? 1.0 : 0.0);
                    XD[n][m] = D1(X[n][m], (n == hit && sd == 3 + m) ? 1.0 : 0.0);
                }
            for (int c = 0; c < 6; c++) d6[c] = D1(0.0);
            quad_force(PD, XD, rf, d6);
            double v = 0.0;
            for (int c = 0; c < 6; c++) v = v + om.w[c] * d6[c].d;
            acc[sd] = acc[sd] + v;
        }
    }
    for (int m = 0; m < 3; m++) {
        lpn3[(long)m * npts + t] = acc[m];
        lxyz[(long)m * npts + t] = acc[3 + m];
    }
}

// pressure traction of a face node: s_lr * n * 2 (p - poo), n the unit normal of
// the summed metric triples of the face directions (Sub1.f90:750-780)
template <class T>
__device__ inline void pn_pressure(T p, T x, T y, T z, int s_lr, double poo, T *out)
{
    T s = sqrt(x * x + y * y + z * z);
    s = s < SML_SSS ? T(SML_SSS) : s;
    T nx = x / s, ny = y / s, nz = z / s;
    T pn = 2.0 * (p - poo);
    out[0] = (double)s_lr * nx * pn;
    out[1] = (double)s_lr * ny * pn;
    out[2] = (double)s_lr * nz * pn;
}

// gradient of the traction w.r.t. p and the metrics of each region point, contracted with lpn3
__global__ void k_force_pn_vjp(DRegion g, DBlock b, DFields F, DParams P, const double *lpn3, int npts, double *gp,
                               double *gm)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    long q = gidx(b, i, j, k);
    const double *tri[3][3] = {{F.kcx, F.kcy, F.kcz}, {F.etx, F.ety, F.etz}, {F.ctx, F.cty, F.ctz}};
    double x = 0, y = 0, z = 0;
    for (int d = 0; d < 3; d++)
        if (g.l3[d]) { x = x + tri[d][0][q]; y = y + tri[d][1][q]; z = z + tri[d][2][q]; }
    double lam[3] = {lpn3[t], lpn3[npts + t], lpn3[2L * npts + t]};
    double gsd[4];
    for (int sd = 0; sd < 4; sd++) {
        D1 out[3];
        pn_pressure(D1(F.p[q], sd == 0 ? 1.0 : 0.0), D1(x, sd == 1 ? 1.0 : 0.0), D1(y, sd == 2 ? 1.0 : 0.0),
                    D1(z, sd == 3 ? 1.0 : 0.0), g.s_lr, P.poo, out);
        gsd[sd] = lam[0] * out[0].d + lam[1] * out[1].d + lam[2] * out[2].d;
    }
    gp[4 * F.total + q] = gp[4 * F.total + q] + gsd[0];
    for (int d = 0; d < 3; d++)
        if (g.l3[d])
            for (int c = 0; c < 3; c++) gm[(3 * d + c) * F.total + q] = gm[(3 * d + c) * F.total + q] + gsd[1 + c];
}

__device__ inline long xidx_of(const DBlock &b, int i, int j, int k)
{
    return (long)(i - 1) + (long)b.ni * ((long)(j - 1) + (long)b.nj * (long)(k - 1));
}

// face-node coordinates of a region from X (PLOT3D order): back = 0 gathers
// xyz <- X, back = 1 adds xyz into X (the transpose, for gradients)
__global__ void k_wall_xyz(DRegion g, DBlock b, double *X, double *xyz, int npts, int back)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    long nn = (long)b.ni * b.nj * b.nk, xi = xidx_of(b, i, j, k);
    for (int m = 0; m < 3; m++) {
        if (back) X[m * nn + xi] = X[m * nn + xi] + xyz[(long)m * npts + t];
        else xyz[(long)m * npts + t] = X[m * nn + xi];
    }
}

// force (PostProcess.f90:1429-1477): the regions' sums added in (block,
// region) order, scaled to coefficients, cl, cd and xcp.  out = cfx, cfy, cfz,
// cmx, cmy, cmz, cd, cl, xcp; with hist != nullptr the 9 values are also
// appended to the ring hist[cap][9] at *count % cap.
__global__ void k_force_coef(const double *out6, int nwalls, ForceCoef c, double *out, double *hist,
                             unsigned long long *count, int cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double f[6] = {0, 0, 0, 0, 0, 0};
    for (int w = 0; w < nwalls; w++)
        for (int m = 0; m < 6; m++) f[m] = f[m] + out6[6 * w + m];
    for (int m = 0; m < 3; m++) f[m] = c.wf * f[m] * c.scf;
    for (int m = 3; m < 6; m++) f[m] = c.wf * f[m] * c.vcm;
    double e[9];
    for (int m = 0; m < 6; m++) e[m] = f[m];
    e[7] = f[1] * c.cosa - f[0] * c.sina;
    e[6] = c.cosb * (f[1] * c.sina + f[0] * c.cosa) + c.sinb * f[2];
    e[8] = (f[1] >= 0 ? 1.0 : -1.0) * f[5] / (fabs(f[1]) + SMALL);
    for (int m = 0; m < 9; m++) out[m] = e[m];
    if (hist) {
        const unsigned long long n = *count;
        double *h = hist + 9 * (long)(n % (unsigned long long)cap);
        for (int m = 0; m < 9; m++) h[m] = e[m];
        *count = n + 1;
    }
}

}  // namespace

int launch_force_coef(const double *out6, int nwalls, const ForceCoef &c, double *out, double *hist,
                      unsigned long long *count, int cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_force_coef, dim3(1), dim3(64), 0, s, out6, nwalls, c, out, hist, count, cap);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

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

int launch_wall_plot(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, double cpk, double *out,
                     hipStream_t s)
{
    if (g.npts <= 0) return 0;
    hipLaunchKernelGGL(k_wall_plot, dim3((g.npts + 255) / 256), dim3(256), 0, s, g, b, F, P, cpk, out, g.npts);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_wall_xyz(const DRegion &g, const DBlock &b, double *X, double *xyz, int back, hipStream_t s)
{
    if (g.npts <= 0) return 0;
    hipLaunchKernelGGL(k_wall_xyz, dim3((g.npts + 255) / 256), dim3(256), 0, s, g, b, X, xyz, g.npts, back);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// the pressure part of the force of one region (pn3 as k_force_pn leaves it) and
// the gradient of sum_c omega[c] * sums[c] into gp (p), gm (metrics), lxyz (coordinates)
int launch_force_vjp(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, const double *pn3,
                     const double *xyz, const double ref[3], const double omega[6], double *lpn3, double *lxyz,
                     double *gp, double *gm, hipStream_t s)
{
    int npts = g.npts;
    if (npts <= 0) return 0;
    Omega6 om;
    for (int c = 0; c < 6; c++) om.w[c] = omega[c];
    dim3 grid((npts + 127) / 128), blk(128);
    hipLaunchKernelGGL(k_force_quad_vjp, grid, blk, 0, s, g, pn3, xyz, npts, make_double3(ref[0], ref[1], ref[2]), om,
                       lpn3, lxyz);
    hipLaunchKernelGGL(k_force_pn_vjp, grid, blk, 0, s, g, b, F, P, (const double *)lpn3, npts, gp, gm);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
