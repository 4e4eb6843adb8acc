// wcns_sa.hip -- the turbulence models as the reference runs them inside
// r_h_s_vis (RHS.f90:22 -> turblence.f90:15-103 turbulence ->
// differential_model_turbulence), ncmpcor 0, ndes 0, method 1: the
// Spalart-Allmaras one-equation model (nameturb 'SA', nlamtur 1) and the SST
// k-omega model (nameturb 'SST', nlamtur 2), and their set-up (wall distance,
// uniform initial field).
//
//   k_wall_dist   parallel.f90:4313-4468 calc_wall_dists_parallel: squared
//                 distance to every no-slip wall node, minimum, sqrt + small;
//                 wall points staged through LDS in chunks.
//   k_sa_init     turblence.f90:2701-2734 init_SA / :3091-3117 init_SST.
//   k_sa_node     per node 1..n: the central-difference gradients of u, v, w
//                 and nu~ (dfidxyz_center, turblence.f90:753-849), the
//                 equivalent viscosity (viscosity_equivalent_SA_r,
//                 turblence.f90:2785-2808) and the contravariant velocities
//                 that the scalar LU-SGS coefficients are made of.
//   k_sa_rhs      per interior node 2..n-1, fused in the reference's order of
//                 accumulation: source (cal_rhs_source_SA, turblence.f90:
//                 2506-2699), diffusion along i, j, k (cal_rhs_viscous_
//                 turbulence_dif_r, :1022-1185), NND convection along i, j, k
//                 (cal_rhs_inviscous_turbulence + NND_turbulence +
//                 flux_turbulence, :291-459) and the diagonal (source part,
//                 then cal_spec_turbulence_r, :218-259).  Face quantities that
//                 two nodes share are formed by both, with the same operations.
//   k_sa_sweep    lusgs_turbulence (turblence.f90:477-583): the scalar
//                 forward and backward sweeps, pipelined over the same
//                 (j,k)-column tiles, work list and exchange slots as the
//                 five-variable sweeps (one word per boundary node).
//   k_sa_update   update_turbulence (:585-623), cal_vist_SA (:2735-2760) and
//                 max_control_turbulence (:625-660), node by node.
//   SST: k_sa_node also forms d(omega)/dx, the blending function F1 and the
//   equivalent viscosities of the first loop of cal_rhs_source_SST
//   (:2810-2923); k_sst_rhs is the per-node fusion of the rest of
//   cal_rhs_source_SST (:2925-3086), cal_rhs_viscous_turbulence_dif (:851-1019),
//   cal_spec_turbulence (:175-215) and the NND convection of both variables;
//   the scalar sweeps run once per variable; k_sa_update applies
//   cal_vist_SST (:3119-3164).
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// cal_rhs_source_SA model constants (turblence.f90:2509-2513)
constexpr double SA_CB1 = 0.1355, SA_CB2 = 0.622, SA_CV1 = 7.1, SA_EPS = 1.0e-20;
constexpr double SA_CGA = 2. / 3.0, SA_KI = 0.419, SA_CW2 = 0.3, SA_CW3 = 2.0;

__device__ inline double amin1(double a, double b) { return a < b ? a : b; }
__device__ inline double amax1(double a, double b) { return a > b ? a : b; }

// Convect.f90:1281-1286 minmod
__device__ inline double minmod(double x, double y)
{
    return 0.5 * (copysign(1.0, x) + copysign(1.0, y)) * amin1(fabs(x), fabs(y));
}

__device__ inline void node_of(const DBlock &b, long t, int lo, int ext_i, int ext_j, int &i, int &j, int &k)
{
    i = (int)(t % ext_i) + lo;
    j = (int)((t / ext_i) % ext_j) + lo;
    k = (int)(t / ((long)ext_i * ext_j)) + lo;
}

// ------------------------------------------------------------ wall distance
constexpr int WD_THREADS = 256;

__global__ void __launch_bounds__(WD_THREADS) k_wall_dist(DBlock b, const double *x, const double *y,
                                                          const double *z, const double *wall, int nw,
                                                          double *dist)
{
    __shared__ double sx[WD_THREADS], sy[WD_THREADS], sz[WD_THREADS];
    const long n = (long)b.ni * b.nj * b.nk;
    const long t = (long)blockIdx.x * WD_THREADS + threadIdx.x;
    const bool act = t < n;
    double xc = 0.0, yc = 0.0, zc = 0.0;
    if (act) { xc = x[t]; yc = y[t]; zc = z[t]; }
    double best = 1.0e30;
    for (int w0 = 0; w0 < nw; w0 += WD_THREADS) {
        const int cnt = nw - w0 < WD_THREADS ? nw - w0 : WD_THREADS;
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            sx[threadIdx.x] = wall[w0 + threadIdx.x];
            sy[threadIdx.x] = wall[(long)nw + w0 + threadIdx.x];
            sz[threadIdx.x] = wall[2L * nw + w0 + threadIdx.x];
        }
        __syncthreads();
        for (int m = 0; m < cnt; m++) {
            const double dx = xc - sx[m], dy = yc - sy[m], dz = zc - sz[m];
            const double d = dx * dx + dy * dy + dz * dz;
            best = d < best ? d : best;
        }
    }
    if (act) {
        int i, j, k;
        node_of(b, t, 1, b.ni, b.nj, i, j, k);
        dist[gidx(b, i, j, k)] = sqrt(best) + SMALL;
    }
}

// ------------------------------------------------------------------ init
__global__ void k_sa_init(DBlock b, DFields F, double q0, double w0, double vt)
{
    const long n = (long)(b.ni + 3) * (b.nj + 3) * (b.nk + 3);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    node_of(b, t, -1, b.ni + 3, b.nj + 3, i, j, k);
    const long g = gidx(b, i, j, k);
    F.qke[g] = q0;
    if (F.qkw) F.qkw[g] = w0;
    F.vist[g] = vt;
}

// cal_rhs_source_SST / cal_vist_SST constants (turblence.f90:2813-2818, 3122)
constexpr double SST_CGAK0 = 0.85, SST_CGAK1 = 1.0, SST_CGAW0 = 0.5, SST_CGAW1 = 0.856;
constexpr double SST_BETA0 = 0.075, SST_BETA1 = 0.0828, SST_CDMIN = 1.0e-20, SST_KI = 0.41;

// ------------------------------------------------------------- node pass
// gradient array m of variable v (u, v, w, nu~) at component c: grad[(3v+c)*total + g]
__global__ void k_sa_node(BlockGrid G, DFields F, SaWork W, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i, j, k;
    node_of(b, t, 1, b.ni, b.nj, i, j, k);
    const long g = gidx(b, i, j, k);
    const long gi0 = gidx(b, i - 1, j, k), gi1 = gidx(b, i + 1, j, k);
    const long gj0 = gidx(b, i, j - 1, k), gj1 = gidx(b, i, j + 1, k);
    const long gk0 = gidx(b, i, j, k - 1), gk1 = gidx(b, i, j, k + 1);
    const double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
    const double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
    const double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
    const double vol2 = 0.5 / F.vol[g];
    const double *fis[4] = {F.u, F.v, F.w, F.qke};
    const long T = F.total;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double *fi = fis[v];
        const double dkc = fi[gi1] - fi[gi0];
        const double det = fi[gj1] - fi[gj0];
        const double dct = fi[gk1] - fi[gk0];
        const double ddx = dkc * kx + det * ex + dct * cx;
        const double ddy = dkc * ky + det * ey + dct * cy;
        const double ddz = dkc * kz + det * ez + dct * cz;
        W.grad[(3 * v) * T + g] = ddx * vol2;
        W.grad[(3 * v + 1) * T + g] = ddy * vol2;
        W.grad[(3 * v + 2) * T + g] = ddz * vol2;
    }
    const double r = F.r[g], um = F.u[g], vm = F.v[g], wm = F.w[g];
    if (P.turb == 2) {
        // d(omega)/dx, then the first loop of cal_rhs_source_SST (nodes 1..n)
        const double *fi = F.qkw;
        const double dkc = fi[gi1] - fi[gi0], det = fi[gj1] - fi[gj0], dct = fi[gk1] - fi[gk0];
        const double gw[3] = {(dkc * kx + det * ex + dct * cx) * vol2, (dkc * ky + det * ey + dct * cy) * vol2,
                              (dkc * kz + det * ez + dct * cz) * vol2};
#pragma unroll
        for (int c = 0; c < 3; c++) W.gradw[c * T + g] = gw[c];
        const double re500 = P.re * 500.0, cgaw24 = SST_CGAW1 * 4.0, cgaw22 = SST_CGAW1 * 2.0;
        const double obetx = 1.0 / 0.09;
        const double qk = F.qke[g], qw = F.qkw[g], vsl = F.visl[g], vst = F.vist[g];
        const double rok = r * qk, row = r * qw;
        const double dn = F.dist[g], dn2 = dn * dn;
        const double dkxdw = W.grad[9 * T + g] * gw[0] + W.grad[10 * T + g] * gw[1] + W.grad[11 * T + g] * gw[2];
        const double rowdkdw = r / qw * dkxdw * cgaw22;
        const double CDkw = amax1(SST_CDMIN, rowdkdw);
        const double sqrtkwdn = obetx * sqrt(qk) / qw / dn;
        const double vlrodn2w = re500 * vsl / row / dn2;
        const double rokdn2 = cgaw24 * rok / CDkw / dn2;
        const double arg2 = amax1(sqrtkwdn, vlrodn2w);
        const double arg1 = amin1(arg2, rokdn2);
        const double f1 = arg1 >= 2.0 ? 1.0 : tanh(pow(arg1, 4.0));
        const double f2 = 1.0 - f1;
        W.f1[g] = f1;
        W.viseq[0][g] = vsl + (f1 * SST_CGAK0 + f2 * SST_CGAK1) * vst;
        W.viseq[1][g] = vsl + (f1 * SST_CGAW0 + f2 * SST_CGAW1) * vst;
    } else {
        W.viseq[0][g] = 1.5 * (F.visl[g] / r + F.qke[g]);
    }
    // mxdq_std_turbulence's l1 = g . (u, v, w) with g the metric of each direction
    W.un[0][g] = kx * um + ky * vm + kz * wm;
    W.un[1][g] = ex * um + ey * vm + ez * wm;
    W.un[2][g] = cx * um + cy * vm + cz * wm;
}

// -------------------------------------------------------------- RHS pass
// flux_turbulence (turblence.f90:437-459): prim = (r, u, v, w, nu~)
__device__ inline double flux_turb(const double *q, double nx, double ny, double nz, double npn)
{
    double l1 = nx * q[1] + ny * q[2] + nz * q[3];
    l1 = 0.5 * (l1 + npn * sqrt(l1 * l1));
    return l1 * q[4] * q[0];
}

// node at offset d from (i,j,k) along direction DIR
template <int DIR>
__device__ inline long nbr(const DBlock &b, int i, int j, int k, int d)
{
    return gidx(b, i + (DIR == 0 ? d : 0), j + (DIR == 1 ? d : 0), k + (DIR == 2 ? d : 0));
}

// diffusion flux difference along DIR of the variable whose gradient arrays are
// gx, gy, gz and whose equivalent viscosity is viseq
template <int DIR>
__device__ inline double sa_diffusion(const DBlock &b, const DFields &F, const double *gx, const double *gy,
                                      const double *gz, const double *viseq, int i, int j, int k)
{
    const double *mx = DIR == 0 ? F.kcx : (DIR == 1 ? F.etx : F.ctx);
    const double *my = DIR == 0 ? F.kcy : (DIR == 1 ? F.ety : F.cty);
    const double *mz = DIR == 0 ? F.kcz : (DIR == 1 ? F.etz : F.ctz);
    double fv[2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        // face between nodes l-1 and l, l = node + f
        const long qa = nbr<DIR>(b, i, j, k, f), qb = nbr<DIR>(b, i, j, k, f - 1);
        const double nx = 0.5 * (mx[qa] + mx[qb]), ny = 0.5 * (my[qa] + my[qb]), nz = 0.5 * (mz[qa] + mz[qb]);
        const double dmudx = gx[qa] + gx[qb], dmudy = gy[qa] + gy[qb], dmudz = gz[qa] + gz[qb];
        const double vis = viseq[qa] + viseq[qb];
        fv[f] = (dmudx * nx + dmudy * ny + dmudz * nz) * vis * 0.25;
    }
    return fv[1] - fv[0];
}

template <int DIR>
__device__ inline double sa_nnd(const DBlock &b, const DFields &F, const DParams &P, const double *qv, int i, int j,
                                int k)
{
    const double *mx = DIR == 0 ? F.kcx : (DIR == 1 ? F.etx : F.ctx);
    const double *my = DIR == 0 ? F.kcy : (DIR == 1 ? F.ety : F.cty);
    const double *mz = DIR == 0 ? F.kcz : (DIR == 1 ? F.etz : F.ctz);
    const double c1 = 0.25 * (1.0 + P.xk), c2 = 0.25 * (1.0 - P.xk), xb = P.xb;
    // Q at offsets -2..2, DQ(p) = Q(p) - Q(p-1) at -1..2, CL/CR at -1..1
    double q[5][5];
#pragma unroll
    for (int o = 0; o < 5; o++) {
        const long g = nbr<DIR>(b, i, j, k, o - 2);
        q[o][0] = F.r[g]; q[o][1] = F.u[g]; q[o][2] = F.v[g]; q[o][3] = F.w[g]; q[o][4] = qv[g];
    }
    double cl[3][5], cr[3][5];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int m = 0; m < 5; m++) {
            const double dq0 = q[p + 1][m] - q[p][m];       // DQ at offset p-1
            const double dq1 = q[p + 2][m] - q[p + 1][m];   // DQ at offset p
            cl[p][m] = minmod(xb * dq1, dq0);
            cr[p][m] = minmod(dq1, xb * dq0);
        }
    double fc[2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        // face between offsets f-1 and f
        const long ga = nbr<DIR>(b, i, j, k, f - 1), gb = nbr<DIR>(b, i, j, k, f);
        const double nx = 0.50 * (mx[ga] + mx[gb]);
        const double ny = 0.50 * (my[ga] + my[gb]);
        const double nz = 0.50 * (mz[ga] + mz[gb]);
        double ql[5], qr[5];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            ql[m] = q[f + 1][m] + c1 * cr[f][m] + c2 * cl[f][m];
            qr[m] = q[f + 2][m] - c1 * cl[f + 1][m] - c2 * cr[f + 1][m];
        }
        fc[f] = flux_turb(ql, nx, ny, nz, 1.0) + flux_turb(qr, nx, ny, nz, -1.0);
    }
    return fc[1] - fc[0];
}

__global__ void __launch_bounds__(128) k_sa_rhs(BlockGrid G, DFields F, SaWork W, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int mi = b.ni - 2, mj = b.nj - 2;
    const long n = (long)mi * mj * (b.nk - 2);
    if (t >= n) return;
    int i, j, k;
    node_of(b, t, 2, mi, mj, i, j, k);
    const long g = gidx(b, i, j, k);
    const long T = F.total;
    const double re = P.re, reynolds = P.reynolds;
    const double cw1 = SA_CB1 / SA_KI / SA_KI + (1.0 + SA_CB2) / SA_CGA;
    const double cv13 = SA_CV1 * SA_CV1 * SA_CV1, ki2 = SA_KI * SA_KI;
    const double cw36 = pow(SA_CW3, 6.0), onesix = 1.0 / 6.0;
    const double ft2 = 0.0;

    // cal_rhs_source_SA
    const double ro = F.r[g], volume = F.vol[g], vsl = F.visl[g], vtbar = F.qke[g];
    const double romu = ro * vtbar;
    const double dn = amax1(F.dist[g], re);
    const double dn2 = dn * dn;
    const double reki2dn2 = reynolds * ki2 * dn2;
    const double kafan = romu / vsl;
    const double kf3 = kafan * kafan * kafan;
    const double fv1 = kf3 / (kf3 + cv13);
    const double fv2 = 1.0 - kafan / (1.0 + kafan * fv1);
    const double dudy = W.grad[1 * T + g], dudz = W.grad[2 * T + g];
    const double dvdx = W.grad[3 * T + g], dvdz = W.grad[5 * T + g];
    const double dwdx = W.grad[6 * T + g], dwdy = W.grad[7 * T + g];
    const double a1 = dudy - dvdx, a2 = dvdz - dwdy, a3 = dwdx - dudz;
    const double omg = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    double sbar = omg + vtbar / reki2dn2 * fv2;
    sbar = amax1(sbar, SA_EPS);
    double rbar = vtbar / (sbar * reki2dn2);
    rbar = amin1(rbar, 10.0);
    const double rbar5 = pow(rbar, 5.0);
    const double gbar = rbar + SA_CW2 * rbar * (rbar5 - 1.0);
    const double gbar6 = pow(gbar, 6.0);
    const double fw = gbar * pow((1.0 + cw36) / (gbar6 + cw36), onesix);
    const double cpmu = SA_CB1 * (1.0 - ft2) * sbar;
    const double cdmu = (cw1 * fw - SA_CB1 * ft2 / ki2) * vtbar / dn2 / reynolds;
    const double Sp = cpmu * vtbar, Sd = cdmu * vtbar;
    const double dmudx = W.grad[9 * T + g], dmudy = W.grad[10 * T + g], dmudz = W.grad[11 * T + g];
    const double Dbar = SA_CB2 / SA_CGA * (dmudx * dmudx + dmudy * dmudy + dmudz * dmudz) * re;
    double dq = 0.0 - ro * (Sp - Sd + Dbar) * volume;
    double sp = 0.0 - amin1(cpmu - 2.0 * cdmu, 0.0) * volume;

    // cal_rhs_viscous_turbulence_dif_r
    {
        const double *gx = W.grad + 9 * T, *gy = W.grad + 10 * T, *gz = W.grad + 11 * T;
        dq = dq - re * ro * sa_diffusion<0>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
        dq = dq - re * ro * sa_diffusion<1>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
        dq = dq - re * ro * sa_diffusion<2>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
    }

    // cal_spec_turbulence_r
    {
        const double um = F.u[g], vm = F.v[g], wm = F.w[g];
        const double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
        const double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
        const double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
        const double unx = um * kx + vm * ky + wm * kz;
        const double uny = um * ex + vm * ey + wm * ez;
        const double unz = um * cx + vm * cy + wm * cz;
        sp = sp + fabs(unx) + fabs(uny) + fabs(unz) + 1.0 / (F.dtdt[g] * P.timedt_turb);
        const double nx = kx * kx + ky * ky + kz * kz;
        const double ny = ex * ex + ey * ey + ez * ez;
        const double nz = cx * cx + cy * cy + cz * cz;
        const double vis = W.viseq[0][g] / volume;
        sp = sp + 2.0 * re * vis * (nx + ny + nz);
    }

    // cal_rhs_inviscous_turbulence
    dq = dq + sa_nnd<0>(b, F, P, F.qke, i, j, k);
    dq = dq + sa_nnd<1>(b, F, P, F.qke, i, j, k);
    dq = dq + sa_nnd<2>(b, F, P, F.qke, i, j, k);

    W.dqke[0][g] = dq;
    W.spec[0][g] = sp;
}

// SST, per interior node 2..n-1: both variables' right-hand sides and diagonals
__global__ void __launch_bounds__(128) k_sst_rhs(BlockGrid G, DFields F, SaWork W, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int mi = b.ni - 2, mj = b.nj - 2;
    const long n = (long)mi * mj * (b.nk - 2);
    if (t >= n) return;
    int i, j, k;
    node_of(b, t, 2, mi, mj, i, j, k);
    const long g = gidx(b, i, j, k);
    const long T = F.total;
    const double re = P.re, reynolds = P.reynolds;
    const double ki2 = SST_KI * SST_KI, twothd = 0.2 / 0.3;
    const double betax = 0.09, obetx = 1.0 / betax, ki2osqrtbetx = ki2 / sqrt(betax);
    const double cgaw22 = SST_CGAW1 * 2.0;

    // cal_rhs_source_SST, second loop
    const double ro = F.r[g], qk = F.qke[g], qw = F.qkw[g], volume = F.vol[g], vst = F.vist[g];
    const double rok = ro * qk, row = ro * qw;
    const double f1 = W.f1[g], f2 = 1.0 - f1;
    const double beita = f1 * SST_BETA0 + f2 * SST_BETA1;
    const double cga1 = f1 * SST_CGAW0 + f2 * SST_CGAW1;
    const double ga = beita * obetx - cga1 * ki2osqrtbetx;
    const double ux = W.grad[0 * T + g], uy = W.grad[1 * T + g], uz = W.grad[2 * T + g];
    const double vx = W.grad[3 * T + g], vy = W.grad[4 * T + g], vz = W.grad[5 * T + g];
    const double wx = W.grad[6 * T + g], wy = W.grad[7 * T + g], wz = W.grad[8 * T + g];
    const double div_u = ux + vy + wz;
    const double dkxdw = W.grad[9 * T + g] * W.gradw[0 * T + g] + W.grad[10 * T + g] * W.gradw[1 * T + g] +
                         W.grad[11 * T + g] * W.gradw[2 * T + g];
    const double omg = vst * ((2.0 * (ux * ux + vy * vy + wz * wz + uy * vx + uz * wx + wy * vz) + uy * uy + uz * uz +
                               vx * vx + vz * vz + wx * wx + wy * wy) -
                              twothd * div_u * div_u) * re -
                       twothd * div_u * rok;
    const double cdk = betax * qw, cdw = beita * qw;
    const double Sd0 = cdk * rok, Sd1 = cdw * row;
    const double Sp0 = amin1(omg, 20.0 * Sd0);
    const double Sp1 = ga * ro * omg / vst * reynolds;
    const double dbw = cgaw22 * f2 * dkxdw / qw / qw;
    const double Dbar0 = 0.0, Dbar1 = dbw * row;
    double sp0 = 0.0 + 2.0 * (cdk + twothd * amax1(ux + vy + wz, 0.0)) * volume;
    double sp1 = 0.0 + 2.0 * (2.0 * cdw + fabs(dbw)) * volume;
    double dq0 = 0.0 - (Sp0 - Sd0 + Dbar0) * volume;
    double dq1 = 0.0 - (Sp1 - Sd1 + Dbar1) * volume;

    // cal_rhs_viscous_turbulence_dif
    {
        const double *gx = W.grad + 9 * T, *gy = W.grad + 10 * T, *gz = W.grad + 11 * T;
        const double *hx = W.gradw, *hy = W.gradw + T, *hz = W.gradw + 2 * T;
        dq0 = dq0 - re * sa_diffusion<0>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
        dq1 = dq1 - re * sa_diffusion<0>(b, F, hx, hy, hz, W.viseq[1], i, j, k);
        dq0 = dq0 - re * sa_diffusion<1>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
        dq1 = dq1 - re * sa_diffusion<1>(b, F, hx, hy, hz, W.viseq[1], i, j, k);
        dq0 = dq0 - re * sa_diffusion<2>(b, F, gx, gy, gz, W.viseq[0], i, j, k);
        dq1 = dq1 - re * sa_diffusion<2>(b, F, hx, hy, hz, W.viseq[1], i, j, k);
    }

    // cal_spec_turbulence
    {
        const double um = F.u[g], vm = F.v[g], wm = F.w[g];
        const double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
        const double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
        const double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
        const double unx = um * kx + vm * ky + wm * kz;
        const double uny = um * ex + vm * ey + wm * ez;
        const double unz = um * cx + vm * cy + wm * cz;
        const double dtt = 1.0 / (F.dtdt[g] * P.timedt_turb);
        const double nx = kx * kx + ky * ky + kz * kz;
        const double ny = ex * ex + ey * ey + ez * ez;
        const double nz = cx * cx + cy * cy + cz * cz;
        sp0 = sp0 + fabs(unx) + fabs(uny) + fabs(unz) + dtt;
        sp0 = sp0 + 2.0 * re * (W.viseq[0][g] / volume / ro) * (nx + ny + nz);
        sp1 = sp1 + fabs(unx) + fabs(uny) + fabs(unz) + dtt;
        sp1 = sp1 + 2.0 * re * (W.viseq[1][g] / volume / ro) * (nx + ny + nz);
    }

    // cal_rhs_inviscous_turbulence, per variable
    dq0 = dq0 + sa_nnd<0>(b, F, P, F.qke, i, j, k);
    dq0 = dq0 + sa_nnd<1>(b, F, P, F.qke, i, j, k);
    dq0 = dq0 + sa_nnd<2>(b, F, P, F.qke, i, j, k);
    dq1 = dq1 + sa_nnd<0>(b, F, P, F.qkw, i, j, k);
    dq1 = dq1 + sa_nnd<1>(b, F, P, F.qkw, i, j, k);
    dq1 = dq1 + sa_nnd<2>(b, F, P, F.qkw, i, j, k);

    W.dqke[0][g] = dq0;
    W.dqke[1][g] = dq1;
    W.spec[0][g] = sp0;
    W.spec[1][g] = sp1;
}

// ------------------------------------------------------------ the sweeps
__device__ inline void xch_put1(double *slot, double v)
{
    __hip_atomic_store(xch_ptr(slot), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned long long xch_issue1(double *slot)
{
    return __hip_atomic_load(xch_ptr(slot), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait for a slot read issued earlier (see the five-variable sweeps).
__device__ inline bool xch_finish1(double *slot, unsigned long long raw, double &v, long spin_limit, int *err)
{
    xch_word *p = xch_ptr(slot);
    long spins = 0;
    unsigned long long x = raw;
    while (x == XCH_EMPTY) {
        if (++spins > spin_limit ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            v = 0.0;
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
        x = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    v = __longlong_as_double((long long)x);
    return true;
}

struct SaOp {
    double d, sp, ui, uj, uk;
};

// Scalar LU-SGS, one tile of (j,k) columns per workgroup of one wavefront; lane
// (jl,kl) walks its column, at step t updating node i = t+1-jl-kl (forward).
// Nodes outside 2..n-1 take no part in lusgs_turbulence: their lanes publish 0,
// the value dqke holds there.  Forward: dq = (A+ dq_lower - dq) / spec;
// backward: dq = dq - (A- dq_upper) / spec, A+- = (l1 +- |l1|)/2 of the
// neighbour (mxdq_std_turbulence, turblence.f90:560-583).
// PM: the operands come from (and the unknown goes to) the plane-major copy
// pm = slots SK_A.. of the FP64 plane-major pool, where the lanes of a tile
// read whole rows of a hyperplane; otherwise from the field pools, where every
// lane reads a line of its own.
template <int FWD, bool PM>
__global__ void __launch_bounds__(64) k_sa_sweep(const SweepBlock *blocks, const int2 *work, SaWork W,
                                                 double *dqv, const double *specv, double *pm, long sktot, int *err,
                                                 long spin_limit)
{
    constexpr int off = FWD ? -1 : 1;
    const double npn = FWD ? 1.0 : -1.0;
    const int2 wk = work[blockIdx.x];
    const SweepBlock sb = blocks[wk.x];
    const DBlock b = sb.b;
    const int TJ = sb.TJ, TK = sb.TK, nTJ = sb.nTJ, nTK = sb.nTK, ni = b.ni;
    const int tile = wk.y;
    const int J = tile % nTJ, K = tile / nTJ;
    const int c = threadIdx.x;
    const int jl = c % TJ, kl = c / TJ;
    const int jlo = 1 + J * TJ, klo = 1 + K * TK;
    const int jn = min(TJ, b.nj - jlo + 1), kn = min(TK, b.nk - klo + 1);
    const bool col = kl < kn && jl < jn;
    const int j = jlo + jl, k = klo + kl;
    const int nsteps = ni + jn + kn - 2;
    const int i0 = FWD ? 1 - jl - kl : ni + (jn - 1 - jl) + (kn - 1 - kl);
    auto i_at = [&](int st) { return FWD ? i0 + st : i0 - st; };
    auto live = [&](int st) { int ii = i_at(st); return col && st < nsteps && ii >= 1 && ii <= ni; };
    const bool colin = j >= 2 && j <= b.nj - 1 && k >= 2 && k <= b.nk - 1;
    auto inner = [&](int ii) { return colin && ii >= 2 && ii <= ni - 1; };

    __shared__ double sh[2][64];
    const int jsrc = FWD ? (jl > 0 ? 0 : (J > 0 ? 1 : 2)) : (jl < jn - 1 ? 0 : (J < nTJ - 1 ? 1 : 2));
    const int ksrc = FWD ? (kl > 0 ? 0 : (K > 0 ? 1 : 2)) : (kl < kn - 1 ? 0 : (K < nTK - 1 ? 1 : 2));
    const bool putj = col && (FWD ? (jl == jn - 1 && J < nTJ - 1) : (jl == 0 && J > 0));
    const bool putk = col && (FWD ? (kl == kn - 1 && K < nTK - 1) : (kl == 0 && K > 0));
    const long rowlen = ni;
    double *xj_in = sb.xj + (jsrc == 1 ? ((long)(tile + off) * TK + kl) * rowlen : 0);
    double *xk_in = sb.xk + (ksrc == 1 ? ((long)(tile + off * nTJ) * TJ + jl) * rowlen : 0);
    double *xj_out = sb.xj + ((long)tile * TK + kl) * rowlen;
    double *xk_out = sb.xk + ((long)tile * TJ + jl) * rowlen;

    const SkIndex ix = skew_index(b, sb.soff);
    auto load = [&](int st, SaOp &o) {
        const int i = i_at(st);
        if (!inner(i)) return;
        if constexpr (PM) {
            const long g = ix(i, j, k);
            o.d = pm[g];
            o.sp = pm[sktot + g];
            o.ui = pm[2 * sktot + ix(i + off, j, k)];
            o.uj = pm[3 * sktot + ix(i, j + off, k)];
            o.uk = pm[4 * sktot + ix(i, j, k + off)];
        } else {
            const long g = gidx(b, i, j, k);
            o.d = dqv[g];
            o.sp = specv[g];
            o.ui = W.un[0][gidx(b, i + off, j, k)];
            o.uj = W.un[1][gidx(b, i, j + off, k)];
            o.uk = W.un[2][gidx(b, i, j, k + off)];
        }
    };
    // operands of this step (o0) and the next two (o1, o2), rotated by value so
    // that no register array is indexed at run time
    SaOp o0{}, o1{}, o2{};
    if (live(0)) load(0, o0);
    if (live(1)) load(1, o1);
    double prev = 0.0;
    unsigned long long rj = 0, rk = 0;
    bool bad = false;
    if (live(0) && inner(i_at(0))) {
        const int i = i_at(0);
        if (jsrc == 1) rj = xch_issue1(xj_in + (i - 1));
        if (ksrc == 1) rk = xch_issue1(xk_in + (i - 1));
    }
    for (int step = 0; step < nsteps; step++) {
        const int slot = step & 1;
        if (live(step + 2)) load(step + 2, o2);
        if (live(step)) {
            const int i = i_at(step);
            double out = 0.0;
            if (inner(i)) {
                const SaOp &o = o0;
                double nj_ = 0.0, nk_ = 0.0;
                if (jsrc == 0) nj_ = sh[slot ^ 1][c + off];
                else if (jsrc == 1 && !xch_finish1(xj_in + (i - 1), rj, nj_, bad ? 0 : spin_limit, err)) {
                    if (!bad) atomicExch(err, 1);
                    bad = true;
                }
                if (ksrc == 0) nk_ = sh[slot ^ 1][c + off * TJ];
                else if (ksrc == 1 && !xch_finish1(xk_in + (i - 1), rk, nk_, bad ? 0 : spin_limit, err)) {
                    if (!bad) atomicExch(err, 1);
                    bad = true;
                }
                const double ai = 0.5 * (o.ui + npn * fabs(o.ui));
                const double aj = 0.5 * (o.uj + npn * fabs(o.uj));
                const double ak = 0.5 * (o.uk + npn * fabs(o.uk));
                const double rhs0 = ai * prev + aj * nj_ + ak * nk_;
                out = FWD ? (rhs0 - o.d) / o.sp : o.d - rhs0 / o.sp;
                if constexpr (PM) pm[ix(i, j, k)] = out;
                else dqv[gidx(b, i, j, k)] = out;
            }
            if (putj) xch_put1(xj_out + (i - 1), out);
            if (putk) xch_put1(xk_out + (i - 1), out);
            sh[slot][c] = out;
            prev = out;
        }
        if (live(step + 1) && inner(i_at(step + 1))) {
            const int i = i_at(step + 1);
            if (jsrc == 1) rj = xch_issue1(xj_in + (i - 1));
            if (ksrc == 1) rk = xch_issue1(xk_in + (i - 1));
        }
        o0 = o1;
        o1 = o2;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- update
__global__ void k_sa_update(BlockGrid G, DFields F, SaWork W, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const long n = (long)(b.ni + 1) * (b.nj + 1) * (b.nk + 1);
    if (t >= n) return;
    int i, j, k;
    node_of(b, t, 0, b.ni + 1, b.nj + 1, i, j, k);
    const long g = gidx(b, i, j, k);
    const bool in1 = i >= 1 && j >= 1 && k >= 1;
    const bool in2 = i >= 2 && i <= b.ni - 1 && j >= 2 && j <= b.nj - 1 && k >= 2 && k <= b.nk - 1;
    const double r = F.r[g];
    double q = F.qke[g];
    double vt = F.vist[g];
    if (P.turb == 2) {
        double qw = F.qkw[g];
        if (in2) {
            q = amax1(q + W.dqke[0][g] / r, 1.0e-20 * P.koo);
            qw = amax1(qw + W.dqke[1][g] / r, 1.0e-20 * P.omgoo);
            F.qke[g] = q;
            F.qkw[g] = qw;
        }
        if (in1) {
            // cal_vist_SST, with the velocity gradients of this step's source term
            const long T = F.total;
            const double a1 = 0.31, obetx = 2.0 / 0.09;
            const double a1reynolds = a1 * P.reynolds, b1 = 500.0 * P.re;
            const double vsl = F.visl[g];
            const double dn = F.dist[g], dn2 = dn * dn;
            const double sqrtkwdn = obetx * sqrt(q) / qw / dn;
            const double vlrodn2w = b1 * vsl / r / dn2 / qw;
            const double arg2 = amax1(sqrtkwdn, vlrodn2w);
            const double f2 = arg2 >= 4.0 ? 1.0 : tanh(arg2 * arg2);
            const double a_ = W.grad[1 * T + g] - W.grad[3 * T + g];
            const double b_ = W.grad[5 * T + g] - W.grad[7 * T + g];
            const double c_ = W.grad[6 * T + g] - W.grad[2 * T + g];
            const double omg = sqrt(a_ * a_ + b_ * b_ + c_ * c_);
            vt = a1reynolds * r * q / amax1(a1 * qw, omg * f2);
        }
    } else {
        if (in2) {
            q = amax1(q + W.dqke[0][g] / r, 1.0e-20 * P.muoo);
            F.qke[g] = q;
        }
        if (in1) {
            const double cv13 = SA_CV1 * SA_CV1 * SA_CV1;
            const double romu = r * q;
            const double kafan = romu / F.visl[g];
            const double kf3 = kafan * kafan * kafan;
            const double fv1 = kf3 / (kf3 + cv13);
            vt = romu * fv1;
        }
    }
    if (P.kmaxlim > 1.0e-15) vt = amin1(vt, P.kmaxlim * F.visl[g]);
    F.vist[g] = vt;
}

}  // namespace

int launch_wall_dist(const DBlock &b, const double *x, const double *y, const double *z, const double *wall, int nw,
                     double *dist, hipStream_t s)
{
    long n = (long)b.ni * b.nj * b.nk;
    hipLaunchKernelGGL(k_wall_dist, grid1(n, WD_THREADS), dim3(WD_THREADS), 0, s, b, x, y, z, wall, nw, dist);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_sa_init(const DBlock &b, const DFields &F, double q0, double w0, double vt, hipStream_t s)
{
    long n = (long)(b.ni + 3) * (b.nj + 3) * (b.nk + 3);
    hipLaunchKernelGGL(k_sa_init, grid1(n, 256), dim3(256), 0, s, b, F, q0, w0, vt);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_sa_rhs(const BlockGrid &nodes, const BlockGrid &inner, const DFields &F, const SaWork &W,
                  const DParams &P, hipStream_t s)
{
    if (nodes.nwg) hipLaunchKernelGGL(k_sa_node, dim3(nodes.nwg), dim3(256), 0, s, nodes, F, W, P);
    if (inner.nwg) {
        if (P.turb == 2) hipLaunchKernelGGL(k_sst_rhs, dim3(inner.nwg), dim3(128), 0, s, inner, F, W, P);
        else hipLaunchKernelGGL(k_sa_rhs, dim3(inner.nwg), dim3(128), 0, s, inner, F, W, P);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_sa_sweep(int fwd, int var, const SweepPlan &S, const SaWork &W, double *pm, long sktot, hipStream_t s)
{
    if (var < 0 || var > 1 || !W.dqke[var] || !W.spec[var]) return -2;
    if (S.lanes > 64 || S.lanes < 1 || S.nwork < 1) return -2;
    if (launch_fill_xch(S, s)) return -3;     // every exchange word back to the empty pattern
    const int2 *work = fwd ? S.work_fwd : S.work_bwd;
    double *dqv = W.dqke[var];
    const double *specv = W.spec[var];
    double *p = pm ? pm + (long)SK_A * sktot : nullptr;
    const dim3 g(S.nwork), bl(64);
    if (pm && fwd)
        hipLaunchKernelGGL((k_sa_sweep<1, true>), g, bl, 0, s, S.blocks, work, W, dqv, specv, p, sktot, S.err,
                           S.spin_limit);
    else if (pm)
        hipLaunchKernelGGL((k_sa_sweep<0, true>), g, bl, 0, s, S.blocks, work, W, dqv, specv, p, sktot, S.err,
                           S.spin_limit);
    else if (fwd)
        hipLaunchKernelGGL((k_sa_sweep<1, false>), g, bl, 0, s, S.blocks, work, W, dqv, specv, p, sktot, S.err,
                           S.spin_limit);
    else
        hipLaunchKernelGGL((k_sa_sweep<0, false>), g, bl, 0, s, S.blocks, work, W, dqv, specv, p, sktot, S.err,
                           S.spin_limit);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_sa_update(const BlockGrid &G, const DFields &F, const SaWork &W, const DParams &P, hipStream_t s)
{
    if (G.nwg) hipLaunchKernelGGL(k_sa_update, dim3(G.nwg), dim3(256), 0, s, G, F, W, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
