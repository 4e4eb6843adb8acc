// wcns_rhs_vis.hip -- laminar viscous residual R -= Re^-1 dFv/dxi along i, j, k
// (reference src/RHS.f90:2-30 -> src/sub_tgh.f90:1511-1810 WCNSE5_VIS_virtual).
// The viscous spectral radii (src/LHS.f90:141-343) are formed by k_timestep.
//
// Pass 1 (k_vis_node_deriv): fourth-order node derivatives of u, v, w, T along
// each direction (UVWT_DER_4th_virtual, sub_tgh.f90:1909-2029), 12 values per
// node, one-sided only at physical boundaries (negative density in ghost 3).
// Pass 2 (k_rhs_vis<DIR>): one thread marches a segment of a pencil; at every
// half node it forms u,v,w,T (VALUE_line_half_virtual), the direct half-node
// derivative along the pencil (DUVWT_half_line_virtual), the cross derivatives,
// metrics, volume and viscosities interpolated from the nodes
// (VALUE_HALF_NODE), the Cartesian gradients (DUVWT_DXYZ) and the viscous flux
// (FLUX_VIS_LINE_new); six fluxes in registers give the FLUX_DXYZ difference.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

constexpr int SEG = PENCIL_SEG;

struct VLine {
    long base, stride;
    int n;
    __device__ long at(int l) const { return base + (long)l * stride; }
};

// Node values along a pencil come from a source q(l), l the pencil index: an
// array (ArrSrc), a unit vector (UnitSrc: probing a linear formula for its
// coefficient of node `at`, which is how the adjoint transposes the linear
// stencils), or an IO policy's variable (IoSrc).
struct ArrSrc {
    const double *a;
    VLine L;
    __device__ double operator()(int l) const { return a[L.at(l)]; }
};
struct UnitSrc {
    int at;
    __device__ double operator()(int l) const { return l == at ? 1.0 : 0.0; }
};

// VALUE_HALF_NODE (WCNSE5.f90:799-852), n > 4, of one node array at face f
template <class S>
__device__ inline auto vhn(const S &q, int n, int f) -> decltype(q(0))
{
    if (f >= 2 && f <= n - 2) return (9.0 * (q(f) + q(f + 1)) + -1.0 * (q(f + 2) + q(f - 1))) / 16.0;
    if (f == 1) return (5.0 * q(1) + 15.0 * q(2) + -5.0 * q(3) + 1.0 * q(4)) / 16.0;
    if (f == n - 1) return (5.0 * q(n) + 15.0 * q(n - 1) + -5.0 * q(n - 2) + 1.0 * q(n - 3)) / 16.0;
    if (f == 0) return (35.0 * q(1) - 35.0 * q(2) + 21.0 * q(3) - 5.0 * q(4)) / 16.0;
    return (35.0 * q(n) - 35.0 * q(n - 1) + 21.0 * q(n - 2) - 5.0 * q(n - 3)) / 16.0;
}

// node derivative along one direction at row l of a pencil (DUVWT_NODE_LINE_virtual)
template <class S>
__device__ inline auto dnode(const S &q, int n, int l, bool fl, bool fr) -> decltype(q(0))
{
    if (fl && l == 1) return (-11. * q(1) + 18. * q(2) - 9. * q(3) + 2. * q(4)) / 6.0;
    if (fl && l == 2) return (-3.0 * q(1) + -10.0 * q(2) + 18.0 * q(3) + -6.0 * q(4) + 1.0 * q(5)) / 12.0;
    if (fr && l == n) return -(-11. * q(n) + 18. * q(n - 1) - 9.0 * q(n - 2) + 2. * q(n - 3)) / 6.0;
    if (fr && l == n - 1)
        return -(-3.0 * q(n) + -10.0 * q(n - 1) + 18.0 * q(n - 2) + -6.0 * q(n - 3) + 1.0 * q(n - 4)) / 12.0;
    return (8.0 * (q(l + 1) - q(l - 1)) + -1.0 * (q(l + 2) - q(l - 2))) / 12.0;
}

// flags of a pencil: uvwt(4,-2) < 0 after the density override (sub_tgh.f90:1923-1935)
__device__ inline void pencil_flags(const DFields &F, const VLine &L, bool &fl, bool &fr)
{
    long lo = L.at(-2), hi = L.at(L.n + 3);
    fl = F.r[lo] < 0.0 || F.t[lo] < 0.0;
    fr = F.r[hi] < 0.0 || F.t[hi] < 0.0;
}

__device__ inline VLine line_of(const DBlock &b, int dir, int i, int j, int k)
{
    VLine L;
    if (dir == 0) { L.base = gidx(b, 0, j, k); L.stride = 1; L.n = b.ni; }
    else if (dir == 1) { L.base = gidx(b, i, 0, k); L.stride = b.sx; L.n = b.nj; }
    else { L.base = gidx(b, i, j, 0); L.stride = b.sy; L.n = b.nk; }
    return L;
}

// dn layout: component c (0..11 = 4*dir + {u,v,w,T}) of node g at dn[c*total + g];
// the derivatives of the four arrays q0..q3 (u, v, w, T for the solver, their
// directions for the tangent), with the pencil flags of the state in F
__device__ inline void node_deriv(const DBlock &b, const DFields &F, const double *const *q, double *dn, long t)
{
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    for (int dir = 0; dir < 3; dir++) {
        VLine L = line_of(b, dir, i, j, k);
        bool fl, fr;
        pencil_flags(F, L, fl, fr);
        int l = dir == 0 ? i : (dir == 1 ? j : k);
        for (int m = 0; m < 4; m++) dn[(long)(4 * dir + m) * F.total + g] = dnode(ArrSrc{q[m], L}, L.n, l, fl, fr);
    }
}

__global__ void k_vis_node_deriv(BlockGrid G, DFields F, double *dn)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    const double *q[4] = {F.u, F.v, F.w, F.t};
    node_deriv(b, F, q, dn, t);
}

// u,v,w,T at half node f (VALUE_line_half_virtual, sub_tgh.f90:2335-2387)
template <class S>
__device__ inline auto value_line_half(const S &q, int n, int f, bool fl, bool fr) -> decltype(q(0))
{
    if (fl && f == 0) return (35.0 * q(1) + -35.0 * q(2) + 21.0 * q(3) + -5.0 * q(4)) / 16.0;
    if (fl && f == 1) return (5.0 * q(1) + 15.0 * q(2) + -5.0 * q(3) + 1.0 * q(4)) / 16.0;
    if (fr && f == n) return (35.0 * q(n) + -35.0 * q(n - 1) + 21.0 * q(n - 2) + -5.0 * q(n - 3)) / 16.0;
    if (fr && f == n - 1) return (5.0 * q(n) + 15.0 * q(n - 1) + -5.0 * q(n - 2) + 1.0 * q(n - 3)) / 16.0;
    return (9.0 * (q(f) + q(f + 1)) + -1.0 * (q(f + 2) + q(f - 1))) / 16.0;
}

// direct half-node derivative (DUVWT_half_line_virtual, sub_tgh.f90:2277-2330)
template <class S>
__device__ inline auto dhalf(const S &q, int n, int f, bool fl, bool fr) -> decltype(q(0))
{
    if (fl && f == 0) return (-71.0 * q(1) + 141.0 * q(2) + -93.0 * q(3) + 23.0 * q(4)) / 24.0;
    if (fl && f == 1) return (-22.0 * q(1) + 17.0 * q(2) + 9.0 * q(3) + -5.0 * q(4) + 1.0 * q(5)) / 24.0;
    if (fr && f == n) return -(-71.0 * q(n) + 141.0 * q(n - 1) + -93.0 * q(n - 2) + 23.0 * q(n - 3)) / 24.0;
    if (fr && f == n - 1)
        return -(-22.0 * q(n) + 17.0 * q(n - 1) + 9.0 * q(n - 2) + -5.0 * q(n - 3) + 1.0 * q(n - 4)) / 24.0;
    return (27.0 * (q(f + 1) - q(f)) + -1.0 * (q(f + 2) - q(f - 1))) / 24.0;
}

// Variables a viscous face reads, as an IO policy numbers them.  Metrics in
// kxyz_line order (kcx, etx, ctx, kcy, ety, cty, kcz, etz, ctz), sub_tgh.f90:1571-1581.
enum { VU = 0, VT = 3, VDN = 4, VM = 16, VVOL = 25, VVISL = 26, VVIST = 27 };
// pool index (kcx, kcy, kcz, etx, ..., ctz) of kxyz_line entry a
__host__ __device__ constexpr int kxyz_pool(int a) { return (a % 3) * 3 + a / 3; }

// The solver's IO: double loads, rows subtracted from R as re * dFv
struct VisIO {
    using T = double;
    const double *q4[4];          // u, v, w, T
    const double *dn;
    const double *ms[9];
    const double *vol, *visl, *vist;
    double *rhs;
    long total;
    double re;
    __device__ double val(int var, long g, int) const
    {
        if (var < 4) return q4[var][g];
        if (var < VM) return dn[(long)(var - VDN) * total + g];
        if (var < VVOL) return ms[var - VM][g];
        if (var == VVOL) return vol[g];
        if (var == VVISL) return visl[g];
        return vist[g];
    }
    __device__ void row(long g, int, const double *df) const
    {
#pragma unroll
        for (int m = 0; m < 5; m++) rhs[m * total + g] = rhs[m * total + g] - re * df[m];
    }
};

// Tangent IO: D1 loads of (state, direction); mu_t is held fixed
struct VisTanIO {
    using T = D1;
    VisIO s;
    const double *d4[4];          // du, dv, dw, dT
    const double *ddn, *dm, *dvisl;   // dm: metrics in pool order, vol at 9
    double *drhs;
    __device__ D1 val(int var, long g, int l) const
    {
        double x = s.val(var, g, l), dx;
        if (var < 4) dx = d4[var][g];
        else if (var < VM) dx = ddn[(long)(var - VDN) * s.total + g];
        else if (var < VVOL) dx = dm[(long)kxyz_pool(var - VM) * s.total + g];
        else if (var == VVOL) dx = dm[9 * s.total + g];
        else if (var == VVISL) dx = dvisl[g];
        else dx = 0.0;
        return D1(x, dx);
    }
    __device__ void row(long g, int, const D1 *df) const
    {
#pragma unroll
        for (int m = 0; m < 5; m++) drhs[m * s.total + g] = drhs[m * s.total + g] - s.re * df[m].d;
    }
};

template <class IO>
struct IoSrc {
    const IO *io;
    int var;
    VLine L;
    __device__ typename IO::T operator()(int l) const { return io->val(var, L.at(l), l); }
};

// The viscous flux of a half node in two parts: visc_inputs interpolates the
// 28 quantities the flux is made of from the pencil's nodes (all linear in the
// node values), visc_core forms the flux from them.  Layout of the inputs:
// 0..3 u, v, w, T; 4..15 the derivatives du[4*dir + {u,v,w,T}] (the pencil
// direction's from the half-node difference, the others interpolated from the
// node derivatives); 16..24 the metrics in kxyz_line order; 25 the volume;
// 26, 27 the two viscosities.
constexpr int VIN = 28;

template <int DIR, class IO>
__device__ inline void visc_inputs(const IO &io, const VLine &L, int f, bool fl, bool fr, bool turb, double cp_prl,
                                   double cp_prt, typename IO::T *in)
{
    using T = typename IO::T;
    const int n = L.n;
    auto src = [&](int var) { return IoSrc<IO>{&io, var, L}; };
#pragma unroll
    for (int m = 0; m < 4; m++) in[m] = value_line_half(src(VU + m), n, f, fl, fr);
#pragma unroll
    for (int m = 0; m < 4; m++) in[4 + 4 * DIR + m] = dhalf(src(VU + m), n, f, fl, fr);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d == DIR) continue;
#pragma unroll
        for (int m = 0; m < 4; m++) in[4 + 4 * d + m] = vhn(src(VDN + 4 * d + m), n, f);
    }
#pragma unroll
    for (int a = 0; a < 9; a++) in[16 + a] = vhn(src(VM + a), n, f);
    in[25] = vhn(src(VVOL), n, f);
    // viscosities (sub_tgh.f90:1562-1566): vslt1 = visl + vist, vslt2 = visl*cp/prl + vist*cp/prt;
    // laminar: vist = 0
    T s1, s2;
    {
        auto V1 = [&](int l) { long g = L.at(l); return io.val(VVISL, g, l) + (turb ? io.val(VVIST, g, l) : T(0.0)); };
        auto V2 = [&](int l) {
            long g = L.at(l);
            return io.val(VVISL, g, l) * cp_prl + (turb ? io.val(VVIST, g, l) * cp_prt : T(0.0));
        };
        if (f >= 2 && f <= n - 2) {
            s1 = (9.0 * (V1(f) + V1(f + 1)) + -1.0 * (V1(f + 2) + V1(f - 1))) / 16.0;
            s2 = (9.0 * (V2(f) + V2(f + 1)) + -1.0 * (V2(f + 2) + V2(f - 1))) / 16.0;
        } else if (f == 1) {
            s1 = (5.0 * V1(1) + 15.0 * V1(2) + -5.0 * V1(3) + 1.0 * V1(4)) / 16.0;
            s2 = (5.0 * V2(1) + 15.0 * V2(2) + -5.0 * V2(3) + 1.0 * V2(4)) / 16.0;
        } else if (f == n - 1) {
            s1 = (5.0 * V1(n) + 15.0 * V1(n - 1) + -5.0 * V1(n - 2) + 1.0 * V1(n - 3)) / 16.0;
            s2 = (5.0 * V2(n) + 15.0 * V2(n - 1) + -5.0 * V2(n - 2) + 1.0 * V2(n - 3)) / 16.0;
        } else if (f == 0) {
            s1 = (35.0 * V1(1) - 35.0 * V1(2) + 21.0 * V1(3) - 5.0 * V1(4)) / 16.0;
            s2 = (35.0 * V2(1) - 35.0 * V2(2) + 21.0 * V2(3) - 5.0 * V2(4)) / 16.0;
        } else {
            s1 = (35.0 * V1(n) - 35.0 * V1(n - 1) + 21.0 * V1(n - 2) - 5.0 * V1(n - 3)) / 16.0;
            s2 = (35.0 * V2(n) - 35.0 * V2(n - 1) + 21.0 * V2(n - 2) - 5.0 * V2(n - 3)) / 16.0;
        }
    }
    in[26] = s1;
    in[27] = s2;
}

template <int DIR, typename T>
__device__ inline void visc_core(const T *in, T *fv)
{
    const T *uvwt = in, *du = in + 4, *kx = in + 16;
    const T vol = in[25];
    // DUVWT_DXYZ (WCNSE5.f90:854-890)
    T d[12];
#pragma unroll
    for (int m = 1; m <= 3; m++) {
        int m1 = m + m + m - 2, m2 = m1 + 1, m3 = m2 + 1;
        d[m - 1] = (kx[m1 - 1] * du[0] + kx[m2 - 1] * du[4] + kx[m3 - 1] * du[8]) / vol;
        d[m + 2] = (kx[m1 - 1] * du[1] + kx[m2 - 1] * du[5] + kx[m3 - 1] * du[9]) / vol;
        d[m + 5] = (kx[m1 - 1] * du[2] + kx[m2 - 1] * du[6] + kx[m3 - 1] * du[10]) / vol;
        d[m + 8] = (kx[m1 - 1] * du[3] + kx[m2 - 1] * du[7] + kx[m3 - 1] * du[11]) / vol;
    }
    // FLUX_VIS_LINE_new (newvis.f90:283-349); the face normal is the pencil direction's metric
    T vs = in[26], kcp = in[27];
    T vscc = vs * (2.0 / 3.0);
    T dudx = d[0], dudy = d[1], dudz = d[2], dvdx = d[3], dvdy = d[4], dvdz = d[5];
    T dwdx = d[6], dwdy = d[7], dwdz = d[8];
    T t1 = vscc * (2.0 * dudx - dvdy - dwdz);
    T t5 = vscc * (2.0 * dvdy - dwdz - dudx);
    T t9 = vscc * (2.0 * dwdz - dudx - dvdy);
    T t2 = vs * (dudy + dvdx);
    T t3 = vs * (dudz + dwdx);
    T t6 = vs * (dvdz + dwdy);
    T nx = kx[DIR], ny = kx[3 + DIR], nz = kx[6 + DIR];
    fv[0] = 0.0;
    fv[1] = t1 * nx + t2 * ny + t3 * nz;
    fv[2] = t2 * nx + t5 * ny + t6 * nz;
    fv[3] = t3 * nx + t6 * ny + t9 * nz;
    fv[4] = uvwt[0] * fv[1] + uvwt[1] * fv[2] + uvwt[2] * fv[3] + kcp * (d[9] * nx + d[10] * ny + d[11] * nz);
}

template <int DIR, class IO>
__device__ inline void face_visc(const IO &io, const VLine &L, int f, bool fl, bool fr, bool turb, double cp_prl,
                                 double cp_prt, typename IO::T *fv)
{
    typename IO::T in[VIN];
    visc_inputs<DIR>(io, L, f, fl, fr, turb, cp_prl, cp_prt, in);
    visc_core<DIR>(in, fv);
}

__device__ inline VisIO vis_io(const DFields &F, const DParams &P, const double *dn)
{
    VisIO io;
    io.q4[0] = F.u; io.q4[1] = F.v; io.q4[2] = F.w; io.q4[3] = F.t;
    io.dn = dn;
    io.ms[0] = F.kcx; io.ms[1] = F.etx; io.ms[2] = F.ctx;
    io.ms[3] = F.kcy; io.ms[4] = F.ety; io.ms[5] = F.cty;
    io.ms[6] = F.kcz; io.ms[7] = F.etz; io.ms[8] = F.ctz;
    io.vol = F.vol; io.visl = F.visl; io.vist = F.vist;
    io.rhs = F.rhs;
    io.total = F.total;
    io.re = P.re;
    return io;
}

// one pencil's rows a..bb: six fluxes in registers give the FLUX_DXYZ rows
// (WCNSE5.f90:229-251), handed to the IO policy
template <int DIR, class IO>
__device__ void vis_pencil(const IO &io, const VLine &L, bool fl, bool fr, int a, int bb, bool turb, double cp_prl,
                           double cp_prt)
{
    using T = typename IO::T;
    const int n = L.n;
    int F0 = a - 3 < 0 ? 0 : a - 3;
    int F1 = bb + 2 > n ? n : bb + 2;
    if (a <= 2 && F1 < 3) F1 = 3;
    if (bb >= n - 1 && F0 > n - 3) F0 = n - 3;

    T fw[6][5];
    for (int fc = F0; fc <= F1; fc++) {
#pragma unroll
        for (int l = 0; l < 5; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) fw[l][m] = fw[l + 1][m];
        face_visc<DIR>(io, L, fc, fl, fr, turb, cp_prl, cp_prt, fw[5]);
        T df[5];
        if (fc == 3) {
            if (a <= 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = (-23.0 * fw[2][m] + 21.0 * fw[3][m] + 3.0 * fw[4][m] - fw[5][m]) / 24.0;
                io.row(L.at(1), 1, df);
            }
            if (a <= 2 && bb >= 2) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = (fw[2][m] - 27.0 * fw[3][m] + 27.0 * fw[4][m] - fw[5][m]) / 24.0;
                io.row(L.at(2), 2, df);
            }
        }
        int row = fc - 2;
        if (row >= 3 && row <= n - 2 && row >= a && row <= bb) {
#pragma unroll
            for (int m = 0; m < 5; m++)
                df[m] = (2250.0 * (fw[3][m] - fw[2][m]) - 125.0 * (fw[4][m] - fw[1][m]) + 9.0 * (fw[5][m] - fw[0][m])) /
                        1920.0;
            io.row(L.at(row), row, df);
        }
        if (fc == n) {
            if (bb >= n - 1 && a <= n - 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = -(fw[5][m] - 27.0 * fw[4][m] + 27.0 * fw[3][m] - fw[2][m]) / 24.0;
                io.row(L.at(n - 1), n - 1, df);
            }
            if (bb >= n) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = -(-23.0 * fw[5][m] + 21.0 * fw[4][m] + 3.0 * fw[3][m] - fw[2][m]) / 24.0;
                io.row(L.at(n), n, df);
            }
        }
    }
}

// DIR0 = 3: the i direction run as j pencils of the transposed view (the same
// code as DIR = 1 on transposed fields; its own instance so profiles tell it apart)
template <int DIR0>
__global__ void __launch_bounds__(128) k_rhs_vis(BlockGrid G, DFields F, DParams P, const double *dn)
{
    constexpr int DIR = DIR0 == 3 ? 1 : DIR0;
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nbk = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nbk;
    const int nseg = (n + SEG - 1) / SEG;
    if (t >= npen * nseg) return;
    int seg = (int)(t / npen);
    long pen = t - (long)seg * npen;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    VLine L = line_of(b, DIR, i0, j0, k0);
    bool fl, fr;
    pencil_flags(F, L, fl, fr);
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    vis_pencil<DIR>(vis_io(F, P, dn), L, fl, fr, a, bb, P.turb != 0, cp / P.prl, cp / P.prt);
}

// The same viscous residual split in two passes (used for the j, k and, on the
// transposed view, the i direction): k_vis_face forms the flux of one half node
// per thread (face f of a pencil stored at the pencil's node f of fvb, 5 pool
// arrays), k_vis_rows applies the FLUX_DXYZ rows (WCNSE5.f90:229-251) per node.
// One face per thread keeps the register footprint of face_visc alone, so the
// kernel runs several waves per SIMD without spilling; the faces that two
// segments of the marching form would share are formed once.
template <int DIR0>
__global__ void __launch_bounds__(256) k_vis_face(BlockGrid G, DFields F, DParams P, const double *dn, double *fvb)
{
    constexpr int DIR = DIR0 == 3 ? 1 : DIR0;
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nbk = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nbk;
    if (t >= npen * (n + 1)) return;
    const int f = (int)(t / npen);
    const long pen = t - (long)f * npen;
    const int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    const int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    const VLine L = line_of(b, DIR, i0, j0, k0);
    bool fl, fr;
    pencil_flags(F, L, fl, fr);
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    double fv[5];
    face_visc<DIR>(vis_io(F, P, dn), L, f, fl, fr, P.turb != 0, cp / P.prl, cp / P.prt, fv);
    const long g = L.at(f);
#pragma unroll
    for (int m = 0; m < 5; m++) fvb[m * F.total + g] = fv[m];
}

template <int DIR0>
__global__ void __launch_bounds__(256) k_vis_rows(BlockGrid G, DFields F, DParams P, const double *fvb)
{
    constexpr int DIR = DIR0 == 3 ? 1 : DIR0;
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nbk = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nbk;
    if (t >= npen * n) return;
    const int l = (int)(t / npen) + 1;
    const long pen = t - (long)(l - 1) * npen;
    const int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    const int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    const VLine L = line_of(b, DIR, i0, j0, k0);
    const double re = P.re;
    const long g = L.at(l);
    auto fa = [&](int m, int o) { return fvb[m * F.total + L.at(o)]; };
#pragma unroll
    for (int m = 0; m < 5; m++) {
        double df;
        if (l == 1) df = (-23.0 * fa(m, 0) + 21.0 * fa(m, 1) + 3.0 * fa(m, 2) - fa(m, 3)) / 24.0;
        else if (l == 2) df = (fa(m, 0) - 27.0 * fa(m, 1) + 27.0 * fa(m, 2) - fa(m, 3)) / 24.0;
        else if (l == n - 1) df = -(fa(m, n) - 27.0 * fa(m, n - 1) + 27.0 * fa(m, n - 2) - fa(m, n - 3)) / 24.0;
        else if (l == n) df = -(-23.0 * fa(m, n) + 21.0 * fa(m, n - 1) + 3.0 * fa(m, n - 2) - fa(m, n - 3)) / 24.0;
        else
            df = (2250.0 * (fa(m, l) - fa(m, l - 1)) - 125.0 * (fa(m, l + 1) - fa(m, l - 2)) +
                  9.0 * (fa(m, l + 2) - fa(m, l - 3))) / 1920.0;
        // DIR0 = 3: the transposed i-direction contribution, every interior
        // row written once here (0 - re df: the bits of the zeroed buffer's)
        F.rhs[m * F.total + g] = (DIR0 == 3 ? 0.0 : F.rhs[m * F.total + g]) - re * df;
    }
}

// ------------------------------------------------ tangent and adjoint (one block)

// dR -= re * d(dFv/dxi) along DIR; the pencil segments of the solver
template <int DIR>
__global__ void __launch_bounds__(128) k_rhs_vis_jvp(DBlock b, DFields F, DParams P, const double *dn, AdTan D)
{
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nbk = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nbk;
    const int nseg = (n + SEG - 1) / SEG;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npen * nseg) return;
    int seg = (int)(t / npen);
    long pen = t - (long)seg * npen;
    int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    VLine L = line_of(b, DIR, i0, j0, k0);
    bool fl, fr;
    pencil_flags(F, L, fl, fr);
    VisTanIO io;
    io.s = vis_io(F, P, dn);
    const long T5 = F.total;
    for (int m = 0; m < 3; m++) io.d4[m] = D.dp + (1 + m) * T5;
    io.d4[3] = D.dt;
    io.ddn = D.ddn;
    io.dm = D.dm;
    io.dvisl = D.dvisl;
    io.drhs = D.dr;
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    vis_pencil<DIR>(io, L, fl, fr, a, bb, P.turb != 0, cp / P.prl, cp / P.prt);
}

// Adjoint of the viscous residual along DIR, per face instead of per pencil:
//  k_visadj_face: the weight of face f's flux, the transposed FLUX_DXYZ rows
//    (coefficients found by probing the row formulas with a unit face), times
//    the transposed Jacobian of the flux core w.r.t. its 28 inputs (formed in
//    dual numbers, one input at a time, from the inputs interpolated afresh);
//    the 28 results go to gf at the face's pencil node f;
//  k_visadj_node: every node gathers them through the transposed
//    interpolations of the faces whose stencils hold it (coefficients by
//    probing with a unit node) into the gradients of u, v, w, T, the node
//    derivatives, the metrics, the volume and mu.
template <class S>
__device__ inline auto vis_row(const S &fa, int n, int l) -> decltype(fa(0))
{
    if (l == 1) return (-23.0 * fa(0) + 21.0 * fa(1) + 3.0 * fa(2) - fa(3)) / 24.0;
    if (l == 2) return (fa(0) - 27.0 * fa(1) + 27.0 * fa(2) - fa(3)) / 24.0;
    if (l == n - 1) return -(fa(n) - 27.0 * fa(n - 1) + 27.0 * fa(n - 2) - fa(n - 3)) / 24.0;
    if (l == n) return -(-23.0 * fa(n) + 21.0 * fa(n - 1) + 3.0 * fa(n - 2) - fa(n - 3)) / 24.0;
    return (2250.0 * (fa(l) - fa(l - 1)) - 125.0 * (fa(l + 1) - fa(l - 2)) + 9.0 * (fa(l + 2) - fa(l - 3))) / 1920.0;
}

template <int DIR>
__global__ void __launch_bounds__(256) k_visadj_face(DBlock b, DFields F, DParams P, const double *dn, AdAdj A)
{
    const int n = DIR == 0 ? b.ni : (DIR == 1 ? b.nj : b.nk);
    const int na = DIR == 0 ? b.nj : b.ni;
    const int nbk = DIR == 2 ? b.nj : b.nk;
    const long npen = (long)na * nbk;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npen * (n + 1)) return;
    const int f = (int)(t / npen);
    const long pen = t - (long)f * npen;
    const int ia = (int)(pen % na) + 1, ib = (int)(pen / na) + 1;
    const int i0 = DIR == 0 ? 0 : ia, j0 = DIR == 1 ? 0 : (DIR == 0 ? ia : ib), k0 = DIR == 2 ? 0 : ib;
    const VLine L = line_of(b, DIR, i0, j0, k0);
    bool fl, fr;
    pencil_flags(F, L, fl, fr);
    // weight of the face's flux: R -= re * (row formulas), rows f-2..f+3
    double lf[5] = {0, 0, 0, 0, 0};
    const int r0 = f - 2 < 1 ? 1 : f - 2, r1 = f + 3 > n ? n : f + 3;
    for (int l = r0; l <= r1; l++) {
        const double c = vis_row(UnitSrc{f}, n, l);
        if (c == 0.0) continue;
        const long gl = L.at(l);
#pragma unroll
        for (int m = 0; m < 5; m++) lf[m] = lf[m] - P.re * c * A.w[m * F.total + gl];
    }
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    double in[VIN];
    visc_inputs<DIR>(vis_io(F, P, dn), L, f, fl, fr, P.turb != 0, cp / P.prl, cp / P.prt, in);
    const long g = L.at(f);
    for (int q = 0; q < VIN; q++) {
        D1 x[VIN], fv[5];
#pragma unroll
        for (int e = 0; e < VIN; e++) x[e] = D1(in[e], e == q ? 1.0 : 0.0);
        visc_core<DIR>(x, fv);
        double acc = 0.0;
#pragma unroll
        for (int m = 1; m < 5; m++) acc = acc + lf[m] * fv[m].d;
        A.gf[(long)q * F.total + g] = acc;
    }
}

template <int DIR>
__global__ void __launch_bounds__(256) k_visadj_node(DBlock b, DFields F, DParams P, AdAdj A)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e) return;
    const int i = (int)(t % (b.ni + 6)) - 2, j = (int)((t / (b.ni + 6)) % (b.nj + 6)) - 2;
    const int k = (int)(t / ((long)(b.ni + 6) * (b.nj + 6))) - 2;
    const bool in_i = i >= 1 && i <= b.ni, in_j = j >= 1 && j <= b.nj, in_k = k >= 1 && k <= b.nk;
    if (!(DIR == 0 ? (in_j && in_k) : (DIR == 1 ? (in_i && in_k) : (in_i && in_j)))) return;
    const VLine L = line_of(b, DIR, i, j, k);
    const int n = L.n;
    const int l = DIR == 0 ? i : (DIR == 1 ? j : k);
    bool fl, fr;
    pencil_flags(F, L, fl, fr);
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    const double cp_prl = cp / P.prl;
    double gu[4] = {0, 0, 0, 0}, gd[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, gmt[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gv = 0.0;
    const int f0 = l - 4 < 0 ? 0 : l - 4, f1 = l + 4 > n ? n : l + 4;
    for (int f = f0; f <= f1; f++) {
        const double cv = value_line_half(UnitSrc{l}, n, f, fl, fr);
        const double ch = dhalf(UnitSrc{l}, n, f, fl, fr);
        const double cn = vhn(UnitSrc{l}, n, f);
        if (cv == 0.0 && ch == 0.0 && cn == 0.0) continue;
        const long gf = L.at(f);
        auto G = [&](int q) { return A.gf[(long)q * F.total + gf]; };
#pragma unroll
        for (int m = 0; m < 4; m++) gu[m] = gu[m] + cv * G(m) + ch * G(4 + 4 * DIR + m);
        if (cn != 0.0) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (d == DIR) continue;
#pragma unroll
                for (int m = 0; m < 4; m++) gd[4 * d + m] = gd[4 * d + m] + cn * G(4 + 4 * d + m);
            }
#pragma unroll
            for (int a = 0; a < 9; a++) gmt[kxyz_pool(a)] = gmt[kxyz_pool(a)] + cn * G(16 + a);
            gmt[9] = gmt[9] + cn * G(25);
            gv = gv + cn * (G(26) + G(27) * cp_prl);
        }
    }
    const long g = L.at(l), T5 = F.total;
#pragma unroll
    for (int m = 0; m < 3; m++) A.gp[(1 + m) * T5 + g] = A.gp[(1 + m) * T5 + g] + gu[m];
    A.gt[g] = A.gt[g] + gu[3];
    if (l >= 1 && l <= n) {
#pragma unroll
        for (int c = 0; c < 12; c++) A.gdn[c * T5 + g] = A.gdn[c * T5 + g] + gd[c];
#pragma unroll
        for (int c = 0; c < 10; c++) A.gm[c * T5 + g] = A.gm[c * T5 + g] + gmt[c];
        A.gvisl[g] = A.gvisl[g] + gv;
    }
}

// tangent of the node derivatives: the same stencils on du, dv, dw, dT
__global__ void k_dn_jvp(DBlock b, DFields F, AdTan D, double *ddn)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)b.ni * b.nj * b.nk) return;
    const double *q[4] = {D.dp + F.total, D.dp + 2 * F.total, D.dp + 3 * F.total, D.dt};
    node_deriv(b, F, q, ddn, t);
}

// adjoint of the node derivatives: every node of the ghosted block gathers, for
// each direction along which its transverse indices are interior, the rows
// l' = l-4..l+4 whose stencil holds it (the coefficient found by probing the
// stencil with a unit vector)
__global__ void k_dn_vjp(DBlock b, DFields F, AdAdj A)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e) return;
    int i = (int)(t % (b.ni + 6)) - 2, j = (int)((t / (b.ni + 6)) % (b.nj + 6)) - 2;
    int k = (int)(t / ((long)(b.ni + 6) * (b.nj + 6))) - 2;
    const long g = gidx(b, i, j, k);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int dir = 0; dir < 3; dir++) {
        bool in_i = i >= 1 && i <= b.ni, in_j = j >= 1 && j <= b.nj, in_k = k >= 1 && k <= b.nk;
        bool ok = dir == 0 ? (in_j && in_k) : (dir == 1 ? (in_i && in_k) : (in_i && in_j));
        if (!ok) continue;
        VLine L = line_of(b, dir, i, j, k);
        bool fl, fr;
        pencil_flags(F, L, fl, fr);
        int l = dir == 0 ? i : (dir == 1 ? j : k);
        int r0 = l - 4 < 1 ? 1 : l - 4, r1 = l + 4 > L.n ? L.n : l + 4;
        for (int r = r0; r <= r1; r++) {
            double c = dnode(UnitSrc{l}, L.n, r, fl, fr);
            if (c == 0.0) continue;
            long gr = L.at(r);
            for (int m = 0; m < 4; m++) acc[m] = acc[m] + c * A.gdn[(long)(4 * dir + m) * F.total + gr];
        }
    }
    for (int m = 0; m < 3; m++) A.gp[(1 + m) * F.total + g] = A.gp[(1 + m) * F.total + g] + acc[m];
    A.gt[g] = A.gt[g] + acc[3];
}


}  // namespace

int launch_vis_node_deriv(const BlockGrid &G, const DFields &F, double *dn, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_vis_node_deriv, dim3(G.nwg), dim3(256), 0, s, G, F, dn);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long rhs_vis_threads(int dir, const DBlock &b)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    return npen * ((n + SEG - 1) / SEG);
}

int launch_rhs_vis_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *dn, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    dim3 grid(G.nwg), blk(128);
    if (dir == 0) hipLaunchKernelGGL(k_rhs_vis<0>, grid, blk, 0, s, G, F, P, dn);
    if (dir == 1) hipLaunchKernelGGL(k_rhs_vis<1>, grid, blk, 0, s, G, F, P, dn);
    if (dir == 2) hipLaunchKernelGGL(k_rhs_vis<2>, grid, blk, 0, s, G, F, P, dn);
    if (dir == 3) hipLaunchKernelGGL(k_rhs_vis<3>, grid, blk, 0, s, G, F, P, dn);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long vis_face_threads(int dir, const DBlock &b)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    return npen * (n + 1);
}

long vis_row_threads(int dir, const DBlock &b)
{
    return (long)b.ni * b.nj * b.nk;
}

int launch_vis_faces(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *dn, double *fvb,
                     hipStream_t s)
{
    if (G.nwg == 0) return 0;
    if (dir == 1) hipLaunchKernelGGL(k_vis_face<1>, dim3(G.nwg), dim3(256), 0, s, G, F, P, dn, fvb);
    else if (dir == 2) hipLaunchKernelGGL(k_vis_face<2>, dim3(G.nwg), dim3(256), 0, s, G, F, P, dn, fvb);
    else if (dir == 3) hipLaunchKernelGGL(k_vis_face<3>, dim3(G.nwg), dim3(256), 0, s, G, F, P, dn, fvb);
    else return -2;
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_vis_rows(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *fvb, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    if (dir == 1) hipLaunchKernelGGL(k_vis_rows<1>, dim3(G.nwg), dim3(256), 0, s, G, F, P, fvb);
    else if (dir == 2) hipLaunchKernelGGL(k_vis_rows<2>, dim3(G.nwg), dim3(256), 0, s, G, F, P, fvb);
    else if (dir == 3) hipLaunchKernelGGL(k_vis_rows<3>, dim3(G.nwg), dim3(256), 0, s, G, F, P, fvb);
    else return -2;
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_rhs_vis_jvp(int dir, const DBlock &b, const DFields &F, const DParams &P, const double *dn, const AdTan &D,
                       hipStream_t s)
{
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    long nt = npen * ((n + SEG - 1) / SEG);
    dim3 grid((unsigned)((nt + 127) / 128)), blk(128);
    if (dir == 0) hipLaunchKernelGGL(k_rhs_vis_jvp<0>, grid, blk, 0, s, b, F, P, dn, D);
    if (dir == 1) hipLaunchKernelGGL(k_rhs_vis_jvp<1>, grid, blk, 0, s, b, F, P, dn, D);
    if (dir == 2) hipLaunchKernelGGL(k_rhs_vis_jvp<2>, grid, blk, 0, s, b, F, P, dn, D);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_rhs_vis_vjp(int dir, const DBlock &b, const DFields &F, const DParams &P, const double *dn, const AdAdj &A,
                       hipStream_t s)
{
    long npen = dir == 0 ? (long)b.nj * b.nk : (dir == 1 ? (long)b.ni * b.nk : (long)b.ni * b.nj);
    int n = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    long nf = npen * (n + 1), e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    dim3 gf((unsigned)((nf + 255) / 256)), gn((unsigned)((e + 255) / 256)), blk(256);
    if (dir == 0) {
        hipLaunchKernelGGL(k_visadj_face<0>, gf, blk, 0, s, b, F, P, dn, A);
        hipLaunchKernelGGL(k_visadj_node<0>, gn, blk, 0, s, b, F, P, A);
    } else if (dir == 1) {
        hipLaunchKernelGGL(k_visadj_face<1>, gf, blk, 0, s, b, F, P, dn, A);
        hipLaunchKernelGGL(k_visadj_node<1>, gn, blk, 0, s, b, F, P, A);
    } else {
        hipLaunchKernelGGL(k_visadj_face<2>, gf, blk, 0, s, b, F, P, dn, A);
        hipLaunchKernelGGL(k_visadj_node<2>, gn, blk, 0, s, b, F, P, A);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_dn_jvp(const DBlock &b, const DFields &F, const AdTan &D, double *ddn, hipStream_t s)
{
    long n = (long)b.ni * b.nj * b.nk;
    hipLaunchKernelGGL(k_dn_jvp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, F, D, ddn);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_dn_vjp(const DBlock &b, const DFields &F, const AdAdj &A, hipStream_t s)
{
    long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    hipLaunchKernelGGL(k_dn_vjp, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, s, b, F, A);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
