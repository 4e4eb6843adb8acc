// wcns_rhs_vis.hip -- laminar viscous residual R -= Re^-1 dFv/dxi along i, j, k
// (reference src/RHS.f90:2-30 -> src/sub_tgh.f90:1511-1810 WCNSE5_VIS_virtual)
// and the viscous spectral radii (src/LHS.f90:243-343 spectvisl).
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

// VALUE_HALF_NODE (WCNSE5.f90:799-852), n > 4, of one node array at face f
__device__ inline double vhn(const double *a, const VLine &L, int f)
{
    int n = L.n;
    if (f >= 2 && f <= n - 2)
        return (9.0 * (a[L.at(f)] + a[L.at(f + 1)]) + -1.0 * (a[L.at(f + 2)] + a[L.at(f - 1)])) / 16.0;
    if (f == 1) return (5.0 * a[L.at(1)] + 15.0 * a[L.at(2)] + -5.0 * a[L.at(3)] + 1.0 * a[L.at(4)]) / 16.0;
    if (f == n - 1)
        return (5.0 * a[L.at(n)] + 15.0 * a[L.at(n - 1)] + -5.0 * a[L.at(n - 2)] + 1.0 * a[L.at(n - 3)]) / 16.0;
    if (f == 0) return (35.0 * a[L.at(1)] - 35.0 * a[L.at(2)] + 21.0 * a[L.at(3)] - 5.0 * a[L.at(4)]) / 16.0;
    return (35.0 * a[L.at(n)] - 35.0 * a[L.at(n - 1)] + 21.0 * a[L.at(n - 2)] - 5.0 * a[L.at(n - 3)]) / 16.0;
}

// node derivative along one direction at row l of a pencil (DUVWT_NODE_LINE_virtual)
__device__ inline double dnode(const double *q, const VLine &L, int l, bool fl, bool fr)
{
    int n = L.n;
    if (fl && l == 1)
        return (-11. * q[L.at(1)] + 18. * q[L.at(2)] - 9. * q[L.at(3)] + 2. * q[L.at(4)]) / 6.0;
    if (fl && l == 2)
        return (-3.0 * q[L.at(1)] + -10.0 * q[L.at(2)] + 18.0 * q[L.at(3)] + -6.0 * q[L.at(4)] + 1.0 * q[L.at(5)]) / 12.0;
    if (fr && l == n)
        return -(-11. * q[L.at(n)] + 18. * q[L.at(n - 1)] - 9.0 * q[L.at(n - 2)] + 2. * q[L.at(n - 3)]) / 6.0;
    if (fr && l == n - 1)
        return -(-3.0 * q[L.at(n)] + -10.0 * q[L.at(n - 1)] + 18.0 * q[L.at(n - 2)] + -6.0 * q[L.at(n - 3)] +
                 1.0 * q[L.at(n - 4)]) / 12.0;
    return (8.0 * (q[L.at(l + 1)] - q[L.at(l - 1)]) + -1.0 * (q[L.at(l + 2)] - q[L.at(l - 2)])) / 12.0;
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

// dn layout: component c (0..11 = 4*dir + {u,v,w,T}) of node g at dn[c*total + g]
__global__ void k_vis_node_deriv(BlockGrid G, DFields F, double *dn)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    const double *q[4] = {F.u, F.v, F.w, F.t};
    for (int dir = 0; dir < 3; dir++) {
        VLine L = line_of(b, dir, i, j, k);
        bool fl, fr;
        pencil_flags(F, L, fl, fr);
        int l = dir == 0 ? i : (dir == 1 ? j : k);
        for (int m = 0; m < 4; m++) dn[(long)(4 * dir + m) * F.total + g] = dnode(q[m], L, l, fl, fr);
    }
}

// u,v,w,T at half node f (VALUE_line_half_virtual, sub_tgh.f90:2335-2387)
__device__ inline double value_line_half(const double *q, const VLine &L, int f, bool fl, bool fr)
{
    int n = L.n;
    if (fl && f == 0) return (35.0 * q[L.at(1)] + -35.0 * q[L.at(2)] + 21.0 * q[L.at(3)] + -5.0 * q[L.at(4)]) / 16.0;
    if (fl && f == 1) return (5.0 * q[L.at(1)] + 15.0 * q[L.at(2)] + -5.0 * q[L.at(3)] + 1.0 * q[L.at(4)]) / 16.0;
    if (fr && f == n)
        return (35.0 * q[L.at(n)] + -35.0 * q[L.at(n - 1)] + 21.0 * q[L.at(n - 2)] + -5.0 * q[L.at(n - 3)]) / 16.0;
    if (fr && f == n - 1)
        return (5.0 * q[L.at(n)] + 15.0 * q[L.at(n - 1)] + -5.0 * q[L.at(n - 2)] + 1.0 * q[L.at(n - 3)]) / 16.0;
    return (9.0 * (q[L.at(f)] + q[L.at(f + 1)]) + -1.0 * (q[L.at(f + 2)] + q[L.at(f - 1)])) / 16.0;
}

// direct half-node derivative (DUVWT_half_line_virtual, sub_tgh.f90:2277-2330)
__device__ inline double dhalf(const double *q, const VLine &L, int f, bool fl, bool fr)
{
    int n = L.n;
    if (fl && f == 0)
        return (-71.0 * q[L.at(1)] + 141.0 * q[L.at(2)] + -93.0 * q[L.at(3)] + 23.0 * q[L.at(4)]) / 24.0;
    if (fl && f == 1)
        return (-22.0 * q[L.at(1)] + 17.0 * q[L.at(2)] + 9.0 * q[L.at(3)] + -5.0 * q[L.at(4)] + 1.0 * q[L.at(5)]) / 24.0;
    if (fr && f == n)
        return -(-71.0 * q[L.at(n)] + 141.0 * q[L.at(n - 1)] + -93.0 * q[L.at(n - 2)] + 23.0 * q[L.at(n - 3)]) / 24.0;
    if (fr && f == n - 1)
        return -(-22.0 * q[L.at(n)] + 17.0 * q[L.at(n - 1)] + 9.0 * q[L.at(n - 2)] + -5.0 * q[L.at(n - 3)] +
                 1.0 * q[L.at(n - 4)]) / 24.0;
    return (27.0 * (q[L.at(f + 1)] - q[L.at(f)]) + -1.0 * (q[L.at(f + 2)] - q[L.at(f - 1)])) / 24.0;
}

template <int DIR>
__device__ void face_visc(const DFields &F, const DParams &P, const double *dn, const VLine &L, int f, bool fl,
                          bool fr, double cp_prl, double cp_prt, double *fv)
{
    const double *q4[4] = {F.u, F.v, F.w, F.t};
    double uvwt[4], du[12], kx[9];
#pragma unroll
    for (int m = 0; m < 4; m++) uvwt[m] = value_line_half(q4[m], L, f, fl, fr);
#pragma unroll
    for (int m = 0; m < 4; m++) du[4 * DIR + m] = dhalf(q4[m], L, f, fl, fr);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d == DIR) continue;
#pragma unroll
        for (int m = 0; m < 4; m++) du[4 * d + m] = vhn(dn + (long)(4 * d + m) * F.total, L, f);
    }
    // kxyz_line order: (kcx, etx, ctx, kcy, ety, cty, kcz, etz, ctz), sub_tgh.f90:1571-1581
    const double *ms[9] = {F.kcx, F.etx, F.ctx, F.kcy, F.ety, F.cty, F.kcz, F.etz, F.ctz};
#pragma unroll
    for (int a = 0; a < 9; a++) kx[a] = vhn(ms[a], L, f);
    double vol = vhn(F.vol, L, f);
    // viscosities (sub_tgh.f90:1562-1566): vslt1 = visl + vist, vslt2 = visl*cp/prl + vist*cp/prt;
    // laminar: vist = 0
    double s1, s2;
    {
        int n = L.n;
        const bool turb = P.turb != 0;
        auto V1 = [&](int l) { return F.visl[L.at(l)] + (turb ? F.vist[L.at(l)] : 0.0); };
        auto V2 = [&](int l) { return F.visl[L.at(l)] * cp_prl + (turb ? F.vist[L.at(l)] * cp_prt : 0.0); };
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
    // DUVWT_DXYZ (WCNSE5.f90:854-890)
    double d[12];
#pragma unroll
    for (int m = 1; m <= 3; m++) {
        int m1 = m + m + m - 2, m2 = m1 + 1, m3 = m2 + 1;
        d[m - 1] = (kx[m1 - 1] * du[0] + kx[m2 - 1] * du[4] + kx[m3 - 1] * du[8]) / vol;
        d[m + 2] = (kx[m1 - 1] * du[1] + kx[m2 - 1] * du[5] + kx[m3 - 1] * du[9]) / vol;
        d[m + 5] = (kx[m1 - 1] * du[2] + kx[m2 - 1] * du[6] + kx[m3 - 1] * du[10]) / vol;
        d[m + 8] = (kx[m1 - 1] * du[3] + kx[m2 - 1] * du[7] + kx[m3 - 1] * du[11]) / vol;
    }
    // FLUX_VIS_LINE_new (newvis.f90:283-349); the face normal is the pencil direction's metric
    double vs = s1, kcp = s2;
    double vscc = vs * (2.0 / 3.0);
    double dudx = d[0], dudy = d[1], dudz = d[2], dvdx = d[3], dvdy = d[4], dvdz = d[5];
    double dwdx = d[6], dwdy = d[7], dwdz = d[8];
    double t1 = vscc * (2.0 * dudx - dvdy - dwdz);
    double t5 = vscc * (2.0 * dvdy - dwdz - dudx);
    double t9 = vscc * (2.0 * dwdz - dudx - dvdy);
    double t2 = vs * (dudy + dvdx);
    double t3 = vs * (dudz + dwdx);
    double t6 = vs * (dvdz + dwdy);
    double nx = kx[DIR], ny = kx[3 + DIR], nz = kx[6 + DIR];
    fv[0] = 0.0;
    fv[1] = t1 * nx + t2 * ny + t3 * nz;
    fv[2] = t2 * nx + t5 * ny + t6 * nz;
    fv[3] = t3 * nx + t6 * ny + t9 * nz;
    fv[4] = uvwt[0] * fv[1] + uvwt[1] * fv[2] + uvwt[2] * fv[3] + kcp * (d[9] * nx + d[10] * ny + d[11] * nz);
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
    const double re = P.re;
    const double cp = 1.0 / ((P.gama - 1.0) * P.moo * P.moo);
    const double cp_prl = cp / P.prl, cp_prt = cp / P.prt;

    int a = seg * SEG + 1;
    int bb = a + SEG - 1;
    if (bb > n) bb = n;
    int F0 = a - 3 < 0 ? 0 : a - 3;
    int F1 = bb + 2 > n ? n : bb + 2;
    if (a <= 2 && F1 < 3) F1 = 3;
    if (bb >= n - 1 && F0 > n - 3) F0 = n - 3;

    double fw[6][5];
    for (int fc = F0; fc <= F1; fc++) {
#pragma unroll
        for (int l = 0; l < 5; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) fw[l][m] = fw[l + 1][m];
        face_visc<DIR>(F, P, dn, L, fc, fl, fr, cp_prl, cp_prt, fw[5]);
        // FLUX_DXYZ rows (WCNSE5.f90:229-251); R -= re * dfv
        double df[5];
        if (fc == 3) {
            if (a <= 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = (-23.0 * fw[2][m] + 21.0 * fw[3][m] + 3.0 * fw[4][m] - fw[5][m]) / 24.0;
                long g = L.at(1);
#pragma unroll
                for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df[m];
            }
            if (a <= 2 && bb >= 2) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = (fw[2][m] - 27.0 * fw[3][m] + 27.0 * fw[4][m] - fw[5][m]) / 24.0;
                long g = L.at(2);
#pragma unroll
                for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df[m];
            }
        }
        int row = fc - 2;
        if (row >= 3 && row <= n - 2 && row >= a && row <= bb) {
#pragma unroll
            for (int m = 0; m < 5; m++)
                df[m] = (2250.0 * (fw[3][m] - fw[2][m]) - 125.0 * (fw[4][m] - fw[1][m]) + 9.0 * (fw[5][m] - fw[0][m])) /
                        1920.0;
            long g = L.at(row);
#pragma unroll
            for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df[m];
        }
        if (fc == n) {
            if (bb >= n - 1 && a <= n - 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = -(fw[5][m] - 27.0 * fw[4][m] + 27.0 * fw[3][m] - fw[2][m]) / 24.0;
                long g = L.at(n - 1);
#pragma unroll
                for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df[m];
            }
            if (bb >= n) {
#pragma unroll
                for (int m = 0; m < 5; m++) df[m] = -(-23.0 * fw[5][m] + 21.0 * fw[4][m] + 3.0 * fw[3][m] - fw[2][m]) / 24.0;
                long g = L.at(n);
#pragma unroll
                for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df[m];
            }
        }
    }
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
    const double cp_prl = cp / P.prl, cp_prt = cp / P.prt;
    double fv[5];
    face_visc<DIR>(F, P, dn, L, f, fl, fr, cp_prl, cp_prt, fv);
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
        F.rhs[m * F.total + g] = F.rhs[m * F.total + g] - re * df;
    }
}

// LHS.f90:243-343 spectvisl (laminar) / LHS.f90:141-241 spectvist (mu + mu_t), interior
// nodes (ghost-layer radii only ever multiply a zero dq)
__global__ void k_spectvis(BlockGrid G, DFields F, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    double rm_vol = F.r[g] * F.vol[g];
    double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
    double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
    double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
    double ri = kx * kx + ky * ky + kz * kz;
    double rj = ex * ex + ey * ey + ez * ez;
    double rk = cx * cx + cy * cy + cz * cz;
    double vis = P.turb ? F.visl[g] + F.vist[g] : F.visl[g];
    double coef = 2.0 * vis / (P.reynolds * rm_vol + SMALL);
    double coef1 = P.csrv * coef;
    F.srva[g] = coef1 * ri;
    F.srvb[g] = coef1 * rj;
    F.srvc[g] = coef1 * rk;
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

int launch_spectvis(const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_spectvis, dim3(G.nwg), dim3(256), 0, s, G, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
