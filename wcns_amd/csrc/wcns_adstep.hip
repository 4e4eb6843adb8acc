// wcns_adstep.hip -- tangent and adjoint of the parts of a pseudo-time step
// around the residual and the sweeps: the boundary conditions (BC.f90 as in
// bc_point), the temperature, the inviscid and viscous spectral radii and the
// LU-SGS diagonal (Convect.f90:1338-1461, LHS.f90:141-343,602-645), the limited update
// (Sub2.f90:603-796) and the transposed interface averagings
// (parallel.f90:3020-3326).  Each pointwise piece is written once, generic in
// the scalar type: double reproduces the solver's arithmetic, D1 gives its
// derivative along one seeded direction, and the adjoints contract a few
// seeded evaluations (one per input) with the output weights.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// ------------------------------------------------------------ boundaries
// Nodes of boundary point (i,j,k): o[0] the face node, o[1..3] the ghosts
// outward; in[0..2] the nodes it reads (inward in its own block, or behind
// the image point in the partner block for an interface).
__device__ inline void bc_nodes(const DRegion &g, const DBlock &b, const DBlock &bt, int i, int j, int k, long *o,
                                long *in)
{
    const int *l3 = g.l3;
#pragma unroll
    for (int l = 0; l < 4; l++) o[l] = gidx(b, i + l * l3[0], j + l * l3[1], k + l * l3[2]);
    if (g.bctype < 0) {
        int it, jt, kt;
        region_image(g, i, j, k, it, jt, kt);
        const int *tl = g.t_l3;
#pragma unroll
        for (int l = 0; l < 3; l++) in[l] = gidx(bt, it - (l + 1) * tl[0], jt - (l + 1) * tl[1], kt - (l + 1) * tl[2]);
    } else {
#pragma unroll
        for (int l = 0; l < 3; l++) in[l] = gidx(b, i - (l + 1) * l3[0], j - (l + 1) * l3[1], k - (l + 1) * l3[2]);
    }
}

// outputs a point writes (bit q: o[q]), how many of its input nodes matter,
// and how many variables of each: the no-slip wall also reads T (the value
// the previous step left), every other BC only r, u, v, w, p
__device__ inline int bc_mask(int bctype) { return bctype < 0 ? 0xE : 0xF; }
__device__ inline int bc_nin(int bctype)
{
    if (bctype == 5) return 0;
    if (bctype == 4 || bctype == 6) return 1;
    return 3;
}
__device__ inline int bc_nvar(int bctype, const DParams &P) { return (bctype == 2 && P.nvis == 1) ? 6 : 5; }

// The values one point writes (r, u, v, w, p) from the values it reads
// (r, u, v, w, p, T of each input node).  nrm: the face's unit normal
// (getnxyz) for BC 4 and BC 2.
template <typename T>
__device__ inline void bc_eval(const DRegion &g, const DParams &P, const double *nrm, const T (&in)[3][6],
                               T (&o)[4][5])
{
    const double gama = P.gama;
    if (g.bctype < 0) {   // Sub2.f90:206-322: the three ghosts from the partner
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
            for (int m = 0; m < 5; m++) o[l + 1][m] = in[l][m];
        return;
    }
    switch (g.bctype) {
    case 4: {   // BC.f90:460-607
        const double nx = nrm[0], ny = nrm[1], nz = nrm[2];
        const T rin = in[0][0], uin = in[0][1], vin = in[0][2], win = in[0][3], pin = in[0][4];
        T s_in = pin / pow(rin, gama);
        double s_oo = P.poo / ::pow(P.roo, gama);
        T cin = sqrt(gama * pin / rin);
        double sq = ::sqrt(nx * nx + ny * ny + nz * nz);
        sq = sq < SML_SSS ? SML_SSS : sq;
        double cgm1 = (double)g.s_lr / sq;
        double nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1;
        double vnoo = nxa * P.uoo + nya * P.voo + nza * P.woo;
        T vnin = nxa * uin + nya * vin + nza * win;
        T mach = sqrt(uin * uin + vin * vin + win * win) / cin;
        T rb, ub, vb, wb, pb;
        if (mach < 1.0) {
            T rp = vnin + 2.0 * cin / (gama - 1.0);
            double rn = vnoo - 2.0 * P.coo / (gama - 1.0);
            T vnb = 0.5 * (rp + rn);
            T ab = 0.25 * (gama - 1.0) * (rp - rn);
            T sb, uref, vref, wref, vnref;
            if (vnb > 0.0) { sb = s_in; uref = uin; vref = vin; wref = win; vnref = vnin; }
            else { sb = T(s_oo); uref = T(P.uoo); vref = T(P.voo); wref = T(P.woo); vnref = T(vnoo); }
            rb = pow(ab * ab / gama / sb, 1.0 / (gama - 1.0));
            ub = uref + nxa * (vnb - vnref);
            vb = vref + nya * (vnb - vnref);
            wb = wref + nza * (vnb - vnref);
            pb = sb * pow(rb, gama);
        } else if (vnin > 0.0) {
            rb = rin; ub = uin; vb = vin; wb = win; pb = pin;
        } else {
            rb = T(P.roo); ub = T(P.uoo); vb = T(P.voo); wb = T(P.woo); pb = T(P.poo);
        }
        const T v[5] = {rb, ub, vb, wb, pb};
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int m = 0; m < 5; m++) o[q][m] = v[m];
        o[3][0] = -rb;
        return;
    }
    case 5: {   // BC.f90:609-665
        const double v[5] = {P.roo, P.uoo, P.voo, P.woo, P.poo};
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int m = 0; m < 5; m++) o[q][m] = T(v[m]);
        return;
    }
    case 6:     // BC.f90:667-749
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int m = 0; m < 5; m++) o[q][m] = in[0][m];
        return;
    case 2:
        if (P.nvis == 1) {   // BC.f90:1244-1409 viscid_wall (adiabatic, or fixed wall temperature)
            const double moocp = P.moo * P.moo * gama;
            T pres, tres;
            if (P.twall_fixed) {
                pres = (4.0 * in[0][4] - in[1][4]) / 3.0;
                if (pres < PMIN_LIMIT) pres = in[0][4];
                tres = T(P.twall_nd);
            } else {
                pres = (4.0 * in[0][4] - in[1][4]) / 3.0;
                tres = (4.0 * in[0][5] - in[1][5]) / 3.0;
                if (pres < PMIN_LIMIT) pres = in[0][4];
                if (tres < TMIN_LIMIT) tres = in[0][5];
            }
            T dres = moocp * pres / tres;
            o[0][0] = dres; o[0][1] = T(0.0); o[0][2] = T(0.0); o[0][3] = T(0.0); o[0][4] = pres;
#pragma unroll
            for (int l = 0; l < 3; l++) {
                o[l + 1][0] = l == 2 ? -in[l][0] : in[l][0];
                o[l + 1][1] = -in[l][1];
                o[l + 1][2] = -in[l][2];
                o[l + 1][3] = -in[l][3];
                o[l + 1][4] = in[l][4];
            }
            return;
        }
        // fall through: BC.f90:1412-1486 inviscid_wall + dif_average
    case 3: {   // BC.f90:227-457 boundary_3
        const bool wall = g.bctype == 2;   // (nvis = 0 here)
        const double nx = wall ? nrm[0] : g.nrm[0], ny = wall ? nrm[1] : g.nrm[1], nz = wall ? nrm[2] : g.nrm[2];
#pragma unroll
        for (int l = 0; l < 3; l++) {
            const T *a = in[l];
            T sss1 = wall ? 2.0 * (a[1] * nx + a[2] * ny + a[3] * nz) : 2.0 * (nx * a[1] + ny * a[2] + nz * a[3]);
            o[l + 1][0] = (wall && l == 2) ? -a[0] : a[0];
            o[l + 1][1] = a[1] - sss1 * nx;
            o[l + 1][2] = a[2] - sss1 * ny;
            o[l + 1][3] = a[3] - sss1 * nz;
            o[l + 1][4] = a[4];
        }
#pragma unroll
        for (int m = 0; m < 5; m++) o[0][m] = 0.5 * (o[1][m] + in[0][m]);
        return;
    }
    default:
        return;
    }
}

__device__ inline void bc_normal(const DRegion &g, const DFields &F, long g0, double *nrm)
{
    nrm[0] = nrm[1] = nrm[2] = 0.0;
    if (g.bctype == 4 || g.bctype == 2) getnxyz(F, g0, g.l3, nrm[0], nrm[1], nrm[2]);
}

__device__ void bc_fwd_point(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F, const DParams &P,
                             double *tan, double *tape, int t)
{
    int i, j, k;
    region_point(g, t, i, j, k);
    long o[4], in[3];
    bc_nodes(g, b, bt, i, j, k, o, in);
    double *pc[6] = {F.r, F.u, F.v, F.w, F.p, F.t}, *po[5] = {F.rn, F.un, F.vn, F.wn, F.pn};
    const long T5 = F.total;
    D1 x[3][6], y[4][5];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
        for (int m = 0; m < 6; m++) {
            const double v = pc[m][in[l]];
            tape[((long)t * 3 + l) * 6 + m] = v;
            x[l][m] = D1(v, tan ? tan[m * T5 + in[l]] : 0.0);
        }
    double nrm[3];
    bc_normal(g, F, o[0], nrm);
    bc_eval<D1>(g, P, nrm, x, y);
    const int mask = bc_mask(g.bctype);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (!(mask & (1 << q))) continue;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            pc[m][o[q]] = y[q][m].v;
            po[m][o[q]] = y[q][m].v;
            if (tan) tan[m * T5 + o[q]] = y[q][m].d;
        }
    }
}

__device__ void bc_rev_point(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F, const DParams &P,
                             double *lam, const double *tape, int t)
{
    int i, j, k;
    region_point(g, t, i, j, k);
    long o[4], in[3];
    bc_nodes(g, b, bt, i, j, k, o, in);
    const long T5 = F.total;
    const int mask = bc_mask(g.bctype);
    double la[4][5];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int m = 0; m < 5; m++) {
            la[q][m] = 0.0;
            if (mask & (1 << q)) {
                la[q][m] = lam[m * T5 + o[q]];
                lam[m * T5 + o[q]] = 0.0;
            }
        }
    double xv[3][6];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
        for (int m = 0; m < 6; m++) xv[l][m] = tape[((long)t * 3 + l) * 6 + m];
    double nrm[3];
    bc_normal(g, F, o[0], nrm);
    const int nin = bc_nin(g.bctype), nvar = bc_nvar(g.bctype, P);
    for (int l = 0; l < nin; l++)
        for (int m = 0; m < nvar; m++) {
            D1 x[3][6], y[4][5];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 6; c++) x[a][c] = D1(xv[a][c], (a == l && c == m) ? 1.0 : 0.0);
            bc_eval<D1>(g, P, nrm, x, y);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (!(mask & (1 << q))) continue;    // an output the point does not write
#pragma unroll
                for (int c = 0; c < 5; c++) s = s + la[q][c] * y[q][c].d;
            }
            lam[m * T5 + in[l]] = lam[m * T5 + in[l]] + s;
        }
}

__global__ void k_bc_fwd(DRegion g, DBlock b, DBlock bt, DFields F, DParams P, double *tan, double *tape)
{
    if (g.serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int t = 0; t < g.npts; t++) bc_fwd_point(g, b, bt, F, P, tan, tape, t);
        return;
    }
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    bc_fwd_point(g, b, bt, F, P, tan, tape, t);
}

__global__ void k_bc_rev(DRegion g, DBlock b, DBlock bt, DFields F, DParams P, double *lam, const double *tape)
{
    if (g.serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int t = g.npts - 1; t >= 0; t--) bc_rev_point(g, b, bt, F, P, lam, tape, t);
        return;
    }
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    bc_rev_point(g, b, bt, F, P, lam, tape, t);
}

// ------------------------------------------------ radii, dt and coed
// The node's T and mu (initial_q_dq_c_t: get_c_t_only RHS.f90:307-409,
// compute_visl_ns GasModel.f90:219-235)
template <typename T>
__device__ inline void t_mu(const DParams &P, const T *pr, T &tm, T &mu)
{
    double moo2 = P.moo * P.moo;
    T a2 = P.gama * pr[4] / pr[0];
    tm = moo2 * a2;
    mu = tm * sqrt(tm) * (1.0 + P.visc) / (tm + P.visc);
}

// spectinv + spectvisl/vist + localdt0 + getcoed of one node: out = sra,
// srb, src, srva, srvb, srvc, coed.  The viscous radii (LHS.f90:141-343)
// follow the primitives through mu(T(p, r)) when csrv > 0 (mu_t held);
// otherwise they are the stored values.
template <typename T>
__device__ inline void tsv_node(const DParams &P, const T *pr, const DFields &F, long g, T (&out)[7])
{
    T um = pr[1], vm = pr[2], wm = pr[3];
    T cm = sqrt(P.gama * pr[4] / pr[0]);
    double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
    double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
    double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
    T sv[3] = {T(F.srva[g]), T(F.srvb[g]), T(F.srvc[g])};
    if (P.nvis == 1 && P.csrv > 0.0) {
        T tm, mu;
        t_mu<T>(P, pr, tm, mu);
        T vis = P.turb ? mu + F.vist[g] : mu;
        T rm_vol = pr[0] * F.vol[g];
        T coef = 2.0 * vis / (P.reynolds * rm_vol + SMALL);
        T coef1 = P.csrv * coef;
        sv[0] = coef1 * (kx * kx + ky * ky + kz * kz);
        sv[1] = coef1 * (ex * ex + ey * ey + ez * ez);
        sv[2] = coef1 * (cx * cx + cy * cy + cz * cz);
    }
    T cita = kx * um + ky * vm + kz * wm + 0.0;
    T citb = ex * um + ey * vm + ez * wm + 0.0;
    T citc = cx * um + cy * vm + cz * wm + 0.0;
    T sra = fabs(cita) + cm * ::sqrt(kx * kx + ky * ky + kz * kz);
    T srb = fabs(citb) + cm * ::sqrt(ex * ex + ey * ey + ez * ez);
    T src = fabs(citc) + cm * ::sqrt(cx * cx + cy * cy + cz * cz);
    T rv = sv[0] + sv[1] + sv[2];
    T rabc = sra + srb + src + rv;
    T dt;
    if (P.timedt_rate > SMALL) {
        T a = P.cfl / rabc;
        double c = P.timedt_rate / F.vol[g];
        dt = a < c ? a : T(c);
    } else {
        dt = P.cfl / rabc;
    }
    T rad_ns = sra + srb + src;
    if (P.nvis >= 1) rad_ns = rad_ns + sv[0] + sv[1] + sv[2];
    T ct = 1.0 / dt;
    out[0] = sra; out[1] = srb; out[2] = src;
    out[3] = sv[0]; out[4] = sv[1]; out[5] = sv[2];
    out[6] = 1.0 / (ct + 1.0 * 1.0 * rad_ns);
}

__device__ inline bool node_of(const BlockGrid &G, DBlock &b, long &g)
{
    long t;
    grid_thread(G, b, t);
    if (t >= (long)b.ni * b.nj * b.nk) return false;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    g = gidx(b, i, j, k);
    return true;
}

// tangent ts (7 arrays, the order of tsv_node's outputs) from tp
__global__ void k_ts_jvp(BlockGrid G, DFields F, DParams P, const double *tp, double *ts)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    const long T5 = F.total;
    const double *pc[5] = {F.r, F.u, F.v, F.w, F.p};
    D1 pr[5], out[7];
#pragma unroll
    for (int m = 0; m < 5; m++) pr[m] = D1(pc[m][g], tp[m * T5 + g]);
    tsv_node<D1>(P, pr, F, g, out);
#pragma unroll
    for (int q = 0; q < 7; q++) ts[q * T5 + g] = out[q].d;
}

// lp += from the weights of the radii (ls, 3), viscous radii (lv, 3) and coed (lc)
__global__ void k_ts_vjp(BlockGrid G, DFields F, DParams P, const double *ls, const double *lv, const double *lc,
                         double *lp)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    const long T5 = F.total;
    double w[7];
#pragma unroll
    for (int q = 0; q < 3; q++) { w[q] = ls[q * T5 + g]; w[3 + q] = lv[q * T5 + g]; }
    w[6] = lc[g];
    bool any = false;
#pragma unroll
    for (int q = 0; q < 7; q++) any = any || w[q] != 0.0;
    if (!any) return;
    const double *pc[5] = {F.r, F.u, F.v, F.w, F.p};
    for (int q = 0; q < 5; q++) {
        D1 pr[5], out[7];
#pragma unroll
        for (int m = 0; m < 5; m++) pr[m] = D1(pc[m][g], m == q ? 1.0 : 0.0);
        tsv_node<D1>(P, pr, F, g, out);
        double s = 0.0;
#pragma unroll
        for (int o = 0; o < 7; o++) s = s + w[o] * out[o].d;
        lp[q * T5 + g] = lp[q * T5 + g] + s;
    }
}

// the state's temperature after the step: T(p, r) of initial_q_dq_c_t at the
// interior nodes (before the primitive averaging)
__global__ void k_tout_jvp(BlockGrid G, DFields F, DParams P, const double *pf, const double *tp, double *tt)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    const long T5 = F.total;
    D1 pr[5], tm, mu;
#pragma unroll
    for (int m = 0; m < 5; m++) pr[m] = D1(pf[m * T5 + g], tp[m * T5 + g]);
    t_mu<D1>(P, pr, tm, mu);
    tt[g] = tm.d;
}

__global__ void k_tout_vjp(BlockGrid G, DFields F, DParams P, const double *pf, const double *lt, double *lp)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    const long T5 = F.total;
    const double w = lt[g];
    if (w == 0.0) return;
    for (int q = 0; q < 5; q++) {
        D1 pr[5], tm, mu;
#pragma unroll
        for (int m = 0; m < 5; m++) pr[m] = D1(pf[m * T5 + g], m == q ? 1.0 : 0.0);
        t_mu<D1>(P, pr, tm, mu);
        lp[q * T5 + g] = lp[q * T5 + g] + w * tm.d;
    }
}

// ------------------------------------------------------------ update
template <typename T>
__device__ inline void upd_node(const DParams &P, const T *p0, const T *d, T *pn)
{
    const double gama = P.gama, gamam1 = gama - 1.0;
    T q[5];
    q[0] = p0[0]; q[1] = p0[0] * p0[1]; q[2] = p0[0] * p0[2]; q[3] = p0[0] * p0[3];
    q[4] = getem(p0[0], p0[1], p0[2], p0[3], p0[4], gama);
    T qt[5];
#pragma unroll
    for (int m = 0; m < 5; m++) qt[m] = q[m] + d[m];
    T rm = qt[0], rm1 = 1.0 / rm;
    T um = qt[1] * rm1, vm = qt[2] * rm1, wm = qt[3] * rm1, em = qt[4];
    T pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
    T dp = -0.3 + fabs((pm - p0[4]) / p0[4]);
    T falpha = 1. / (1. + 2. * (dp < 0.0 ? T(0.0) : dp));
#pragma unroll
    for (int m = 0; m < 5; m++) qt[m] = q[m] + d[m] * falpha;
    rm = qt[0]; rm1 = 1.0 / rm;
    um = qt[1] * rm1; vm = qt[2] * rm1; wm = qt[3] * rm1; em = qt[4];
    pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
    pn[0] = rm; pn[1] = um; pn[2] = vm; pn[3] = wm; pn[4] = pm;
}

__global__ void k_upd_jvp(BlockGrid G, DParams P, const double *pf, const double *dq, const double *tpf,
                          const double *tdq, double *tpn, long total)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    D1 p0[5], d[5], pn[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        p0[m] = D1(pf[m * total + g], tpf[m * total + g]);
        d[m] = D1(dq[m * total + g], tdq[m * total + g]);
    }
    upd_node<D1>(P, p0, d, pn);
#pragma unroll
    for (int m = 0; m < 5; m++) tpn[m * total + g] = pn[m].d;
}

__global__ void k_upd_vjp(BlockGrid G, DParams P, const double *pf, const double *dq, const double *lpn, double *lp,
                          double *ld, long total)
{
    DBlock b;
    long g;
    if (!node_of(G, b, g)) return;
    double w[5], pv[5], dv[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        w[m] = lpn[m * total + g];
        pv[m] = pf[m * total + g];
        dv[m] = dq[m * total + g];
    }
    for (int q = 0; q < 10; q++) {
        D1 p0[5], d[5], pn[5];
#pragma unroll
        for (int m = 0; m < 5; m++) {
            p0[m] = D1(pv[m], q == m ? 1.0 : 0.0);
            d[m] = D1(dv[m], q == 5 + m ? 1.0 : 0.0);
        }
        upd_node<D1>(P, p0, d, pn);
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 5; m++) s = s + w[m] * pn[m].d;
        if (q < 5) lp[q * total + g] = lp[q * total + g] + s;
        else ld[(q - 5) * total + g] = s;
    }
}

// ---------------------------------------------- interface averagings
// forward: value(qs) = 0.5 (value(qt) vp + value(qs)), vp = vol(qs)/vol(qt)
// for dQ, 1 for the primitives
__device__ inline void avg2pm_t_point(const DRegion &g, const DBlock &b, const DBlock &bt, const double *vol,
                                      double *lam, long total, int use_vol, int nv, int t)
{
    int i, j, k, it, jt, kt;
    region_point(g, t, i, j, k);
    region_image(g, i, j, k, it, jt, kt);
    const long qs = gidx(b, i, j, k), qt = gidx(bt, it, jt, kt);
    const double vp = use_vol ? vol[qs] / vol[qt] : 1.0;
    for (int m = 0; m < nv; m++) {
        const double l = lam[m * total + qs];
        lam[m * total + qt] = lam[m * total + qt] + 0.5 * vp * l;
        lam[m * total + qs] = 0.5 * l;
    }
}

__global__ void k_avg2pm_t(DRegion g, DBlock b, DBlock bt, const double *vol, double *lam, long total, int use_vol,
                           int nv)
{
    if (g.serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int t = g.npts - 1; t >= 0; t--) avg2pm_t_point(g, b, bt, vol, lam, total, use_vol, nv, t);
        return;
    }
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    avg2pm_t_point(g, b, bt, vol, lam, total, use_vol, nv, t);
}

// forward: every written member of a group takes mean(buf) (* vol for dQ)
__global__ void k_npp_assign_t(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok,
                               int ngroups, const DBlock *blocks, const double *vol, double *lam, long total,
                               double *bufb, int use_vol, int nv)
{
    int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    const int s0 = grp_start[gi], n = grp_n[gi];
    double S[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < n; s++) {
        if (!write_ok[s0 + s]) continue;
        const int *p = pts + 4 * (s0 + s);
        const long q = gidx(blocks[p[0]], p[1], p[2], p[3]);
        const double f = use_vol ? vol[q] : 1.0;
        for (int m = 0; m < nv; m++) {
            S[m] = S[m] + lam[m * total + q] * f;
            lam[m * total + q] = 0.0;
        }
    }
    for (int s = 0; s < n; s++)
        for (int m = 0; m < nv; m++) bufb[6 * (s0 + s) + m] = S[m] / (n * 1.0);
}

// forward: buf(s) = value(q_s) (/ vol for dQ), before the pairwise averages.
// A node can be a member of several groups, so the entries are taken per
// node: node u owns entries ent[ustart[u] .. ustart[u+1]-1], summed in order.
__global__ void k_npp_gather_t(const int *pts, const int *ustart, const int *ent, int nuniq, const DBlock *blocks,
                               const double *vol, double *lam, long total, const double *bufb, int use_vol, int nv)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nuniq) return;
    const int e0 = ustart[u], e1 = ustart[u + 1];
    const int *p = pts + 4 * ent[e0];
    const long q = gidx(blocks[p[0]], p[1], p[2], p[3]);
    const double f = use_vol ? vol[q] : 1.0;
    for (int m = 0; m < nv; m++) {
        double a = 0.0;
        for (int e = e0; e < e1; e++) a = a + bufb[6 * ent[e] + m] / f;
        lam[m * total + q] = lam[m * total + q] + a;
    }
}

__global__ void k_zero_ghosts(DBlock b, double *arr, long total, int nv)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= e) return;
    const int i = (int)(t % (b.ni + 6)) - 2, j = (int)((t / (b.ni + 6)) % (b.nj + 6)) - 2;
    const int k = (int)(t / ((long)(b.ni + 6) * (b.nj + 6))) - 2;
    if (i >= 1 && i <= b.ni && j >= 1 && j <= b.nj && k >= 1 && k <= b.nk) return;
    for (int m = 0; m < nv; m++) arr[m * total + b.off + t] = 0.0;
}

}  // namespace

int launch_bc_fwd(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, double *tan,
                  double *tape, hipStream_t s)
{
    const DBlock &b = blocks[g.nb];
    const DBlock &bt = blocks[g.bctype < 0 ? g.nbt : g.nb];
    if (g.serial) hipLaunchKernelGGL(k_bc_fwd, dim3(1), dim3(64), 0, s, g, b, bt, F, P, tan, tape);
    else hipLaunchKernelGGL(k_bc_fwd, grid1(g.npts, 128), dim3(128), 0, s, g, b, bt, F, P, tan, tape);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_bc_rev(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, double *lam,
                  const double *tape, hipStream_t s)
{
    const DBlock &b = blocks[g.nb];
    const DBlock &bt = blocks[g.bctype < 0 ? g.nbt : g.nb];
    if (g.serial) hipLaunchKernelGGL(k_bc_rev, dim3(1), dim3(64), 0, s, g, b, bt, F, P, lam, tape);
    else hipLaunchKernelGGL(k_bc_rev, grid1(g.npts, 128), dim3(128), 0, s, g, b, bt, F, P, lam, tape);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_ts_jvp(const BlockGrid &G, const DFields &F, const DParams &P, const double *tp, double *ts, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_ts_jvp, dim3(G.nwg), dim3(256), 0, s, G, F, P, tp, ts);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_ts_vjp(const BlockGrid &G, const DFields &F, const DParams &P, const double *ls, const double *lv,
                  const double *lc, double *lp, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_ts_vjp, dim3(G.nwg), dim3(256), 0, s, G, F, P, ls, lv, lc, lp);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_tout_jvp(const BlockGrid &G, const DFields &F, const DParams &P, const double *pf, const double *tp,
                    double *tt, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_tout_jvp, dim3(G.nwg), dim3(256), 0, s, G, F, P, pf, tp, tt);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_tout_vjp(const BlockGrid &G, const DFields &F, const DParams &P, const double *pf, const double *lt,
                    double *lp, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_tout_vjp, dim3(G.nwg), dim3(256), 0, s, G, F, P, pf, lt, lp);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_upd_jvp(const BlockGrid &G, const DParams &P, const double *pf, const double *dq, const double *tpf,
                   const double *tdq, double *tpn, long total, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_upd_jvp, dim3(G.nwg), dim3(256), 0, s, G, P, pf, dq, tpf, tdq, tpn, total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_upd_vjp(const BlockGrid &G, const DParams &P, const double *pf, const double *dq, const double *lpn,
                   double *lp, double *ld, long total, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_upd_vjp, dim3(G.nwg), dim3(256), 0, s, G, P, pf, dq, lpn, lp, ld, total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_avg2pm_t(const DRegion &g, const DBlock *blocks, const double *vol, double *lam, long total, int use_vol,
                    int nv, hipStream_t s)
{
    const DBlock &b = blocks[g.nb], &bt = blocks[g.nbt];
    if (g.serial) hipLaunchKernelGGL(k_avg2pm_t, dim3(1), dim3(64), 0, s, g, b, bt, vol, lam, total, use_vol, nv);
    else hipLaunchKernelGGL(k_avg2pm_t, grid1(g.npts, 128), dim3(128), 0, s, g, b, bt, vol, lam, total, use_vol, nv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_npp_assign_t(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok, int ngroups,
                        const DBlock *blocks, const double *vol, double *lam, long total, double *bufb, int use_vol,
                        int nv, hipStream_t s)
{
    if (ngroups <= 0) return 0;
    hipLaunchKernelGGL(k_npp_assign_t, grid1(ngroups, 128), dim3(128), 0, s, grp_start, grp_n, pts, write_ok,
                       ngroups, blocks, vol, lam, total, bufb, use_vol, nv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_npp_gather_t(const int *pts, const int *ustart, const int *ent, int nuniq, const DBlock *blocks,
                        const double *vol, double *lam, long total, const double *bufb, int use_vol, int nv,
                        hipStream_t s)
{
    if (nuniq <= 0) return 0;
    hipLaunchKernelGGL(k_npp_gather_t, grid1(nuniq, 128), dim3(128), 0, s, pts, ustart, ent, nuniq, blocks, vol, lam,
                       total, bufb, use_vol, nv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_zero_ghosts(const DBlock &b, double *arr, long total, int nv, hipStream_t s)
{
    const long e = (long)(b.ni + 6) * (b.nj + 6) * (b.nk + 6);
    hipLaunchKernelGGL(k_zero_ghosts, grid1(e, 256), dim3(256), 0, s, b, arr, total, nv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
