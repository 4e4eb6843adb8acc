// wcns_step.hip -- per-step kernels around the inviscid residual: boundary
// ghost fill, speed of sound / temperature, spectral radii and local time
// step, interface averaging, LU-SGS + SGS sweeps, limited update and the
// residual monitor.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

// ------------------------------------------------------------------ BCs
// A boundary value goes to both primitive buffers.  The update writes only
// the interior nodes into the other buffer, so after the step's buffer swap
// the ghosts read back are the ones this step's boundary conditions set, as
// in the reference (whose update leaves the ghosts alone).
__device__ inline void setp(const DFields &F, long g, double r, double u, double v, double w, double p)
{
    F.r[g] = r; F.u[g] = u; F.v[g] = v; F.w[g] = w; F.p[g] = p;
    F.rn[g] = r; F.un[g] = u; F.vn[g] = v; F.wn[g] = w; F.pn[g] = p;
}

// turbulence variable m (nu~ or k, then SST omega) and its BC_farfield_turbulence
// value (turblence.f90:2321-2338)
__device__ inline double *tvar(const DFields &F, int m) { return m == 0 ? F.qke : F.qkw; }
__device__ inline double tinf(const DParams &P, int m) { return P.turb == 1 ? P.muoo : (m == 0 ? P.koo : P.omgoo); }

__device__ void bc_point(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F,
                         const DParams &P, int i, int j, int k)
{
    const int *l3 = g.l3;
    long g0 = gidx(b, i, j, k);
    long s1 = gidx(b, i + l3[0], j + l3[1], k + l3[2]);
    long s2 = gidx(b, i + 2 * l3[0], j + 2 * l3[1], k + 2 * l3[2]);
    long s3 = gidx(b, i + 3 * l3[0], j + 3 * l3[1], k + 3 * l3[2]);
    long t1 = gidx(b, i - l3[0], j - l3[1], k - l3[2]);
    long t2 = gidx(b, i - 2 * l3[0], j - 2 * l3[1], k - 2 * l3[2]);
    long t3 = gidx(b, i - 3 * l3[0], j - 3 * l3[1], k - 3 * l3[2]);
    const double gama = P.gama;
    if (g.bctype < 0) {
        // Sub2.f90:206-322 boundary_n1_vir3 (cic1 = 0)
        int it, jt, kt;
        region_image(g, i, j, k, it, jt, kt);
        const int *tl = g.t_l3;
        long q1 = gidx(bt, it - tl[0], jt - tl[1], kt - tl[2]);
        long q2 = gidx(bt, it - 2 * tl[0], jt - 2 * tl[1], kt - 2 * tl[2]);
        long q3 = gidx(bt, it - 3 * tl[0], jt - 3 * tl[1], kt - 3 * tl[2]);
        if (P.turb) {
            // Sub2.f90:266-280: first ghost from the partner, face node = mean of the two neighbours
            const double vt = F.vist[q1];
            F.vist[s1] = vt;
            F.vist[g0] = (vt + F.vist[t1]) * 0.5;
            for (int m = 0; m < P.turb; m++) {
                double *q = tvar(F, m);
                const double qk = q[q1];
                q[s1] = qk;
                q[g0] = (qk + q[t1]) * 0.5;
            }
        }
        setp(F, s1, F.r[q1], F.u[q1], F.v[q1], F.w[q1], F.p[q1]);
        setp(F, s2, F.r[q2], F.u[q2], F.v[q2], F.w[q2], F.p[q2]);
        setp(F, s3, F.r[q3], F.u[q3], F.v[q3], F.w[q3], F.p[q3]);
        return;
    }
    switch (g.bctype) {
    case 4: {   // BC.f90:460-607 boundary_4
        double nx, ny, nz;
        getnxyz(F, g0, l3, nx, ny, nz);
        double rin = F.r[t1], uin = F.u[t1], vin = F.v[t1], win = F.w[t1], pin = F.p[t1];
        double s_in = pin / pow(rin, gama);
        double s_oo = P.poo / pow(P.roo, gama);
        double cin = sqrt(gama * pin / rin);
        double sq = sqrt(nx * nx + ny * ny + nz * nz);
        sq = sq < SML_SSS ? SML_SSS : sq;
        double cgm1 = (double)g.s_lr / sq;
        double nxa = nx * cgm1, nya = ny * cgm1, nza = nz * cgm1;
        double vnoo = nxa * P.uoo + nya * P.voo + nza * P.woo;
        double vnin = nxa * uin + nya * vin + nza * win;
        double mach = sqrt(uin * uin + vin * vin + win * win) / cin;
        double rb, ub, vb, wb, pb;
        if (mach < 1.0) {
            double rp = vnin + 2.0 * cin / (gama - 1.0);
            double rn = vnoo - 2.0 * P.coo / (gama - 1.0);
            double vnb = 0.5 * (rp + rn);
            double ab = 0.25 * (gama - 1.0) * (rp - rn);
            double sb, uref, vref, wref, vnref;
            if (vnb > 0.0) { sb = s_in; uref = uin; vref = vin; wref = win; vnref = vnin; }
            else { sb = s_oo; uref = P.uoo; vref = P.voo; wref = P.woo; vnref = vnoo; }
            rb = pow(ab * ab / gama / sb, 1.0 / (gama - 1.0));
            ub = uref + nxa * (vnb - vnref);
            vb = vref + nya * (vnb - vnref);
            wb = wref + nza * (vnb - vnref);
            pb = sb * pow(rb, gama);
        } else if (vnin > 0.0) {
            rb = rin; ub = uin; vb = vin; wb = win; pb = pin;
        } else {
            rb = P.roo; ub = P.uoo; vb = P.voo; wb = P.woo; pb = P.poo;
        }
        setp(F, g0, rb, ub, vb, wb, pb);
        if (P.turb) {
            // sub_tgh.f90:713-787 BC_farfield_turbulence_tgh: inflow (ut < 0) takes the
            // freestream value, outflow the interior neighbour's, face node and first ghost
            const double tem_lr = (double)l3[g.s_nd - 1];
            double ut;
            if (g.s_nd == 1) ut = tem_lr * (F.kcx[g0] * ub + F.kcy[g0] * vb + F.kcz[g0] * wb);
            else if (g.s_nd == 2) ut = tem_lr * (F.etx[g0] * ub + F.ety[g0] * vb + F.etz[g0] * wb);
            else ut = tem_lr * (F.ctx[g0] * ub + F.cty[g0] * vb + F.ctz[g0] * wb);
            for (int m = 0; m < P.turb; m++) {
                double *q = tvar(F, m);
                const double qv = ut < 0.0 ? tinf(P, m) : q[t1];
                q[g0] = qv;
                q[s1] = qv;
            }
        }
        setp(F, s1, rb, ub, vb, wb, pb);
        setp(F, s2, rb, ub, vb, wb, pb);
        setp(F, s3, -rb, ub, vb, wb, pb);
        break;
    }
    case 5:     // BC.f90:609-665 boundary_5
        setp(F, g0, P.roo, P.uoo, P.voo, P.woo, P.poo);
        setp(F, s1, P.roo, P.uoo, P.voo, P.woo, P.poo);
        setp(F, s2, P.roo, P.uoo, P.voo, P.woo, P.poo);
        setp(F, s3, P.roo, P.uoo, P.voo, P.woo, P.poo);
        for (int m = 0; m < P.turb; m++) { tvar(F, m)[g0] = tinf(P, m); tvar(F, m)[s1] = tinf(P, m); }   // turblence.f90:2321-2338
        break;
    case 6: {   // BC.f90:667-749 boundary_6
        double rm = F.r[t1], um = F.u[t1], vm = F.v[t1], wm = F.w[t1], pm = F.p[t1];
        setp(F, g0, rm, um, vm, wm, pm);
        setp(F, s1, rm, um, vm, wm, pm);
        setp(F, s2, rm, um, vm, wm, pm);
        setp(F, s3, rm, um, vm, wm, pm);
        for (int m = 0; m < P.turb; m++) {   // turblence.f90:2340-2355
            double *q = tvar(F, m);
            const double qv = q[t1];
            q[g0] = qv;
            q[s1] = qv;
        }
        break;
    }
    case 2:
        if (P.nvis == 1) {   // BC.f90:1244-1409 viscid_wall
            double moocp = P.moo * P.moo * gama;
            setp(F, s1, F.r[t1], -F.u[t1], -F.v[t1], -F.w[t1], F.p[t1]);
            // turblence.f90:2267-2306 BC_wall_turbulence: odd reflection of nu~ / k in the
            // first ghost; SST omega there 1600 mu/(rho Re d^2) - omega(first interior), with
            // mu from the face node's temperature and the face node's density before this
            // BC sets it.  BC_face_dif_turbulence: face value = mean of ghost and first interior
            double qs1 = 0.0, ws1 = 0.0;
            if (P.turb) { qs1 = -F.qke[t1]; F.qke[s1] = qs1; }
            if (P.turb == 2) {
                const double tm = F.t[g0];
                const double vsl = tm * sqrt(tm) * (1.0 + P.visc) / (tm + P.visc);
                const double d = F.dist[t1];
                ws1 = 1600.0 * vsl / F.r[g0] / P.reynolds / (d * d) - F.qkw[t1];
                F.qkw[s1] = ws1;
            }
            double pres, tres;
            if (P.twall_fixed) {
                pres = (4.0 * F.p[t1] - F.p[t2]) / 3.0;
                if (pres < PMIN_LIMIT) pres = F.p[t1];
                tres = P.twall_nd;
            } else {
                pres = (4.0 * F.p[t1] - F.p[t2]) / 3.0;
                tres = (4.0 * F.t[t1] - F.t[t2]) / 3.0;
                if (pres < PMIN_LIMIT) pres = F.p[t1];
                if (tres < TMIN_LIMIT) tres = F.t[t1];
            }
            double dres = moocp * pres / tres;
            setp(F, g0, dres, 0.0, 0.0, 0.0, pres);
            if (P.turb) F.qke[g0] = 0.5 * (F.qke[t1] + qs1);
            if (P.turb == 2) F.qkw[g0] = 0.5 * (F.qkw[t1] + ws1);
            setp(F, s2, F.r[t2], -F.u[t2], -F.v[t2], -F.w[t2], F.p[t2]);
            setp(F, s3, -F.r[t3], -F.u[t3], -F.v[t3], -F.w[t3], F.p[t3]);
        } else {             // BC.f90:1412-1486 inviscid_wall + dif_average
            double nx, ny, nz;
            getnxyz(F, g0, l3, nx, ny, nz);
            long ss[3] = {s1, s2, s3}, tt[3] = {t1, t2, t3};
            for (int l = 0; l < 3; l++) {
                long a = tt[l];
                double sss1 = 2.0 * (F.u[a] * nx + F.v[a] * ny + F.w[a] * nz);
                setp(F, ss[l], l == 2 ? -F.r[a] : F.r[a], F.u[a] - sss1 * nx, F.v[a] - sss1 * ny,
                     F.w[a] - sss1 * nz, F.p[a]);
            }
            setp(F, g0, 0.5 * (F.r[s1] + F.r[t1]), 0.5 * (F.u[s1] + F.u[t1]), 0.5 * (F.v[s1] + F.v[t1]),
                 0.5 * (F.w[s1] + F.w[t1]), 0.5 * (F.p[s1] + F.p[t1]));
        }
        break;
    case 3: {   // BC.f90:227-457 boundary_3, 3-D branch (plane normal from the face centre)
        double nx = g.nrm[0], ny = g.nrm[1], nz = g.nrm[2];
        long ss[3] = {s1, s2, s3}, tt[3] = {t1, t2, t3};
        for (int l = 0; l < 3; l++) {
            long a = tt[l];
            double sss1 = 2.0 * (nx * F.u[a] + ny * F.v[a] + nz * F.w[a]);
            setp(F, ss[l], F.r[a], F.u[a] - sss1 * nx, F.v[a] - sss1 * ny, F.w[a] - sss1 * nz, F.p[a]);
        }
        setp(F, g0, 0.5 * (F.r[s1] + F.r[t1]), 0.5 * (F.u[s1] + F.u[t1]), 0.5 * (F.v[s1] + F.v[t1]),
             0.5 * (F.w[s1] + F.w[t1]), 0.5 * (F.p[s1] + F.p[t1]));
        for (int m = 0; m < P.turb; m++) {
            // turblence.f90:2308-2319 BC_symmetry_turbulence (first ghost) + dif_average
            double *q = tvar(F, m);
            const double qt = q[t1];
            q[s1] = qt;
            q[g0] = 0.5 * (qt + qt);
        }
        break;
    }
    default:
        break;
    }
}

__global__ void k_bc_region(DRegion g, DBlock b, DBlock bt, DFields F, DParams P)
{
    if (g.serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int t = 0; t < g.npts; t++) {
            int i, j, k;
            region_point(g, t, i, j, k);
            bc_point(g, b, bt, F, P, i, j, k);
        }
        return;
    }
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    bc_point(g, b, bt, F, P, i, j, k);
}

// One dependency level of a region sequence in one launch.  Workgroup w runs
// points work[w].y .. work[w].y + blockDim.x - 1 of region work[w].x, or, with
// work[w].y < 0, every point of that region in order on one thread (a region
// that reads its own writes).  The host puts regions that touch no node
// another region of the level writes on one level.
__global__ void k_bc_batch(const DRegion *regs, const int2 *work, const DBlock *blocks, DFields F, DParams P)
{
    const int2 w = work[blockIdx.x];
    const DRegion g = regs[w.x];
    const DBlock b = blocks[g.nb];
    const DBlock bt = blocks[g.bctype < 0 ? g.nbt : g.nb];
    int i, j, k;
    if (w.y < 0) {
        if (threadIdx.x != 0) return;
        for (int t = 0; t < g.npts; t++) {
            region_point(g, t, i, j, k);
            bc_point(g, b, bt, F, P, i, j, k);
        }
        return;
    }
    const int t = w.y + (int)threadIdx.x;
    if (t >= g.npts) return;
    region_point(g, t, i, j, k);
    bc_point(g, b, bt, F, P, i, j, k);
}

// --------------------------------------------------------- init stage
// set_dq_to_0 (PreProcess.f90:1189-1205) on the residual; with nvis == 1 also
// get_c_t_only's temperature (RHS.f90:307-409) and compute_visl_ns
// (GasModel.f90:219-235).
__global__ void k_init_interior(BlockGrid G, DFields F, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) F.rhs[m * F.total + g] = 0.0;
    if (P.nvis == 1) {
        double moo2 = P.moo * P.moo;
        double a2 = P.gama * F.p[g] / F.r[g];
        double tm = moo2 * a2;
        F.t[g] = tm;
        F.visl[g] = tm * sqrt(tm) * (1.0 + P.visc) / (tm + P.visc);
    }
}

__global__ void k_init_ghost_t(BlockGrid G, DFields F, DParams P, int dir)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    int na = dir == 0 ? b.nj : b.ni, nb = dir == 2 ? b.nj : b.nk;
    int nd = dir == 0 ? b.ni : (dir == 1 ? b.nj : b.nk);
    if (t >= 2L * na * nb) return;
    int side = (int)(t % 2);
    long r = t / 2;
    int ia = (int)(r % na) + 1, ib = (int)(r / na) + 1;
    double moo2 = P.moo * P.moo, gama = P.gama;
    int l1 = side == 0 ? 0 : nd + 1, l2 = side == 0 ? -1 : nd + 2, l3 = side == 0 ? -2 : nd + 3;
    auto at = [&](int l) {
        int i = dir == 0 ? l : ia, j = dir == 1 ? l : (dir == 0 ? ia : ib), k = dir == 2 ? l : ib;
        return gidx(b, i, j, k);
    };
    long q1 = at(l1), q2 = at(l2), q3 = at(l3);
    double a2 = gama * F.p[q1] / F.r[q1];
    F.t[q1] = moo2 * a2;
    a2 = gama * F.p[q2] / F.r[q2];
    F.t[q2] = moo2 * a2;
    if (dir != 0) a2 = gama * F.p[q3] / fabs(F.r[q3]);
    F.t[q3] = moo2 * a2;
}

// ---------------------------------------------------- spectral radius, dt
// Convect.f90:1338-1461 spectinv (interior) + LHS.f90:1343-1377 localdt0;
// VIS: also the viscous radii of spectvisl (laminar, LHS.f90:243-343) /
// spectvist (mu + mu_t, LHS.f90:141-241) of the node, formed here rather than
// read back from a kernel of their own; interior nodes only (ghost-layer radii
// only ever multiply a zero dq)
template <bool VIS>
__global__ void k_timestep(BlockGrid G, DFields F, DParams P)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    double um = F.u[g], vm = F.v[g], wm = F.w[g];
    double cm = sqrt(P.gama * F.p[g] / F.r[g]);
    double kx = F.kcx[g], ky = F.kcy[g], kz = F.kcz[g];
    double ex = F.etx[g], ey = F.ety[g], ez = F.etz[g];
    double cx = F.ctx[g], cy = F.cty[g], cz = F.ctz[g];
    double cita = kx * um + ky * vm + kz * wm + 0.0;
    double citb = ex * um + ey * vm + ez * wm + 0.0;
    double citc = cx * um + cy * vm + cz * wm + 0.0;
    double sra = fabs(cita) + cm * sqrt(kx * kx + ky * ky + kz * kz);
    double srb = fabs(citb) + cm * sqrt(ex * ex + ey * ey + ez * ez);
    double src = fabs(citc) + cm * sqrt(cx * cx + cy * cy + cz * cz);
    F.sra[g] = sra;
    F.srb[g] = srb;
    F.src[g] = src;
    double sva, svb, svc;
    if constexpr (VIS) {
        const double rm_vol = F.r[g] * F.vol[g];
        const double vis = P.turb ? F.visl[g] + F.vist[g] : F.visl[g];
        const double coef = 2.0 * vis / (P.reynolds * rm_vol + SMALL);
        const double coef1 = P.csrv * coef;
        sva = coef1 * (kx * kx + ky * ky + kz * kz);
        svb = coef1 * (ex * ex + ey * ey + ez * ez);
        svc = coef1 * (cx * cx + cy * cy + cz * cz);
        F.srva[g] = sva;
        F.srvb[g] = svb;
        F.srvc[g] = svc;
    } else {
        sva = F.srva[g];
        svb = F.srvb[g];
        svc = F.srvc[g];
    }
    double rv = sva + svb + svc;
    double rabc = sra + srb + src + rv;
    double dt;
    if (P.timedt_rate > SMALL) {
        double a = P.cfl / rabc, c = P.timedt_rate / F.vol[g];
        dt = a < c ? a : c;
    } else {
        dt = P.cfl / rabc;
    }
    F.dtdt[g] = dt;
    if (F.coed) {
        // LHS.f90:602-645 getcoed, the same for all four sweeps of the step
        double rad_ns = sra + srb + src;
        if (P.nvis >= 1) rad_ns = rad_ns + sva + svb + svc;
        double ct = 1.0 / dt;
        F.coed[g] = 1.0 / (ct + 1.0 * 1.0 * rad_ns);
    }
}

// ------------------------------------------------- interface averaging
// parallel.f90:3028-3036 / 3174-3188: pre-averaging values of the multi-points
__global__ void k_npp_gather(const int *pts, int npts, const DBlock *blocks, DFields F, double *buf,
                             int nvar, int use_vol)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    const int *p = pts + 4 * t;
    const DBlock b = blocks[p[0]];
    long q = gidx(b, p[1], p[2], p[3]);
    if (use_vol) {
        double vol_p = F.vol[q];
        for (int m = 0; m < 5; m++) buf[nvar * t + m] = F.rhs[m * F.total + q] / vol_p;
    } else {
        double vol_p = 1.0;
        buf[nvar * t + 0] = F.r[q] / vol_p;
        buf[nvar * t + 1] = F.u[q] / vol_p;
        buf[nvar * t + 2] = F.v[q] / vol_p;
        buf[nvar * t + 3] = F.w[q] / vol_p;
        buf[nvar * t + 4] = F.p[q] / vol_p;
        buf[nvar * t + 5] = F.t[q] / vol_p;
    }
}

// parallel.f90:3060-3120 boundary_match_dq_2pm / :3217-3280 boundary_match_pv_2pm
__device__ inline void avg2pm_point(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F,
                                    int pv, int i, int j, int k)
{
    int it, jt, kt;
    region_image(g, i, j, k, it, jt, kt);
    long qs = gidx(b, i, j, k), qt = gidx(bt, it, jt, kt);
    if (!pv) {
        double vol_p = F.vol[qs] / F.vol[qt];
#pragma unroll
        for (int m = 0; m < 5; m++)
            F.rhs[m * F.total + qs] = 0.5 * (F.rhs[m * F.total + qt] * vol_p + F.rhs[m * F.total + qs]);
    } else {
        F.r[qs] = 0.5 * (F.r[qt] + F.r[qs]);
        F.u[qs] = 0.5 * (F.u[qt] + F.u[qs]);
        F.v[qs] = 0.5 * (F.v[qt] + F.v[qs]);
        F.w[qs] = 0.5 * (F.w[qt] + F.w[qs]);
        F.p[qs] = 0.5 * (F.p[qt] + F.p[qs]);
        F.t[qs] = 0.5 * (F.t[qt] + F.t[qs]);
    }
}

__global__ void k_avg2pm(DRegion g, DBlock b, DBlock bt, DFields F, int pv)
{
    if (g.serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int k = g.s_lo[2]; k <= g.s_hi[2]; k++)
            for (int j = g.s_lo[1]; j <= g.s_hi[1]; j++)
                for (int i = g.s_lo[0]; i <= g.s_hi[0]; i++) avg2pm_point(g, b, bt, F, pv, i, j, k);
        return;
    }
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    avg2pm_point(g, b, bt, F, pv, i, j, k);
}

__global__ void k_avg2pm_batch(const DRegion *regs, const int2 *work, const DBlock *blocks, DFields F, int pv)
{
    const int2 w = work[blockIdx.x];
    const DRegion g = regs[w.x];
    const DBlock b = blocks[g.nb];
    const DBlock bt = blocks[g.nbt];
    int i, j, k;
    if (w.y < 0) {
        if (threadIdx.x != 0) return;
        for (int t = 0; t < g.npts; t++) {
            region_point(g, t, i, j, k);
            avg2pm_point(g, b, bt, F, pv, i, j, k);
        }
        return;
    }
    const int t = w.y + (int)threadIdx.x;
    if (t >= g.npts) return;
    region_point(g, t, i, j, k);
    avg2pm_point(g, b, bt, F, pv, i, j, k);
}

// set_nonconnect_boundary_dq_0 (CIC.f90:2-17) on every physical face at once:
// the same value everywhere, so the order of the regions does not matter.
template <typename T>
__global__ void k_dq_zero_batch(const DRegion *regs, const int2 *work, const DBlock *blocks, T *arr, long total)
{
    const int2 w = work[blockIdx.x];
    const DRegion g = regs[w.x];
    const int t = w.y + (int)threadIdx.x;
    if (t >= g.npts) return;
    int i, j, k;
    region_point(g, t, i, j, k);
    const long q = gidx(blocks[g.nb], i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) arr[m * total + q] = T(0);
}

// parallel.f90:3122-3164 boundary_match_dq_3pp / :3282-3326 boundary_match_pv_3pp
__global__ void k_npp_assign(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok,
                             int ngroups, const double *buf, const DBlock *blocks, DFields F, int nvar,
                             int use_vol)
{
    int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups) return;
    int s0 = grp_start[gi], n = grp_n[gi];
    double d[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < n; s++)
        for (int m = 0; m < nvar; m++) d[m] = d[m] + buf[nvar * (s0 + s) + m];
    for (int m = 0; m < nvar; m++) d[m] = d[m] / (n * 1.0);
    for (int s = 0; s < n; s++) {
        if (!write_ok[s0 + s]) continue;
        const int *p = pts + 4 * (s0 + s);
        const DBlock b = blocks[p[0]];
        long q = gidx(b, p[1], p[2], p[3]);
        if (use_vol) {
            double vol_p = F.vol[q];
            for (int m = 0; m < 5; m++) F.rhs[m * F.total + q] = d[m] * vol_p;
        } else {
            F.r[q] = d[0] * 1.0; F.u[q] = d[1] * 1.0; F.v[q] = d[2] * 1.0;
            F.w[q] = d[3] * 1.0; F.p[q] = d[4] * 1.0; F.t[q] = d[5] * 1.0;
        }
    }
}

// -------------------------------------------------------------- update
// Sub2.f90:603-796 update_limited.  q is prim_to_q(prim) of the step's start
// (get_q_only, RHS.f90:75-100), recomputed here instead of stored.  New
// primitives go to the other buffer; nodes outside the limits are listed for
// the ordered repair below.  All blocks in one launch; block nb lists its
// nodes in flag_list[nb * cap ...] and counts them in nflag[nb].
__global__ void k_update(BlockGrid G, DFields F, DParams P, int *flag_list, unsigned long long *nflag, int cap)
{
    DBlock b;
    long t;
    grid_thread(G, b, t);
    const int nb = G.wg[blockIdx.x].x;
    long n = (long)b.ni * b.nj * b.nk;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
    const double gama = P.gama, gamam1 = gama - 1.0;
    double r0 = F.r[g], u0 = F.u[g], v0 = F.v[g], w0 = F.w[g], p0 = F.p[g];
    double q[5];
    q[0] = r0; q[1] = r0 * u0; q[2] = r0 * v0; q[3] = r0 * w0; q[4] = getem(r0, u0, v0, w0, p0, gama);
    double d[5];
#pragma unroll
    for (int m = 0; m < 5; m++) d[m] = F.dq32 ? (double)F.dq32[m * F.total + g] : F.dq[m * F.total + g];
    double qt[5];
#pragma unroll
    for (int m = 0; m < 5; m++) qt[m] = q[m] + d[m];
    double rm = qt[0], rm1 = 1.0 / rm;
    double um = qt[1] * rm1, vm = qt[2] * rm1, wm = qt[3] * rm1, em = qt[4];
    double pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
    double dp = -0.3 + fabs((pm - p0) / p0);
    double falpha = 1. / (1. + 2. * (0.0 > dp ? 0.0 : dp));
#pragma unroll
    for (int m = 0; m < 5; m++) qt[m] = q[m] + d[m] * falpha;
    rm = qt[0]; rm1 = 1.0 / rm;
    um = qt[1] * rm1; vm = qt[2] * rm1; wm = qt[3] * rm1; em = qt[4];
    pm = gamam1 * (em - 0.5 * rm * (um * um + vm * vm + wm * wm));
    F.rn[g] = rm; F.un[g] = um; F.vn[g] = vm; F.wn[g] = wm; F.pn[g] = pm;
    if (pm <= PMIN_LIMIT || rm <= RMIN_LIMIT || pm >= PMAX_LIMIT || rm >= RMAX_LIMIT) {
        unsigned long long slot = atomicAdd(&nflag[nb], 1ull);
        if (slot < (unsigned long long)cap) flag_list[(long)nb * cap + slot] = (int)t;
    }
}

// Sub2.f90:661-767: a flagged node takes the mean of its six neighbours, those
// earlier in lexicographic order already updated (new buffer), later ones and
// ghosts as before the update (old buffer).  One thread per block, in order;
// the repair reads only its own block, so the blocks run side by side.
__global__ void k_repair(const DBlock *blocks, DFields F, DParams P, int *flags, const unsigned long long *nflag,
                         int cap)
{
    if (threadIdx.x != 0) return;
    const DBlock b = blocks[blockIdx.x];
    int *flag_list = flags + (long)blockIdx.x * cap;
    unsigned long long nf = nflag[blockIdx.x];
    if (nf == 0) return;
    int n = (int)(nf < (unsigned long long)cap ? nf : (unsigned long long)cap);
    for (int a = 1; a < n; a++) {          // insertion sort: lexicographic = linear order
        int v = flag_list[a], c = a - 1;
        while (c >= 0 && flag_list[c] > v) { flag_list[c + 1] = flag_list[c]; c--; }
        flag_list[c + 1] = v;
    }
    for (int a = 0; a < n; a++) {
        long t = flag_list[a];
        int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
        double rm = 0.0, um = 0.0, vm = 0.0, wm = 0.0, pm = 0.0;
        int np = 0;
        for (int ii = -1; ii <= 1; ii++)
            for (int jj = -1; jj <= 1; jj++)
                for (int kk = -1; kk <= 1; kk++) {
                    if (abs(ii) + abs(jj) + abs(kk) != 1) continue;
                    np++;
                    int i0 = i + ii, j0 = j + jj, k0 = k + kk;
                    long g2 = gidx(b, i0, j0, k0);
                    bool interior = i0 >= 1 && i0 <= b.ni && j0 >= 1 && j0 <= b.nj && k0 >= 1 && k0 <= b.nk;
                    bool earlier = (ii + jj + kk) < 0;
                    bool use_new = interior && earlier;
                    rm = rm + (use_new ? F.rn[g2] : F.r[g2]);
                    um = um + (use_new ? F.un[g2] : F.u[g2]);
                    vm = vm + (use_new ? F.vn[g2] : F.v[g2]);
                    wm = wm + (use_new ? F.wn[g2] : F.w[g2]);
                    pm = pm + (use_new ? F.pn[g2] : F.p[g2]);
                }
        double rnp = (double)np;
        rm = rm / rnp; um = um / rnp; vm = vm / rnp; wm = wm / rnp; pm = pm / rnp;
        long g = gidx(b, i, j, k);
        F.rn[g] = rm; F.un[g] = um; F.vn[g] = vm; F.wn[g] = wm; F.pn[g] = pm;
    }
    atomicAdd(&F.counters[1], nf);
    // Sub2.f90:781-785: too many abnormal nodes stops the run
    if (nf > (unsigned long long)(((long)b.ni * b.nj * b.nk) / 64) || nf >= 500ull) F.counters[4] = 1ull;
}

// ------------------------------------------------------------ residual
// PostProcess.f90:1240-1293: sum and max of (|dq|/(dtdt*vol))^2 over nodes 2..n-1.
constexpr int RES_THREADS = 256;

// workgroups of the residual monitor per block: about 8 interior nodes per
// thread, at most 1024 workgroups
__host__ __device__ inline int residual_parts_dev(const DBlock &b)
{
    long n = (long)(b.ni - 2) * (b.nj - 2) * (b.nk - 2);
    long p = (n + RES_THREADS * 8 - 1) / (RES_THREADS * 8);
    return (int)(p < 1 ? 1 : (p > 1024 ? 1024 : p));
}

// The same monitor for every block in one launch over the GRID_RESID list:
// block b owns workgroups wg0[b] .. wg0[b+1]-1 (in order, never permuted), and
// each of them strides over the block's interior nodes 2..n-1.
__global__ void __launch_bounds__(RES_THREADS) k_residual_all(BlockGrid G, DFields F, double *partial)
{
    __shared__ double ssum[RES_THREADS];
    __shared__ double smax[RES_THREADS];
    __shared__ long sarg[RES_THREADS];
    const int2 w = G.wg[blockIdx.x];
    const DBlock b = G.blocks[w.x];
    const int mi = b.ni - 2, mj = b.nj - 2, mk = b.nk - 2;
    const long n = (long)mi * mj * mk;
    const long stride = (long)residual_parts_dev(b) * RES_THREADS;
    double sum = 0.0, mx = 0.0;
    long arg = -1;
    for (long t = (long)w.y * RES_THREADS + threadIdx.x; t < n; t += stride) {
        int i = (int)(t % mi) + 2, j = (int)((t / mi) % mj) + 2, k = (int)(t / ((long)mi * mj)) + 2;
        long g = gidx(b, i, j, k);
        double dtime1 = 1.0 / (F.dtdt[g] * F.vol[g]);
        for (int m = 0; m < 5; m++) {
            double dqm = F.dq32 ? (double)F.dq32[m * F.total + g] : F.dq[m * F.total + g];
            double dresm = fabs(dqm) * dtime1;
            if (dresm > mx) { mx = dresm; arg = t * 5 + m; }
            sum = sum + dresm * dresm;
        }
    }
    ssum[threadIdx.x] = sum; smax[threadIdx.x] = mx; sarg[threadIdx.x] = arg;
    __syncthreads();
    for (int h = RES_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            ssum[threadIdx.x] += ssum[threadIdx.x + h];
            double om = smax[threadIdx.x + h];
            long oa = sarg[threadIdx.x + h];
            if (om > smax[threadIdx.x] || (om == smax[threadIdx.x] && oa >= 0 && (sarg[threadIdx.x] < 0 || oa < sarg[threadIdx.x]))) {
                smax[threadIdx.x] = om; sarg[threadIdx.x] = oa;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = ssum[0];
        partial[3 * blockIdx.x + 1] = smax[0];
        partial[3 * blockIdx.x + 2] = (double)sarg[0];
    }
}

// One workgroup per block: its partials -> resblk[3b..3b+2] = (sum, max,
// argmax), in a fixed order (strided per thread, then a tree), the largest
// maximum with the smallest node index winning ties as in the serial loop.
__global__ void __launch_bounds__(RES_THREADS) k_residual_blocks(const double *partial, const int *wg0, double *resblk)
{
    __shared__ double ssum[RES_THREADS];
    __shared__ double smax[RES_THREADS];
    __shared__ long sarg[RES_THREADS];
    const int b = blockIdx.x;
    double s = 0.0, mx = 0.0;
    long arg = -1;
    for (int p = wg0[b] + threadIdx.x; p < wg0[b + 1]; p += RES_THREADS) {
        s += partial[3 * p];
        const double om = partial[3 * p + 1];
        const long oa = (long)partial[3 * p + 2];
        if (om > mx || (om == mx && oa >= 0 && (arg < 0 || oa < arg))) { mx = om; arg = oa; }
    }
    ssum[threadIdx.x] = s; smax[threadIdx.x] = mx; sarg[threadIdx.x] = arg;
    __syncthreads();
    for (int h = RES_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            ssum[threadIdx.x] += ssum[threadIdx.x + h];
            double om = smax[threadIdx.x + h];
            long oa = sarg[threadIdx.x + h];
            if (om > smax[threadIdx.x] || (om == smax[threadIdx.x] && oa >= 0 && (sarg[threadIdx.x] < 0 || oa < sarg[threadIdx.x]))) {
                smax[threadIdx.x] = om; sarg[threadIdx.x] = oa;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { resblk[3 * b] = ssum[0]; resblk[3 * b + 1] = smax[0]; resblk[3 * b + 2] = (double)sarg[0]; }
}

// residual_curve's combination over the blocks (PostProcess.f90:1280-1292):
// the sums added in block order, the first block with the largest maximum
// names the location.  out = rms, max, i, j, k, m, nb (1-based), step number.
// With hist != nullptr the entry is also appended to the history ring
// hist[cap][8] at position *count % cap, and *count is advanced.
__global__ void k_residual_hist(const double *resblk, const DBlock *blocks, int nblocks, double *out,
                                double *hist, unsigned long long *count, int cap)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sum = 0.0, mx = 0.0;
    long npt = 0;
    double l[5] = {2.0, 2.0, 2.0, 1.0, 1.0};
    for (int nb = 0; nb < nblocks; nb++) {
        const DBlock b = blocks[nb];
        sum = sum + resblk[3 * nb];
        npt += (long)(b.ni - 2) * (b.nj - 2) * (b.nk - 2) * 5;
        if (mx < resblk[3 * nb + 1]) {
            mx = resblk[3 * nb + 1];
            const long a = (long)resblk[3 * nb + 2];
            const long t = a / 5;
            const int mi = b.ni - 2, mj = b.nj - 2;
            l[0] = (double)(t % mi + 2);
            l[1] = (double)((t / mi) % mj + 2);
            l[2] = (double)(t / ((long)mi * mj) + 2);
            l[3] = (double)(a % 5 + 1);
            l[4] = (double)(nb + 1);
        }
    }
    double e[8];
    e[0] = sqrt(sum / (double)npt);
    e[1] = mx;
    for (int m = 0; m < 5; m++) e[2 + m] = l[m];
    const unsigned long long n = hist ? *count : 0ull;
    e[7] = (double)(n + 1);
    for (int m = 0; m < 8; m++) out[m] = e[m];
    if (hist) {
        double *h = hist + 8 * (long)(n % (unsigned long long)cap);
        for (int m = 0; m < 8; m++) h[m] = e[m];
        *count = n + 1;
    }
}

__global__ void k_zero_interior5(DBlock b, double *arr, long total)
{
    long n = (long)b.ni * b.nj * b.nk;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    long g = gidx(b, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) arr[m * total + g] = 0.0;
}

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

int launch_bc_region(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, hipStream_t s)
{
    const DBlock &b = blocks[g.nb];
    const DBlock &bt = blocks[g.bctype < 0 ? g.nbt : g.nb];
    if (g.serial) hipLaunchKernelGGL(k_bc_region, dim3(1), dim3(64), 0, s, g, b, bt, F, P);
    else hipLaunchKernelGGL(k_bc_region, grid1(g.npts, 128), dim3(128), 0, s, g, b, bt, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_init_stage(const BlockGrid &nodes, const BlockGrid *ghost_t, const DFields &F, const DParams &P,
                      hipStream_t s)
{
    if (nodes.nwg) hipLaunchKernelGGL(k_init_interior, dim3(nodes.nwg), dim3(256), 0, s, nodes, F, P);
    if (P.nvis == 1) {
        for (int dir = 0; dir < 3; dir++)
            if (ghost_t[dir].nwg)
                hipLaunchKernelGGL(k_init_ghost_t, dim3(ghost_t[dir].nwg), dim3(128), 0, s, ghost_t[dir], F, P, dir);
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_timestep(const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s)
{
    const bool vis = P.nvis == 1 && P.csrv > 0.0;
    if (G.nwg && vis) hipLaunchKernelGGL(k_timestep<true>, dim3(G.nwg), dim3(256), 0, s, G, F, P);
    if (G.nwg && !vis) hipLaunchKernelGGL(k_timestep<false>, dim3(G.nwg), dim3(256), 0, s, G, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long rhs_inv_threads(int dir, const DBlock &b);
long rhs_vis_threads(int dir, const DBlock &b);
long vis_face_threads(int dir, const DBlock &b);
long vis_row_threads(int dir, const DBlock &b);

long grid_threads(int kind, const DBlock &b)
{
    const long n = (long)b.ni * b.nj * b.nk;
    if (kind == GRID_NODES) return n;
    if (kind >= GRID_GHOST_T && kind < GRID_GHOST_T + 3) {
        const int dir = kind - GRID_GHOST_T;
        const long na = dir == 0 ? b.nj : b.ni, nb = dir == 2 ? b.nj : b.nk;
        return 2 * na * nb;
    }
    if (kind >= GRID_RHS_VIS && kind < GRID_RHS_VIS + 3) return rhs_vis_threads(kind - GRID_RHS_VIS, b);
    if (kind >= GRID_RHS_INV && kind < GRID_RHS_INV + 3) return rhs_inv_threads(kind - GRID_RHS_INV, b);
    if (kind == GRID_SA_RHS) return b.ni > 2 && b.nj > 2 && b.nk > 2 ? (long)(b.ni - 2) * (b.nj - 2) * (b.nk - 2) : 0;
    if (kind == GRID_SA_UPD) return (long)(b.ni + 1) * (b.nj + 1) * (b.nk + 1);
    if (kind == GRID_TRANS) return transpose_tiles(b) * 256;
    if (kind == GRID_SKEW) return skew_stage_wgs(b) * 256;
    if (kind == GRID_RESID) return (long)residual_parts_dev(b) * RES_THREADS;
    // b is the transposed view here: its j pencils are the original i pencils
    if (kind == GRID_RHS_INV_T) return rhs_inv_threads(1, b);
    if (kind == GRID_RHS_VIS_T) return rhs_vis_threads(1, b);
    // d = 0 is the i direction as the transposed view's j pencils (b is that view)
    if (kind >= GRID_VFACE && kind < GRID_VFACE + 3) return vis_face_threads(kind == GRID_VFACE ? 1 : kind - GRID_VFACE, b);
    if (kind >= GRID_VROW && kind < GRID_VROW + 3) return vis_row_threads(kind == GRID_VROW ? 1 : kind - GRID_VROW, b);
    return 0;
}

int grid_block_size(int kind)
{
    // GRID_SKEW and GRID_RESID too
    return (kind == GRID_NODES || kind == GRID_SA_UPD || kind == GRID_TRANS || kind >= GRID_VFACE) ? 256 : 128;
}

int launch_zero_rhs(const DBlock &b, const DFields &F, hipStream_t s)
{
    long n = (long)b.ni * b.nj * b.nk;
    hipLaunchKernelGGL(k_zero_interior5, grid1(n, 256), dim3(256), 0, s, b, F.rhs, F.total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_bc_level(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, const DFields &F,
                    const DParams &P, hipStream_t s)
{
    if (nwork > 0) hipLaunchKernelGGL(k_bc_batch, dim3(nwork), dim3(REGION_CHUNK), 0, s, regs, work, blocks, F, P);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_avg2pm_level(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, const DFields &F,
                        int pv, hipStream_t s)
{
    if (nwork > 0)
        hipLaunchKernelGGL(k_avg2pm_batch, dim3(nwork), dim3(REGION_CHUNK), 0, s, regs, work, blocks, F, pv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_dq_zero_faces(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, double *arr,
                         float *arr32, long total, hipStream_t s)
{
    if (nwork <= 0) return 0;
    if (arr32) hipLaunchKernelGGL(k_dq_zero_batch<float>, dim3(nwork), dim3(REGION_CHUNK), 0, s, regs, work, blocks,
                                  arr32, total);
    else hipLaunchKernelGGL(k_dq_zero_batch<double>, dim3(nwork), dim3(REGION_CHUNK), 0, s, regs, work, blocks,
                            arr, total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_npp_gather(const int *pts, int npts, const DBlock *dblocks, const DFields &F, double *buf, int nvar,
                      int use_vol, hipStream_t s)
{
    if (npts == 0) return 0;
    hipLaunchKernelGGL(k_npp_gather, grid1(npts, 128), dim3(128), 0, s, pts, npts, dblocks, F, buf, nvar, use_vol);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_avg2pm(const DRegion &g, const DBlock *blocks, const DFields &F, int pv, hipStream_t s)
{
    const DBlock &b = blocks[g.nb], &bt = blocks[g.nbt];
    if (g.serial) hipLaunchKernelGGL(k_avg2pm, dim3(1), dim3(64), 0, s, g, b, bt, F, pv);
    else hipLaunchKernelGGL(k_avg2pm, grid1(g.npts, 128), dim3(128), 0, s, g, b, bt, F, pv);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_npp_assign(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok, int ngroups,
                      const double *buf, const DBlock *dblocks, const DFields &F, int nvar, int use_vol,
                      hipStream_t s)
{
    if (ngroups == 0) return 0;
    hipLaunchKernelGGL(k_npp_assign, grid1(ngroups, 64), dim3(64), 0, s, grp_start, grp_n, pts, write_ok,
                       ngroups, buf, dblocks, F, nvar, use_vol);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_update(const BlockGrid &nodes, const DFields &F, const DParams &P, int *flag_list,
                  unsigned long long *nflag, int cap, hipStream_t s)
{
    if (nodes.nwg) hipLaunchKernelGGL(k_update, dim3(nodes.nwg), dim3(256), 0, s, nodes, F, P, flag_list, nflag, cap);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_repair(const DBlock *blocks, int nblocks, const DFields &F, const DParams &P, int *flag_list,
                  const unsigned long long *nflag, int cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_repair, dim3(nblocks), dim3(64), 0, s, blocks, F, P, flag_list, nflag, cap);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int residual_parts(const DBlock &b) { return residual_parts_dev(b); }

int launch_residual_all(const BlockGrid &G, const DFields &F, double *partial, const int *wg0, int nblocks,
                        double *resblk, const DBlock *blocks, double *out, double *hist, unsigned long long *count,
                        int cap, hipStream_t s)
{
    if (G.nwg <= 0 || nblocks <= 0) return -3;
    hipLaunchKernelGGL(k_residual_all, dim3(G.nwg), dim3(RES_THREADS), 0, s, G, F, partial);
    if (hipGetLastError() != hipSuccess) return -3;
    hipLaunchKernelGGL(k_residual_blocks, dim3(nblocks), dim3(RES_THREADS), 0, s, partial, wg0, resblk);
    if (hipGetLastError() != hipSuccess) return -3;
    hipLaunchKernelGGL(k_residual_hist, dim3(1), dim3(64), 0, s, resblk, blocks, nblocks, out, hist, count, cap);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
