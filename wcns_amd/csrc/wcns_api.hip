// wcns_api.hip -- the C-ABI (include/wcns_amd.h): context, host-side setup
// (boundary topology, interface maps, multi-point groups) and the ordering of
// the per-step kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/wcns_amd.h"
#include "wcns_kernels.hpp"

using namespace wcns;

namespace {

struct HRegion {
    int bctype;
    int raw_s[6], raw_t[6];
    int nbt, ibcwin;         // 1-based as in the .inp file
    DRegion d;               // analysed form
};

struct HBlock {
    int ni, nj, nk;
    std::vector<double> x, y, z;
    std::vector<HRegion> regs;
    std::vector<int> bcorder;  // region indices in -1,4,5,6,2,3 order
    DBlock db;
    long g3;
};

#define CK(call)                                                                    \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return WCNS_E_HIP;                                                      \
        }                                                                           \
    } while (0)

}  // namespace

// Scratch of wcns_residual_noalloc: a file-scope device array, so that the
// entry point needs no allocation.  Every GPU holds its own copy until the
// process ends; host code reaches it only through the device address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch_ptr),
// never through the array's name.  Layout: 3 partials per residual workgroup
// (at most NOALLOC_WGS), then 3 per block (at most NOALLOC_BLOCKS), then the
// 8-double result.
constexpr int NOALLOC_WGS = 65536;
constexpr int NOALLOC_BLOCKS = 4096;
constexpr size_t NOALLOC_DOUBLES = 3 * (size_t)NOALLOC_WGS + 3 * (size_t)NOALLOC_BLOCKS + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The scratch array's address on `device` (which must be current), looked up
// once per device; nullptr if the lookup fails.
double *noalloc_scratch_ptr(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        noalloc_ptr[device] = (double *)p;
    }
    return noalloc_ptr[device];
}
}  // namespace

// accelerator complex dies of an MI355X, each with its own L2; workgroup ids
// are dispatched to them in turn
constexpr long XCDS = 8;
constexpr long XCD_SLICE_MIN = 32768;

struct wcns_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *noalloc_scratch = nullptr;      // this device's copy of `scratch` (device address)
    wcns_params prm{};
    DParams dp{};
    std::vector<HBlock> blocks;
    bool setup_done = false;
    int lhs_prec = WCNS_LHS_FP64;
    std::string err;

    long total = 0;
    double *prim[2][5] = {{nullptr}};
    int cur = 0;
    double *t = nullptr, *metric[10] = {nullptr}, *rhs = nullptr, *dq = nullptr;
    double *dtdt = nullptr, *sr[3] = {nullptr}, *srv[3] = {nullptr}, *visl = nullptr;
    double *dn = nullptr;                   // 12 node derivatives (viscous); SA gradients before them
    double *qke = nullptr, *vist = nullptr, *dist = nullptr;   // turbulence-model fields
    double *qkw = nullptr;                  // SST omega
    SaWork sa{};                            // turbulence scratch (grad aliases dn)
    double *sa_pool = nullptr;              // viseq, un[3], dqke, spec (+ SST arrays)
    double *coed = nullptr;                 // LU-SGS diagonal per node
    float *dq32 = nullptr;                  // dQ of the FP32 implicit operator
    void *skew[2] = {nullptr, nullptr};     // plane-major sweep operands, FP64 / FP32 implicit operator
    std::vector<long> skoff;                // per block offset in each plane-major array
    long *d_skoff = nullptr;                // the same, on the device
    long sktot = 0;
    SweepPlan plan{};                       // batched pipelined sweeps over all blocks
    long spin_limit = 2000000;              // re-reads before a sweep's wait gives up (WCNS_SPIN_LIMIT)
    bool sgs_forced = false;                // WCNS_SGS_PAIR fixes plan.sgs_paired; the autotuner keeps it
    std::vector<std::pair<int, int>> tile_shape;   // per block (TJ, TK)
    SweepBlock *d_sblocks = nullptr;
    int2 *d_work = nullptr;                 // forward list, then backward list
    unsigned long long *d_xch = nullptr;
    int *d_sweep_err = nullptr;
    bool tiled_sweeps = true;
    unsigned long long *counters = nullptr;
    DBlock *d_blocks = nullptr;
    BlockGrid grid[GRID_N] = {};            // block work-lists of the batched launches
    int2 *d_grid = nullptr;
    // i-direction residuals on i<->j transposed copies (WCNS_IDIR=direct: in place)
    bool idir_t = true;
    DBlock *d_blocks_t = nullptr;           // transposed views
    double *tpool = nullptr;                // the copies below, total elements each
    double *mt[10] = {nullptr};             // kcx kcy kcz etx ety etz ctx cty ctz vol (set-up)
    double *pt[5] = {nullptr}, *t_t = nullptr, *visl_t = nullptr, *vist_t = nullptr;
    double *dn_t = nullptr;                 // 12 slots; 0..3 <- j derivatives, 8..11 <- k derivatives
    bool pt_fresh = false;                  // pt[] holds the current primitives (no write since the copy)
    int ncu = 256;                          // compute units of the device
    int sa_staged = -1;                     // WCNS_SA_STAGED=0|1 fixes the turbulence sweeps' operand path
    double *df_t = nullptr;                 // 5: the i-direction contribution, transposed
    // schemes 45 / 47: pencil-end ghost metrics and each block's first record;
    // flux 4: the pressure ratios (3 x total)
    double *d_gmet = nullptr;
    long *d_gmoff = nullptr;
    double *d_rpres = nullptr;
    double *fvb = nullptr;                  // 5: viscous half-node fluxes of one direction
    int *d_flag = nullptr;                  // flag_cap nodes per block for the ordered repair
    int flag_cap = 512;
    unsigned long long *d_nflag = nullptr;  // flagged-node count per block
    // surface stages as dependency levels: regions (BC sequence, interfaces,
    // physical faces) and per-level work lists; WCNS_SURFACE=sequence launches
    // the BC and interface regions one at a time instead
    bool surf_levels = true;
    DRegion *d_regs = nullptr;
    int2 *d_rwork = nullptr;
    std::vector<int> bc_lvl, if_lvl;        // work offsets of each level, then the end
    int zero_w0 = 0, zero_nw = 0;           // the physical faces' work entries
    // residual monitor: 3 partials per GRID_RESID workgroup, each block's
    // first workgroup (nblocks + 1), per-block (sum, max, argmax), the latest
    // entry (8 doubles) and the on-device history ring of whole steps
    double *d_partial = nullptr;
    int *d_res_wg0 = nullptr;
    int res_nwg = 0;                        // residual workgroups of all blocks
    double *d_resblk = nullptr;
    double *d_res_out = nullptr;
    double *d_hist = nullptr;
    unsigned long long *d_hist_n = nullptr;
    int hist_cap = 8192;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // two-step graphs, by starting primitive buffer
    int graph_prec = -1;                            // implicit-operator precision they were captured with

    std::vector<DRegion> iface;              // interfaces in (nb, nr) order
    std::vector<DRegion> phys;               // bctype > 0 regions
    struct Wall { DRegion d; double *xyz; };  // no-slip walls with their face-node coordinates
    std::vector<Wall> walls;
    double *d_force_pn = nullptr, *d_force_part = nullptr, *d_force_out = nullptr;
    double *d_force_last = nullptr;         // the latest coefficients (9)
    double *d_fhist = nullptr;              // force history ring (fhist_cap x 9), and its count
    unsigned long long *d_fhist_n = nullptr;
    int fhist_cap = 8192;
    // tangent / adjoint of the residual, allocated on first use
    double *ad_pool = nullptr;              // AD_N arrays of total elements
    double *ad_work = nullptr;              // metric slots of the largest block
    double *ad_force = nullptr;             // traction / coordinate gradients and coordinates of a wall
    // tangent / adjoint of a whole step, allocated on first use: AS_N arrays
    // of total elements, and the boundary tape (15 doubles per point of every
    // region of the BC sequence, at as_tape_off[region])
    double *as_pool = nullptr;
    double *as_tape = nullptr;
    std::vector<long> as_tape_off;
    int *as_ustart = nullptr, *as_uent = nullptr;   // multi-point entries per node (CSR)
    int as_nuniq = 0;
    int npp = 0, ngroups = 0;
    int *d_npp_pts = nullptr, *d_grp_start = nullptr, *d_grp_n = nullptr, *d_write_ok = nullptr;
    double *d_npp_buf = nullptr;
    std::vector<int> h_npp_pts, h_grp_n;

    // per-kernel timing with HIP events on this context's stream
    bool timing = false;
    struct Rec { int slot, launches; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    hipEvent_t event()
    {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[pool_used++];
    }

    DFields fields() const
    {
        DFields F{};
        F.r = prim[cur][0]; F.u = prim[cur][1]; F.v = prim[cur][2]; F.w = prim[cur][3]; F.p = prim[cur][4];
        F.rn = prim[cur ^ 1][0]; F.un = prim[cur ^ 1][1]; F.vn = prim[cur ^ 1][2];
        F.wn = prim[cur ^ 1][3]; F.pn = prim[cur ^ 1][4];
        F.t = t;
        F.kcx = metric[0]; F.kcy = metric[1]; F.kcz = metric[2];
        F.etx = metric[3]; F.ety = metric[4]; F.etz = metric[5];
        F.ctx = metric[6]; F.cty = metric[7]; F.ctz = metric[8]; F.vol = metric[9];
        F.rhs = rhs; F.dq = dq; F.dtdt = dtdt;
        F.sra = sr[0]; F.srb = sr[1]; F.src = sr[2];
        F.srva = srv[0]; F.srvb = srv[1]; F.srvc = srv[2];
        F.visl = visl; F.counters = counters; F.total = total; F.coed = coed;
        F.dq32 = lhs_prec == WCNS_LHS_FP32 ? dq32 : nullptr;
        F.qke = qke; F.vist = vist; F.dist = dist; F.qkw = qkw;
        F.gmet = d_gmet; F.gmoff = d_gmoff; F.rpres = d_rpres;
        return F;
    }
};

namespace {

// The fields as the j-direction kernels see them on the transposed view: the
// transposed copies, with the i and j metric triples exchanged so that the
// "j" direction of the view is the original i direction; the residual output
// goes to the transposed contribution df_t (zeroed before the launch).
DFields fields_t(const wcns_ctx *ctx, const DFields &F)
{
    DFields T = F;
    T.r = ctx->pt[0]; T.u = ctx->pt[1]; T.v = ctx->pt[2]; T.w = ctx->pt[3]; T.p = ctx->pt[4];
    T.t = ctx->t_t;
    T.kcx = ctx->mt[3]; T.kcy = ctx->mt[4]; T.kcz = ctx->mt[5];
    T.etx = ctx->mt[0]; T.ety = ctx->mt[1]; T.etz = ctx->mt[2];
    T.ctx = ctx->mt[6]; T.cty = ctx->mt[7]; T.ctz = ctx->mt[8]; T.vol = ctx->mt[9];
    T.visl = ctx->visl_t;
    T.vist = ctx->vist_t;
    T.rhs = ctx->df_t;
    return T;
}

// Copies the i-direction inputs to the transposed view.  The i-direction
// kernels of the view write every interior row of df_t (its ghost rows keep
// the zeros of the allocation), so it needs no clearing.
int stage_transposed(wcns_ctx *ctx, const DFields &F, bool vis)
{
    TransList L{};
    const double *src[5] = {F.r, F.u, F.v, F.w, F.p};
    // the inviscid stage after the viscous one of the same step reuses its copies
    if (vis || !ctx->pt_fresh)
        for (int m = 0; m < 5; m++) { L.src[L.n] = src[m]; L.dst[L.n++] = ctx->pt[m]; }
    ctx->pt_fresh = true;
    if (vis) {
        L.src[L.n] = F.t; L.dst[L.n++] = ctx->t_t;
        L.src[L.n] = F.visl; L.dst[L.n++] = ctx->visl_t;
        if (F.vist) { L.src[L.n] = F.vist; L.dst[L.n++] = ctx->vist_t; }
        // derivative slots: the view's direction 0 is the original j, direction 2 the original k
        for (int m = 0; m < 4; m++) { L.src[L.n] = ctx->dn + (4 + m) * F.total; L.dst[L.n++] = ctx->dn_t + m * F.total; }
        for (int m = 0; m < 4; m++) { L.src[L.n] = ctx->dn + (8 + m) * F.total; L.dst[L.n++] = ctx->dn_t + (8 + m) * F.total; }
    }
    return launch_transpose(0, ctx->grid[GRID_TRANS], L, ctx->stream) ? -3 : 0;
}

// rhs += the transposed i-direction contribution
int add_transposed(wcns_ctx *ctx, const DFields &F)
{
    TransList L{};
    for (int m = 0; m < 5; m++) { L.src[L.n] = ctx->df_t + m * F.total; L.dst[L.n++] = F.rhs + m * F.total; }
    return launch_transpose(1, ctx->grid[GRID_TRANS], L, ctx->stream);
}

// Records HIP events around the launches of one timing slot.
struct Scope {
    wcns_ctx *c;
    int slot, launches;
    hipEvent_t a{};
    Scope(wcns_ctx *ctx, int s, int n = 1) : c(ctx), slot(s), launches(n)
    {
        if (c->timing) { a = c->event(); hipEventRecord(a, c->stream); }
    }
    ~Scope()
    {
        if (c->timing) {
            hipEvent_t b = c->event();
            hipEventRecord(b, c->stream);
            c->recs.push_back({slot, launches, a, b});
        }
    }
};

int fail(wcns_ctx *ctx, int code, const std::string &msg)
{
    ctx->err = msg;
    return code;
}

// src/parallel.f90:1123-1301 analyze_bc + analyze_bc_connect (method = 1)
int analyze_region(wcns_ctx *ctx, int nb, int nr, HRegion &r)
{
    DRegion &d = r.d;
    memset(&d, 0, sizeof d);
    d.bctype = r.bctype;
    d.nb = nb;
    int s_st[3], s_ed[3];
    for (int m = 0; m < 3; m++) { s_st[m] = r.raw_s[2 * m]; s_ed[m] = r.raw_s[2 * m + 1]; }
    d.s_nd = 0;
    for (int m = 0; m < 3; m++) {
        d.s_lo[m] = std::min(std::abs(s_st[m]), std::abs(s_ed[m]));
        d.s_hi[m] = std::max(std::abs(s_st[m]), std::abs(s_ed[m]));
        d.l3[m] = 0;
        if (s_st[m] == s_ed[m]) {
            d.s_nd = m + 1;
            if (s_st[m] == 1) { d.s_lr = -1; d.l3[m] = -1; }
            else { d.s_lr = 1; d.l3[m] = 1; }
        }
    }
    if (d.s_nd == 0) return fail(ctx, WCNS_E_ARG, "region " + std::to_string(nr + 1) + " of block " +
                                                   std::to_string(nb + 1) + " is not a face");
    const HBlock &b = ctx->blocks[nb];
    int dims[3] = {b.ni, b.nj, b.nk};
    for (int m = 0; m < 3; m++)
        if (d.s_lo[m] < 1 || d.s_hi[m] > dims[m])
            return fail(ctx, WCNS_E_ARG, "region outside its block");
    d.npts = (d.s_hi[0] - d.s_lo[0] + 1) * (d.s_hi[1] - d.s_lo[1] + 1) * (d.s_hi[2] - d.s_lo[2] + 1);
    if (r.bctype >= 0) {
        if (r.bctype < 2 || r.bctype > 6) return fail(ctx, WCNS_E_BC, "unsupported bctype " + std::to_string(r.bctype));
        d.nbt = nb;
        return 0;
    }
    if (r.bctype != -1) return fail(ctx, WCNS_E_BC, "only bctype -1 interfaces (PreProcess.f90:2566)");
    if (r.nbt < 1 || r.nbt > (int)ctx->blocks.size()) return fail(ctx, WCNS_E_ARG, "bad target block");
    d.nbt = r.nbt - 1;
    int t_st[3], t_ed[3], t_nd = 0;
    for (int m = 0; m < 3; m++) { t_st[m] = r.raw_t[2 * m]; t_ed[m] = r.raw_t[2 * m + 1]; }
    for (int m = 0; m < 3; m++) {
        d.t_l3[m] = 0;
        if (t_st[m] == t_ed[m]) {
            t_nd = m + 1;
            d.t_l3[m] = t_st[m] == 1 ? -1 : 1;
        }
    }
    if (t_nd == 0) return fail(ctx, WCNS_E_ARG, "interface target is not a face");
    int sdir[3][3] = {{0}};
    bool found = false;
    for (int m = 1; m <= 3 && !found; m++)
        for (int n = 1; n <= 3 && !found; n++) {
            if (m == d.s_nd || n == t_nd) continue;
            int js1 = s_st[m - 1];
            if (std::abs(js1) < std::abs(s_ed[m - 1])) js1 = s_ed[m - 1];
            int js2 = t_st[n - 1];
            if (std::abs(js2) < std::abs(t_ed[n - 1])) js2 = t_ed[n - 1];
            if (js1 * js2 > 0) {
                sdir[n - 1][m - 1] = 1;
                sdir[t_nd - 1][d.s_nd - 1] = 1;
                sdir[6 - n - t_nd - 1][6 - m - d.s_nd - 1] = 1;
                found = true;
            }
        }
    int s_sign[3], t_sign[3], st_sign[3];
    for (int m = 0; m < 3; m++) {
        s_sign[m] = 1; s_st[m] = std::abs(s_st[m]); s_ed[m] = std::abs(s_ed[m]);
        if (m + 1 != d.s_nd && s_st[m] > s_ed[m]) s_sign[m] = -1;
        t_sign[m] = 1; t_st[m] = std::abs(t_st[m]); t_ed[m] = std::abs(t_ed[m]);
        if (m + 1 != t_nd && t_st[m] > t_ed[m]) t_sign[m] = -1;
    }
    for (int m = 0; m < 3; m++) {
        int co = 0;
        for (int n = 0; n < 3; n++) if (sdir[n][m] == 1) co = n + 1;
        st_sign[m] = (co > 0 ? t_sign[co - 1] : 1) * s_sign[m];
    }
    // image = t_st + sum_m sdir[a][m] * (idx_m - s_st_m) * st_sign_m
    for (int a = 0; a < 3; a++) {
        d.t0[a] = t_st[a];
        d.s0[a] = s_st[a];
        for (int m = 0; m < 3; m++) d.M[3 * a + m] = sdir[a][m] * st_sign[m];
    }
    // every image must land on the target block
    const HBlock &bt = ctx->blocks[d.nbt];
    int tdims[3] = {bt.ni, bt.nj, bt.nk};
    for (int t = 0; t < d.npts; t += std::max(1, d.npts - 1)) {
        int i, j, k, it, jt, kt;
        region_point(d, t, i, j, k);
        region_image(d, i, j, k, it, jt, kt);
        int im[3] = {it, jt, kt};
        for (int m = 0; m < 3; m++)
            if (im[m] < 1 || im[m] > tdims[m]) return fail(ctx, WCNS_E_ARG, "interface image outside target block");
    }
    return 0;
}

typedef std::tuple<int, int, int, int> Pt;

// parallel.f90:177-393 + 4202-4311: greedy multi-point groups, kept in order
void build_groups(wcns_ctx *ctx, std::vector<std::vector<Pt>> &groups)
{
    int nbk = (int)ctx->blocks.size();
    std::vector<std::vector<int>> iused(nbk);
    for (int nb = 0; nb < nbk; nb++) iused[nb].assign(ctx->blocks[nb].regs.size(), 0);
    for (int nb = 0; nb < nbk; nb++)
        for (size_t r = 0; r < ctx->blocks[nb].regs.size(); r++) {
            const HRegion &g = ctx->blocks[nb].regs[r];
            if (g.bctype < 0 && iused[nb][r] == 0) {
                iused[nb][r] = 1;
                int tb = g.nbt - 1, tr = g.ibcwin - 1;
                if (tb >= 0 && tb < nbk && tr >= 0 && tr < (int)iused[tb].size()) iused[tb][tr] = -1;
            }
        }
    std::vector<std::vector<Pt>> comp;
    std::map<Pt, std::vector<int>> where;        // point -> groups containing it (ascending)
    const int mcyc[5] = {1, 2, 3, 1, 2};
    for (int nb = 0; nb < nbk; nb++)
        for (size_t r = 0; r < ctx->blocks[nb].regs.size(); r++) {
            const HRegion &g = ctx->blocks[nb].regs[r];
            if (g.bctype >= 0 || iused[nb][r] != 1) continue;
            const DRegion &d = g.d;
            int id2 = mcyc[d.s_nd] - 1, id3 = mcyc[d.s_nd + 1] - 1;
            for (int k = d.s_lo[2]; k <= d.s_hi[2]; k++)
                for (int j = d.s_lo[1]; j <= d.s_hi[1]; j++)
                    for (int i = d.s_lo[0]; i <= d.s_hi[0]; i++) {
                        int idx[3] = {i, j, k};
                        bool inner = true;
                        for (int m = 0; m < 3; m++) {
                            int lo = d.s_lo[m], hi = d.s_hi[m];
                            if (m == id2 || m == id3) { lo++; hi--; }
                            if (idx[m] < lo || idx[m] > hi) inner = false;
                        }
                        if (inner) continue;
                        int it, jt, kt;
                        region_image(d, i, j, k, it, jt, kt);
                        Pt a{nb, i, j, k}, bt{d.nbt, it, jt, kt};
                        int first = -1;
                        auto ia = where.find(a), ib = where.find(bt);
                        if (ia != where.end()) first = ia->second.front();
                        if (ib != where.end() && (first < 0 || ib->second.front() < first)) first = ib->second.front();
                        if (first >= 0) {
                            auto &cg = comp[first];
                            bool ea = std::find(cg.begin(), cg.end(), a) != cg.end();
                            bool eb = std::find(cg.begin(), cg.end(), bt) != cg.end();
                            if (!ea) { cg.push_back(a); auto &w = where[a]; w.push_back(first); std::sort(w.begin(), w.end()); }
                            if (!eb) { cg.push_back(bt); auto &w = where[bt]; w.push_back(first); std::sort(w.begin(), w.end()); }
                        } else {
                            comp.push_back({a, bt});
                            int gi = (int)comp.size() - 1;
                            where[a].push_back(gi);
                            where[bt].push_back(gi);
                        }
                    }
        }
    groups.clear();
    for (auto &c : comp)
        if (c.size() > 2) groups.push_back(c);
}

long g3_of(int ni, int nj, int nk) { return (long)(ni + 6) * (nj + 6) * (nk + 6); }

// Every device operation of a context goes through its (non-blocking) stream,
// so that fills and copies are ordered with the kernels around them; the
// copies wait for completion because the host buffers are the caller's.
template <typename T>
int dalloc(wcns_ctx *ctx, T **p, size_t n)
{
    if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
        ctx->err = "hipMalloc of " + std::to_string(sizeof(T) * n) + " bytes failed";
        return WCNS_E_NOMEM;
    }
    if (hipMemsetAsync(*p, 0, sizeof(T) * (n ? n : 1), ctx->stream) != hipSuccess) return WCNS_E_HIP;
    return 0;
}

int copy_sync(wcns_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (hipMemcpyAsync(dst, src, bytes, kind, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctx->err = "device copy failed";
        return WCNS_E_HIP;
    }
    return 0;
}

#define TRY(x)                 \
    do {                       \
        int e__ = (x);         \
        if (e__) return e__;   \
    } while (0)

int check_counters(wcns_ctx *ctx)
{
    unsigned long long c[8];
    CK(hipMemcpyAsync(c, ctx->counters, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (c[2]) return fail(ctx, WCNS_E_ROE, "sqrt(c2) is a math error in flux_Roe / flux_vanleer (flux.f90:179-182, 437-441)");
    if (c[4]) return fail(ctx, WCNS_E_REPAIR, "too many abnormal nodes in update_limited (Sub2.f90:781-785)");
    if (ctx->d_sweep_err) {
        int e = 0;
        TRY(copy_sync(ctx, &e, ctx->d_sweep_err, sizeof(int), hipMemcpyDeviceToHost));
        if (e) return fail(ctx, WCNS_E_HIP, "a pipelined sweep wait expired");
    }
    return 0;
}

}  // namespace

// ======================================================================
extern "C" {

int wcns_create(wcns_ctx **out, int device, const wcns_params *prm)
{
    if (!out || !prm) return WCNS_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return WCNS_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return WCNS_E_HIP;
    wcns_ctx *ctx = new wcns_ctx();
    ctx->device = device;
    ctx->prm = *prm;
    if (const char *sw = getenv("WCNS_SWEEP")) ctx->tiled_sweeps = strcmp(sw, "plane") != 0;
    if (const char *id = getenv("WCNS_IDIR")) ctx->idir_t = strcmp(id, "direct") != 0;
    if (const char *sf = getenv("WCNS_SURFACE")) ctx->surf_levels = strcmp(sf, "sequence") != 0;
    if (const char *sp = getenv("WCNS_SGS_PAIR")) { ctx->plan.sgs_paired = atoi(sp) != 0; ctx->sgs_forced = true; }
    if (const char *sl = getenv("WCNS_SPIN_LIMIT")) ctx->spin_limit = std::max(0L, atol(sl));
    if (const char *ss = getenv("WCNS_SA_STAGED")) ctx->sa_staged = atoi(ss) != 0;
    if (const char *iw = getenv("WCNS_INV_WAVES")) ctx->dp.inv_waves = atoi(iw) == 2 ? 2 : (atoi(iw) == 1 ? 1 : 0);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return WCNS_E_HIP;
    }
    // the no-allocation scratch of this device (device current: hipSetDevice above)
    ctx->noalloc_scratch = noalloc_scratch_ptr(device);
    if (!ctx->noalloc_scratch) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return WCNS_E_HIP;
    }
    if (prm->nscheme != 4 && prm->nscheme != 41 && prm->nscheme != 42 && prm->nscheme != 45 && prm->nscheme != 47)
        ctx->err = "nscheme must be 4, 41, 42, 45 or 47";
    *out = ctx;
    return WCNS_OK;
}

int wcns_destroy(wcns_ctx *ctx)
{
    if (!ctx) return WCNS_E_ARG;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (int b = 0; b < 2; b++) for (int m = 0; m < 5; m++) hipFree(ctx->prim[b][m]);
    for (int m = 0; m < 10; m++) hipFree(ctx->metric[m]);
    for (int m = 0; m < 3; m++) { hipFree(ctx->sr[m]); hipFree(ctx->srv[m]); }
    void *pl[] = {ctx->d_sblocks, ctx->d_work, ctx->d_xch};
    for (void *p : pl) if (p) hipFree(p);
    void *ps[] = {ctx->fvb, ctx->tpool, ctx->d_blocks_t, ctx->d_grid, ctx->qkw, ctx->qke, ctx->vist, ctx->dist, ctx->sa_pool, ctx->skew[0], ctx->skew[1], ctx->dq32, ctx->coed, ctx->d_sweep_err, ctx->dn, ctx->t, ctx->rhs, ctx->dq, ctx->dtdt, ctx->visl, ctx->counters, ctx->d_blocks, ctx->d_flag, ctx->d_nflag, ctx->d_regs, ctx->d_rwork, ctx->d_skoff,
                  ctx->d_partial, ctx->d_res_wg0, ctx->d_resblk, ctx->d_res_out, ctx->d_hist, ctx->d_hist_n, ctx->d_npp_pts, ctx->d_grp_start, ctx->d_grp_n,
                  ctx->d_write_ok, ctx->d_npp_buf};
    for (void *p : ps) if (p) hipFree(p);
    for (auto &w : ctx->walls) hipFree(w.xyz);
    void *fs[] = {ctx->d_force_pn, ctx->d_force_part, ctx->d_force_out, ctx->ad_pool, ctx->ad_work, ctx->ad_force,
                  ctx->d_force_last, ctx->d_fhist, ctx->d_fhist_n, ctx->as_pool, ctx->as_tape, ctx->as_ustart, ctx->as_uent,
                  ctx->d_gmet, ctx->d_gmoff, ctx->d_rpres};
    for (void *p : fs) if (p) hipFree(p);
    for (hipEvent_t e : ctx->pool) hipEventDestroy(e);
    for (int c = 0; c < 2; c++)
        if (ctx->graph[c]) hipGraphExecDestroy(ctx->graph[c]);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return WCNS_OK;
}

const char *wcns_last_error(const wcns_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int wcns_add_block(wcns_ctx *ctx, int ni, int nj, int nk, const double *x, const double *y, const double *z, int *nb)
{
    if (!ctx) return WCNS_E_ARG;
    if (ctx->setup_done) return fail(ctx, WCNS_E_STATE, "blocks must be added before wcns_setup");
    if (ni < 2 || nj < 2 || nk < 2 || !x || !y || !z) return fail(ctx, WCNS_E_ARG, "bad block");
    HBlock b;
    b.ni = ni; b.nj = nj; b.nk = nk;
    size_t n = (size_t)ni * nj * nk;
    b.x.assign(x, x + n); b.y.assign(y, y + n); b.z.assign(z, z + n);
    b.g3 = g3_of(ni, nj, nk);
    ctx->blocks.push_back(std::move(b));
    if (nb) *nb = (int)ctx->blocks.size() - 1;
    return WCNS_OK;
}

int wcns_add_region(wcns_ctx *ctx, int nb, const int s_raw[6], int bctype, const int t_raw[6], int nbt, int ibcwin)
{
    if (!ctx || !s_raw) return WCNS_E_ARG;
    if (ctx->setup_done) return fail(ctx, WCNS_E_STATE, "regions must be added before wcns_setup");
    if (nb < 0 || nb >= (int)ctx->blocks.size()) return fail(ctx, WCNS_E_ARG, "bad block index");
    HRegion r{};
    r.bctype = bctype;
    memcpy(r.raw_s, s_raw, sizeof r.raw_s);
    if (bctype < 0) {
        if (!t_raw) return WCNS_E_ARG;
        memcpy(r.raw_t, t_raw, sizeof r.raw_t);
        r.nbt = nbt;
        r.ibcwin = ibcwin;
    }
    ctx->blocks[nb].regs.push_back(r);
    return WCNS_OK;
}

// Plane-major copy of the sweep operands for the pipelined sweeps; the
// metrics are staged once here, the per-step operands before every sweep set.
static int alloc_skew(wcns_ctx *ctx, int prec)
{
    size_t n = (size_t)SK_N * ctx->sktot;
    if (prec == 1) TRY(dalloc(ctx, (float **)&ctx->skew[1], n));
    else TRY(dalloc(ctx, (double **)&ctx->skew[0], n));
    DFields F = ctx->fields();
    for (size_t nb = 0; nb < ctx->blocks.size(); nb++)
        if (launch_skew_stage(SKEW_METRICS, ctx->blocks[nb].db, ctx->skoff[nb], F, prec, ctx->skew[prec], ctx->sktot,
                              ctx->stream))
            return fail(ctx, WCNS_E_HIP, "staging the metrics");
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// (Re)builds the batched sweep plan from ctx->tile_shape: per-block tile
// grids and exchange slots, and the work lists.  The forward list orders all
// tiles of all blocks by tile diagonal J+K, then block, then J -- a
// topological order of every block's tile graph, which interleaves the blocks
// so that their sweeps run side by side; the backward list is its reverse.
static int build_sweep_plan(wcns_ctx *ctx)
{
    std::vector<SweepBlock> sb;
    std::vector<long> xoff;
    long nx = 0;
    int lanes = 1;
    struct W { int key, nb, J, t; };
    std::vector<W> wl;
    for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
        const HBlock &b = ctx->blocks[nb];
        SweepBlock q{};
        q.b = b.db;
        q.soff = ctx->skoff[nb];
        q.TJ = ctx->tile_shape[nb].first;
        q.TK = ctx->tile_shape[nb].second;
        if (q.TJ < 1 || q.TK < 1 || q.TJ * q.TK > 64) return fail(ctx, WCNS_E_ARG, "bad sweep tile shape");
        q.nTJ = (b.nj + q.TJ - 1) / q.TJ;
        q.nTK = (b.nk + q.TK - 1) / q.TK;
        lanes = std::max(lanes, q.TJ * q.TK);
        xoff.push_back(nx);
        nx += sweep_exchange_words(q.TJ, q.TK, q.nTJ, q.nTK, b.ni);
        for (int t = 0; t < q.nTJ * q.nTK; t++) wl.push_back({t % q.nTJ + t / q.nTJ, (int)nb, t % q.nTJ, t});
        sb.push_back(q);
    }
    std::sort(wl.begin(), wl.end(), [](const W &a, const W &c) {
        return std::make_tuple(a.key, a.nb, a.J) < std::make_tuple(c.key, c.nb, c.J);
    });
    std::vector<int2> work;
    for (auto &w : wl) work.push_back(make_int2(w.nb, w.t));
    for (auto it = wl.rbegin(); it != wl.rend(); ++it) work.push_back(make_int2(it->nb, it->t));
    if (ctx->d_sblocks) { hipFree(ctx->d_sblocks); ctx->d_sblocks = nullptr; }
    if (ctx->d_work) { hipFree(ctx->d_work); ctx->d_work = nullptr; }
    if (ctx->d_xch) { hipFree(ctx->d_xch); ctx->d_xch = nullptr; }
    TRY(dalloc(ctx, &ctx->d_xch, nx));
    for (size_t nb = 0; nb < sb.size(); nb++) {
        sb[nb].xj = (double *)(ctx->d_xch + xoff[nb]);
        sb[nb].xk = sb[nb].xj + (long)sb[nb].nTJ * sb[nb].nTK * sb[nb].TK * sb[nb].b.ni * 5;
    }
    TRY(dalloc(ctx, &ctx->d_sblocks, sb.size()));
    TRY(dalloc(ctx, &ctx->d_work, work.size()));
    TRY(copy_sync(ctx, ctx->d_sblocks, sb.data(), sizeof(SweepBlock) * sb.size(), hipMemcpyHostToDevice));
    TRY(copy_sync(ctx, ctx->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
    SweepPlan &P = ctx->plan;
    P.blocks = ctx->d_sblocks;
    P.nwork = (int)wl.size();
    P.work_fwd = ctx->d_work;
    P.work_bwd = ctx->d_work + wl.size();
    P.lanes = lanes;
    P.xch = ctx->d_xch;
    P.nxch = nx;
    P.err = ctx->d_sweep_err;
    P.spin_limit = ctx->spin_limit;
    return 0;
}

int wcns_set_lhs_precision(wcns_ctx *ctx, int prec)
{
    if (!ctx || (prec != WCNS_LHS_FP64 && prec != WCNS_LHS_FP32)) return WCNS_E_ARG;
    if (prec == WCNS_LHS_FP32 && !ctx->tiled_sweeps)
        return fail(ctx, WCNS_E_UNSUPPORTED, "the FP32 implicit operator runs in the pipelined sweeps only");
    if (prec == WCNS_LHS_FP32 && ctx->setup_done && !ctx->dq32) TRY(dalloc(ctx, &ctx->dq32, 5 * ctx->total));
    if (prec == WCNS_LHS_FP32 && ctx->setup_done && ctx->tiled_sweeps && !ctx->skew[1]) TRY(alloc_skew(ctx, 1));
    ctx->lhs_prec = prec;
    return WCNS_OK;
}

// Dependency levels of region operations that must give the results of
// running them one after another (boundary_sequence's region order, or the
// interface averaging's).  touch(g, rd, wr) lists the pool nodes operation g
// reads and writes.  An operation goes one level above every earlier one that
// wrote a node it reads or writes, or read a node it writes, so the
// operations of one level touch no node that another of them writes.
extern "C++" {
template <typename Touch>
static bool region_levels(long total, const std::vector<DRegion> &seq, Touch touch, std::vector<int> &lvl)
{
    std::vector<uint16_t> lw(total, 0), lr(total, 0);   // 1 + level of the last write / read
    std::vector<long> rd, wr;
    lvl.assign(seq.size(), 0);
    for (size_t n = 0; n < seq.size(); n++) {
        rd.clear();
        wr.clear();
        touch(seq[n], rd, wr);
        int L = 0;
        for (long q : rd) L = std::max(L, (int)lw[q]);
        for (long q : wr) L = std::max(L, std::max((int)lw[q], (int)lr[q]));
        if (L + 1 > 0xffff) return false;       // more levels than the 16-bit stamps hold
        lvl[n] = L;
        const uint16_t L1 = (uint16_t)(L + 1);
        for (long q : rd) lr[q] = std::max(lr[q], L1);
        for (long q : wr) lw[q] = std::max(lw[q], L1);
    }
    return true;
}
}

// Work entries (region, first point) of the regions seq[n] + base, level by
// level; a serial region is one entry (region, -1).
static void level_work(const std::vector<DRegion> &seq, const std::vector<int> &lvl, int base,
                       std::vector<int2> &work, std::vector<int> &start)
{
    int nl = 0;
    for (int l : lvl) nl = std::max(nl, l + 1);
    for (int l = 0; l < nl; l++) {
        start.push_back((int)work.size());
        for (size_t n = 0; n < seq.size(); n++) {
            if (lvl[n] != l) continue;
            if (seq[n].serial) { work.push_back(make_int2(base + (int)n, -1)); continue; }
            for (int t = 0; t < seq[n].npts; t += REGION_CHUNK) work.push_back(make_int2(base + (int)n, t));
        }
    }
    start.push_back((int)work.size());
}

static int build_region_plans(wcns_ctx *ctx)
{
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    std::vector<DRegion> bcs;
    for (auto &b : ctx->blocks)
        for (int r : b.bcorder) bcs.push_back(b.regs[r].d);
    // bc_point: the face node and three ghost layers written; three interior
    // layers read, and for an interface three layers of the partner block
    auto bc_touch = [&](const DRegion &g, std::vector<long> &rd, std::vector<long> &wr) {
        const DBlock &b = dbs[g.nb];
        const int *l3 = g.l3;
        for (int t = 0; t < g.npts; t++) {
            int i, j, k;
            region_point(g, t, i, j, k);
            for (int a = 0; a <= 3; a++) wr.push_back(gidx(b, i + a * l3[0], j + a * l3[1], k + a * l3[2]));
            for (int a = 1; a <= 3; a++) rd.push_back(gidx(b, i - a * l3[0], j - a * l3[1], k - a * l3[2]));
            if (g.bctype < 0) {
                const DBlock &bt = dbs[g.nbt];
                const int *tl = g.t_l3;
                int it, jt, kt;
                region_image(g, i, j, k, it, jt, kt);
                for (int a = 1; a <= 3; a++) rd.push_back(gidx(bt, it - a * tl[0], jt - a * tl[1], kt - a * tl[2]));
            }
        }
    };
    // avg2pm_point: the node written, its image read
    auto if_touch = [&](const DRegion &g, std::vector<long> &rd, std::vector<long> &wr) {
        for (int t = 0; t < g.npts; t++) {
            int i, j, k, it, jt, kt;
            region_point(g, t, i, j, k);
            region_image(g, i, j, k, it, jt, kt);
            wr.push_back(gidx(dbs[g.nb], i, j, k));
            rd.push_back(gidx(dbs[g.nbt], it, jt, kt));
        }
    };
    std::vector<int> lb, li;
    if (!region_levels(ctx->total, bcs, bc_touch, lb) || !region_levels(ctx->total, ctx->iface, if_touch, li))
        return fail(ctx, WCNS_E_UNSUPPORTED, "more than 65534 dependency levels of boundary regions");
    std::vector<DRegion> regs = bcs;
    regs.insert(regs.end(), ctx->iface.begin(), ctx->iface.end());
    regs.insert(regs.end(), ctx->phys.begin(), ctx->phys.end());
    std::vector<int2> work;
    ctx->bc_lvl.clear();
    ctx->if_lvl.clear();
    level_work(bcs, lb, 0, work, ctx->bc_lvl);
    level_work(ctx->iface, li, (int)bcs.size(), work, ctx->if_lvl);
    ctx->zero_w0 = (int)work.size();
    std::vector<int> l0(ctx->phys.size(), 0);
    std::vector<DRegion> ph = ctx->phys;
    for (auto &g : ph) g.serial = 0;
    std::vector<int> zs;
    level_work(ph, l0, (int)(bcs.size() + ctx->iface.size()), work, zs);
    ctx->zero_nw = (int)work.size() - ctx->zero_w0;
    TRY(dalloc(ctx, &ctx->d_regs, std::max<size_t>(regs.size(), 1)));
    TRY(dalloc(ctx, &ctx->d_rwork, std::max<size_t>(work.size(), 1)));
    if (!regs.empty()) TRY(copy_sync(ctx, ctx->d_regs, regs.data(), sizeof(DRegion) * regs.size(), hipMemcpyHostToDevice));
    if (!work.empty()) TRY(copy_sync(ctx, ctx->d_rwork, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
    return WCNS_OK;
}

int wcns_setup(wcns_ctx *ctx)
{
    if (!ctx) return WCNS_E_ARG;
    if (ctx->setup_done) return fail(ctx, WCNS_E_STATE, "setup already done");
    if (ctx->blocks.empty()) return fail(ctx, WCNS_E_STATE, "no blocks");
    CK(hipSetDevice(ctx->device));
    CK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const wcns_params &p = ctx->prm;
    if (p.nscheme != 4 && p.nscheme != 41 && p.nscheme != 42 && p.nscheme != 45 && p.nscheme != 47)
        return fail(ctx, WCNS_E_UNSUPPORTED, "nscheme must be 4, 41, 42, 45 or 47");
    if (p.nflux < 1 || p.nflux > 4)
        return fail(ctx, WCNS_E_UNSUPPORTED, "nflux must be 1 (SW), 2 (van Leer), 3 (Roe) or 4 (RoeMx)");
    // invcode dispatches flux_roemx from schemes 42 and 47 only (Convect.f90:34, 47)
    if (p.nflux == 4 && p.nscheme != 42 && p.nscheme != 47)
        return fail(ctx, WCNS_E_UNSUPPORTED, "nflux 4 (RoeMx) goes with nscheme 42 or 47");
    if (p.nvis != 0 && p.nvis != 1) return fail(ctx, WCNS_E_ARG, "nvis must be 0 or 1");
    if (p.nlamtur >= 0 && !((p.nlamtur == 1 || p.nlamtur == 2) && p.nvis == 1))
        return fail(ctx, WCNS_E_UNSUPPORTED, "turbulence models: SA (nlamtur = 1) or SST (nlamtur = 2), with nvis = 1");
    const bool turb = p.nlamtur == 1 || p.nlamtur == 2;
    const bool sst = p.nlamtur == 2;
    if (p.lfref <= 0.0 || p.lref <= 0.0) return fail(ctx, WCNS_E_ARG, "lfref and lref must be positive");
    if (p.ntmst != 0) return fail(ctx, WCNS_E_UNSUPPORTED, "only local time stepping (ntmst = 0)");
    // layout
    long off = 0;
    for (auto &b : ctx->blocks) {
        for (int n : {b.ni, b.nj, b.nk}) {
            if (n > 4 && n < 12) return fail(ctx, WCNS_E_UNSUPPORTED, "directions with 5..11 nodes are not supported");
            if (p.nvis == 1 && n <= 4) return fail(ctx, WCNS_E_UNSUPPORTED, "2-D (n <= 4) viscous directions are not supported");
        }
        b.db.ni = b.ni; b.db.nj = b.nj; b.db.nk = b.nk; b.db.pad = 0;
        b.db.sx = b.ni + 6;
        b.db.sy = (long)(b.ni + 6) * (b.nj + 6);
        b.db.off = off;
        off += b.g3;
    }
    ctx->total = off;
    // regions
    static const int order[6] = {-1, 4, 5, 6, 2, 3};      // PreProcess.f90:2543-2573
    for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
        HBlock &b = ctx->blocks[nb];
        for (size_t r = 0; r < b.regs.size(); r++) TRY(analyze_region(ctx, (int)nb, (int)r, b.regs[r]));
        b.bcorder.clear();
        for (int o : order)
            for (size_t r = 0; r < b.regs.size(); r++)
                if ((b.regs[r].bctype < 0 ? -1 : b.regs[r].bctype) == o) b.bcorder.push_back((int)r);
        if (b.bcorder.size() != b.regs.size()) return fail(ctx, WCNS_E_BC, "incorrect BC types");
    }
    // device fields
    long T = ctx->total;
    for (int bb = 0; bb < 2; bb++) for (int m = 0; m < 5; m++) TRY(dalloc(ctx, &ctx->prim[bb][m], T));
    for (int m = 0; m < 10; m++) TRY(dalloc(ctx, &ctx->metric[m], T));
    for (int m = 0; m < 3; m++) { TRY(dalloc(ctx, &ctx->sr[m], T)); TRY(dalloc(ctx, &ctx->srv[m], T)); }
    TRY(dalloc(ctx, &ctx->t, T));
    TRY(dalloc(ctx, &ctx->rhs, 5 * T));
    TRY(dalloc(ctx, &ctx->dq, 5 * T));
    if (ctx->lhs_prec == WCNS_LHS_FP32) TRY(dalloc(ctx, &ctx->dq32, 5 * T));
    TRY(dalloc(ctx, &ctx->dtdt, T));
    TRY(dalloc(ctx, &ctx->visl, T));
    if (p.nvis == 1) TRY(dalloc(ctx, &ctx->dn, 12 * T));
    if (p.nvis == 1) TRY(dalloc(ctx, &ctx->fvb, 5 * T));
    if (turb) {
        TRY(dalloc(ctx, &ctx->qke, T));
        TRY(dalloc(ctx, &ctx->vist, T));
        TRY(dalloc(ctx, &ctx->dist, T));
        TRY(dalloc(ctx, &ctx->sa_pool, (sst ? 13 : 6) * T));
        ctx->sa.grad = ctx->dn;
        ctx->sa.viseq[0] = ctx->sa_pool;
        for (int d = 0; d < 3; d++) ctx->sa.un[d] = ctx->sa_pool + (1 + d) * T;
        ctx->sa.dqke[0] = ctx->sa_pool + 4 * T;
        ctx->sa.spec[0] = ctx->sa_pool + 5 * T;
        if (sst) {
            TRY(dalloc(ctx, &ctx->qkw, T));
            ctx->sa.gradw = ctx->sa_pool + 6 * T;
            ctx->sa.viseq[1] = ctx->sa_pool + 9 * T;
            ctx->sa.dqke[1] = ctx->sa_pool + 10 * T;
            ctx->sa.spec[1] = ctx->sa_pool + 11 * T;
            ctx->sa.f1 = ctx->sa_pool + 12 * T;
        }
    }
    TRY(dalloc(ctx, &ctx->coed, T));
    TRY(dalloc(ctx, &ctx->d_sweep_err, 1));
    // sweep tiles: TJ x TK columns per workgroup (16 x 4 unless tuned)
    ctx->tile_shape.clear();
    for (auto &b : ctx->blocks) {
        int TJ = std::min(16, b.nj);
        ctx->tile_shape.push_back({TJ, std::max(1, std::min(64 / TJ, b.nk))});
    }
    TRY(dalloc(ctx, &ctx->counters, 8));
    TRY(dalloc(ctx, &ctx->d_flag, (size_t)ctx->flag_cap * ctx->blocks.size()));
    TRY(dalloc(ctx, &ctx->d_nflag, ctx->blocks.size()));
    std::vector<DBlock> hb;
    std::vector<int> res_wg0(1, 0);
    for (auto &b : ctx->blocks) { hb.push_back(b.db); res_wg0.push_back(res_wg0.back() + residual_parts(b.db)); }
    TRY(dalloc(ctx, &ctx->d_blocks, hb.size()));
    TRY(copy_sync(ctx, ctx->d_blocks, hb.data(), sizeof(DBlock) * hb.size(), hipMemcpyHostToDevice));
    // transposed views of the blocks (i <-> j), for the i-direction residuals
    std::vector<DBlock> hbt;
    for (auto &b : hb) hbt.push_back(transposed_view(b));
    TRY(dalloc(ctx, &ctx->d_blocks_t, hbt.size()));
    TRY(copy_sync(ctx, ctx->d_blocks_t, hbt.data(), sizeof(DBlock) * hbt.size(), hipMemcpyHostToDevice));
    // block work-lists: every block's workgroups of a launch kind, block after block
    {
        std::vector<int2> all;
        size_t start[GRID_N];
        for (int kind = 0; kind < GRID_N; kind++) {
            start[kind] = all.size();
            const long bs = grid_block_size(kind);
            const bool tview = kind == GRID_RHS_INV_T || kind == GRID_RHS_VIS_T || kind == GRID_VFACE || kind == GRID_VROW;
            std::vector<int2> lst;
            for (size_t nb = 0; nb < hb.size(); nb++) {
                const long nw = (grid_threads(kind, tview ? hbt[nb] : hb[nb]) + bs - 1) / bs;
                for (long w = 0; w < nw; w++) lst.push_back(make_int2((int)nb, (int)w));
            }
            // Workgroup ids go round-robin to the XCDs, each with an L2 of its
            // own.  Entry p of the launch takes item (p % XCDS) * share + p / XCDS
            // of the list, so that each XCD works through one contiguous slice
            // of the domain and stencil neighbours (the k +- 1, 2 planes) are
            // fetched into one L2 instead of all of them.  Measured: it pays for
            // the transposes always, and for the stencil kernels on large
            // launches (64 blocks, 30 M nodes: viscous faces and rows -8 %);
            // on one 2 M-node block the viscous kernels ran 10-30 % slower with
            // it, so launches of at most XCD_SLICE_MIN workgroups keep the
            // plain order.
            const long n = (long)lst.size(), q = n / XCDS, r = n % XCDS;
            // the residual monitor's partials are combined per block in list order
            const bool slices = kind != GRID_RESID && (kind == GRID_TRANS || n > XCD_SLICE_MIN);
            for (long pos = 0; pos < n; pos++) {
                const long x = pos % XCDS, idx = pos / XCDS;
                const long first = x * q + std::min(x, r);
                all.push_back(slices ? lst[first + idx] : lst[pos]);
            }
        }
        TRY(dalloc(ctx, &ctx->d_grid, std::max<size_t>(all.size(), 1)));
        if (!all.empty())
            TRY(copy_sync(ctx, ctx->d_grid, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice));
        for (int kind = 0; kind < GRID_N; kind++) {
            const size_t end = kind + 1 < GRID_N ? start[kind + 1] : all.size();
            const bool tview = kind == GRID_RHS_INV_T || kind == GRID_RHS_VIS_T || kind == GRID_VFACE || kind == GRID_VROW;
            ctx->grid[kind] = BlockGrid{tview ? ctx->d_blocks_t : ctx->d_blocks, ctx->d_grid + start[kind],
                                        (int)(end - start[kind])};
        }
    }
    if (ctx->idir_t) {
        // metrics 10, primitives 5, t, visl, vist, 12 derivative slots, 5 residual components
        const int ntp = 10 + 5 + 3 + 12 + 5;
        TRY(dalloc(ctx, &ctx->tpool, (size_t)ntp * T));
        double *q = ctx->tpool;
        for (int m = 0; m < 10; m++, q += T) ctx->mt[m] = q;
        for (int m = 0; m < 5; m++, q += T) ctx->pt[m] = q;
        ctx->t_t = q; q += T;
        ctx->visl_t = q; q += T;
        ctx->vist_t = q; q += T;
        ctx->dn_t = q; q += 12 * T;
        ctx->df_t = q;
    }
    ctx->res_nwg = res_wg0.back();
    TRY(dalloc(ctx, &ctx->d_partial, 3 * (size_t)res_wg0.back()));
    TRY(dalloc(ctx, &ctx->d_res_wg0, res_wg0.size()));
    TRY(copy_sync(ctx, ctx->d_res_wg0, res_wg0.data(), sizeof(int) * res_wg0.size(), hipMemcpyHostToDevice));
    TRY(dalloc(ctx, &ctx->d_resblk, 3 * ctx->blocks.size()));
    TRY(dalloc(ctx, &ctx->d_res_out, 8));
    TRY(dalloc(ctx, &ctx->d_hist, 8 * (size_t)ctx->hist_cap));
    TRY(dalloc(ctx, &ctx->d_hist_n, 1));
    CK(hipMemsetAsync(ctx->d_hist_n, 0, sizeof(unsigned long long), ctx->stream));

    // freestream (PreProcess.f90:968-1113, height < 0)
    DParams &d = ctx->dp;
    const double pai = 3.14159265;
    double att = p.attack * pai / 180.0, ssl = p.sideslip * pai / 180.0;
    d.gama = p.gama; d.efix = p.efix; d.cfl = p.cfl; d.timedt_rate = p.timedt_rate; d.moo = p.moo;
    d.reynolds = (p.reynolds / p.lfref) * p.lref;       // PreProcess.f90:994
    d.csrv = p.csrv; d.re = 1.0 / d.reynolds; d.visc = 110.4 / p.tref; d.prl = p.prl; d.prt = p.prt;
    d.roo = 1.0; d.uoo = cos(att) * cos(ssl); d.voo = sin(att) * cos(ssl); d.woo = sin(ssl);
    d.poo = 1.0 / (p.gama * p.moo * p.moo);
    d.coo = sqrt(p.gama * d.poo / d.roo);
    d.twall_fixed = p.twall > 0.0;
    d.twall_nd = p.twall / p.tref;
    d.nvis = p.nvis; d.nscheme = p.nscheme; d.nflux = p.nflux;
    d.turb = turb ? p.nlamtur : 0;
    // turblence.f90:3166-3181 inflow_turbulence
    d.vtoo = 1.0e-6;
    d.muoo = 1.0e-1;
    if (sst) {
        d.omgoo = 10.0 * (p.lref / p.lfref);
        d.koo = d.omgoo * d.vtoo / d.reynolds;
    }
    d.kmaxlim = p.kmaxlim; d.timedt_turb = p.timedt_turb; d.xk = p.xk; d.xb = p.xb;

    // wall distance (SA): every node of every no-slip wall region (bctype 2, 20, 21),
    // parallel.f90:4313-4468
    double *d_wall = nullptr;
    int nwall = 0;
    if (turb) {
        std::vector<double> wx, wy, wz;
        for (auto &b : ctx->blocks)
            for (auto &r : b.regs) {
                if (!(r.bctype == 2 || r.bctype == 20 || r.bctype == 21)) continue;
                for (int t = 0; t < r.d.npts; t++) {
                    int i, j, k;
                    region_point(r.d, t, i, j, k);
                    size_t q = (size_t)(i - 1) + (size_t)b.ni * ((size_t)(j - 1) + (size_t)b.nj * (size_t)(k - 1));
                    wx.push_back(b.x[q]); wy.push_back(b.y[q]); wz.push_back(b.z[q]);
                }
            }
        nwall = (int)wx.size();
        TRY(dalloc(ctx, &d_wall, 3 * (size_t)std::max(nwall, 1)));
        if (nwall) {
            wx.insert(wx.end(), wy.begin(), wy.end());
            wx.insert(wx.end(), wz.begin(), wz.end());
            TRY(copy_sync(ctx, d_wall, wx.data(), 8 * wx.size(), hipMemcpyHostToDevice));
        }
    }

    // GCL metrics per block on the device
    DFields F = ctx->fields();
    for (auto &b : ctx->blocks) {
        long n = (long)b.ni * b.nj * b.nk;
        double *dx, *dy, *dz, *work;
        TRY(dalloc(ctx, &dx, n)); TRY(dalloc(ctx, &dy, n)); TRY(dalloc(ctx, &dz, n));
        TRY(dalloc(ctx, &work, 21 * b.g3 + 1));
        TRY(copy_sync(ctx, dx, b.x.data(), 8 * n, hipMemcpyHostToDevice));
        TRY(copy_sync(ctx, dy, b.y.data(), 8 * n, hipMemcpyHostToDevice));
        TRY(copy_sync(ctx, dz, b.z.data(), 8 * n, hipMemcpyHostToDevice));
        if (launch_gcl_metrics(b.db, dx, dy, dz, F, work, ctx->stream)) return fail(ctx, WCNS_E_HIP, "gcl launch");
        if (turb && launch_wall_dist(b.db, dx, dy, dz, d_wall, nwall, ctx->dist, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "wall distance launch");
        int nbad = 0;
        CK(hipMemcpyAsync(&nbad, (int *)(work + 21 * b.g3), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        hipFree(dx); hipFree(dy); hipFree(dz); hipFree(work);
        if (nbad) return fail(ctx, WCNS_E_JACOBIAN, std::to_string(nbad) + " nodes with non-positive Jacobian");
    }
    if (d_wall) hipFree(d_wall);
    if (turb) {
        // init_SA (turblence.f90:2701-2734): nu~ = muoo, mu_t = muoo fv1(muoo);
        // init_SST (:3091-3117): k = koo, omega = omgoo, mu_t = vtoo; on -1..n+1
        const double cv13 = 7.1 * 7.1 * 7.1, kafan = d.muoo;
        const double kf3 = kafan * kafan * kafan;
        const double fv1 = kf3 / (kf3 + cv13);
        F = ctx->fields();
        for (auto &b : ctx->blocks) {
            int e = sst ? launch_sa_init(b.db, F, d.koo, d.omgoo, d.vtoo, ctx->stream)
                        : launch_sa_init(b.db, F, d.muoo, 0.0, d.muoo * fv1, ctx->stream);
            if (e) return fail(ctx, WCNS_E_HIP, "turbulence init");
        }
    }
    if (ctx->idir_t) {
        TransList L{};
        for (int m = 0; m < 10; m++) { L.src[L.n] = ctx->metric[m]; L.dst[L.n++] = ctx->mt[m]; }
        if (launch_transpose(0, ctx->grid[GRID_TRANS], L, ctx->stream)) return fail(ctx, WCNS_E_HIP, "transposing the metrics");
    }
    // symmetry-plane normals (BC.f90:381-396) need the metrics at the face centre
    for (auto &b : ctx->blocks)
        for (auto &r : b.regs) {
            if (r.bctype != 3) continue;
            DRegion &g = r.d;
            int dims[3] = {b.ni, b.nj, b.nk};
            if (dims[g.s_nd - 1] <= 4) return fail(ctx, WCNS_E_UNSUPPORTED, "2-D symmetry branch not built");
            int c[3] = {(g.s_lo[0] + g.s_hi[0]) / 2, (g.s_lo[1] + g.s_hi[1]) / 2, (g.s_lo[2] + g.s_hi[2]) / 2};
            long q = gidx(b.db, c[0], c[1], c[2]);
            double m[9];
            for (int a = 0; a < 9; a++) TRY(copy_sync(ctx, &m[a], ctx->metric[a] + q, 8, hipMemcpyDeviceToHost));
            double x = 0, y = 0, z = 0;
            for (int a = 0; a < 3; a++)
                if (g.l3[a]) { x = x + m[3 * a]; y = y + m[3 * a + 1]; z = z + m[3 * a + 2]; }
            double s = std::max(sqrt(x * x + y * y + z * z), 1.0e-30);
            double nx = x / s, ny = y / s, nz = z / s;
            double s2 = sqrt(nx * nx + ny * ny + nz * nz);
            g.nrm[0] = nx / s2; g.nrm[1] = ny / s2; g.nrm[2] = nz / s2;
        }
    // schemes 45 / 47: the ghost metrics at both ends of every pencil
    // (Convect.f90:180-297), region by region in the BC order
    if (d.nscheme == 45 || d.nscheme == 47) {
        std::vector<long> gmoff;
        long tot = 0;
        for (auto &b : ctx->blocks) {
            gmoff.push_back(tot);
            tot += 24 * ((long)b.nj * b.nk + (long)b.ni * b.nk + (long)b.ni * b.nj);
        }
        TRY(dalloc(ctx, &ctx->d_gmet, (size_t)tot));
        TRY(dalloc(ctx, &ctx->d_gmoff, gmoff.size()));
        TRY(copy_sync(ctx, ctx->d_gmoff, gmoff.data(), sizeof(long) * gmoff.size(), hipMemcpyHostToDevice));
        F = ctx->fields();
        for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
            const HBlock &b = ctx->blocks[nb];
            for (int r : b.bcorder) {
                const DRegion &g = b.regs[r].d;
                if (launch_ghost_metrics(g, b.db, ctx->blocks[g.nbt].db, F, ctx->d_gmet, gmoff[nb], ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "ghost metrics");
            }
        }
    }
    // flux 4: the pressure-ratio arrays; the entries the per-step kernel
    // never writes stay +inf
    if (d.nflux == 4) {
        TRY(dalloc(ctx, &ctx->d_rpres, 3 * (size_t)T));
        if (launch_fill_u64(reinterpret_cast<unsigned long long *>(ctx->d_rpres), 3 * T, 0x7FF0000000000000ull,
                            ctx->stream))
            return fail(ctx, WCNS_E_HIP, "pressure-ratio fill");
    }
    // plane-major sweep operands
    ctx->skoff.clear();
    ctx->sktot = 0;
    for (auto &b : ctx->blocks) { ctx->skoff.push_back(ctx->sktot); ctx->sktot += skew_size(b.db); }
    TRY(dalloc(ctx, &ctx->d_skoff, ctx->skoff.size()));
    TRY(copy_sync(ctx, ctx->d_skoff, ctx->skoff.data(), sizeof(long) * ctx->skoff.size(), hipMemcpyHostToDevice));
    if (ctx->tiled_sweeps) {
        TRY(alloc_skew(ctx, 0));
        if (ctx->lhs_prec == WCNS_LHS_FP32) TRY(alloc_skew(ctx, 1));
    }
    if (ctx->tiled_sweeps || turb) TRY(build_sweep_plan(ctx));
    // interface lists: 2pm regions in (nb, nr) order, serial when a region reads its own writes
    ctx->iface.clear();
    ctx->phys.clear();
    for (auto &b : ctx->blocks)
        for (auto &r : b.regs) {
            if (r.bctype > 0) ctx->phys.push_back(r.d);
            if (r.bctype == 2 || r.bctype / 10 == 2) ctx->walls.push_back({r.d, nullptr});
            if (r.bctype >= 0) continue;
            DRegion g = r.d;
            if (g.nbt == g.nb) {
                std::set<std::tuple<int, int, int>> w;
                for (int t = 0; t < g.npts; t++) { int i, j, k; region_point(g, t, i, j, k); w.insert({i, j, k}); }
                for (int t = 0; t < g.npts && !g.serial; t++) {
                    int i, j, k, it, jt, kt;
                    region_point(g, t, i, j, k);
                    region_image(g, i, j, k, it, jt, kt);
                    if (w.count({it, jt, kt}) && !(it == i && jt == j && kt == k)) g.serial = 1;
                }
            }
            ctx->iface.push_back(g);
        }
    // multi-point groups
    std::vector<std::vector<Pt>> groups;
    build_groups(ctx, groups);
    std::vector<int> pts, gstart, gn, ok;
    std::map<Pt, int> last;
    for (auto &g : groups) {
        gstart.push_back((int)pts.size() / 4);
        gn.push_back((int)g.size());
        for (auto &q : g) {
            last[q] = (int)pts.size() / 4;
            pts.push_back(std::get<0>(q)); pts.push_back(std::get<1>(q));
            pts.push_back(std::get<2>(q)); pts.push_back(std::get<3>(q));
        }
    }
    int np = (int)pts.size() / 4;
    ok.assign(np, 0);
    for (auto &kv : last) ok[kv.second] = 1;   // a point written by several groups keeps the last write
    ctx->npp = np;
    ctx->ngroups = (int)groups.size();
    ctx->h_npp_pts = pts;
    ctx->h_grp_n = gn;
    TRY(dalloc(ctx, &ctx->d_npp_pts, std::max<size_t>(pts.size(), 4)));
    TRY(dalloc(ctx, &ctx->d_grp_start, std::max<size_t>(gstart.size(), 1)));
    TRY(dalloc(ctx, &ctx->d_grp_n, std::max<size_t>(gn.size(), 1)));
    TRY(dalloc(ctx, &ctx->d_write_ok, std::max<size_t>(ok.size(), 1)));
    TRY(dalloc(ctx, &ctx->d_npp_buf, 6 * std::max(np, 1)));
    if (np) {
        TRY(copy_sync(ctx, ctx->d_npp_pts, pts.data(), 4 * pts.size(), hipMemcpyHostToDevice));
        TRY(copy_sync(ctx, ctx->d_grp_start, gstart.data(), 4 * gstart.size(), hipMemcpyHostToDevice));
        TRY(copy_sync(ctx, ctx->d_grp_n, gn.data(), 4 * gn.size(), hipMemcpyHostToDevice));
        TRY(copy_sync(ctx, ctx->d_write_ok, ok.data(), 4 * ok.size(), hipMemcpyHostToDevice));
    }
    TRY(build_region_plans(ctx));
    // walls: face-node coordinates in region-point order, force scratch
    size_t maxw = 1, maxq = 1;
    for (auto &w : ctx->walls) {
        const HBlock &b = ctx->blocks[w.d.nb];
        std::vector<double> xyz(3 * (size_t)w.d.npts);
        for (int t = 0; t < w.d.npts; t++) {
            int i, j, k;
            region_point(w.d, t, i, j, k);
            size_t q = (size_t)(i - 1) + (size_t)b.ni * ((size_t)(j - 1) + (size_t)b.nj * (size_t)(k - 1));
            xyz[t] = b.x[q]; xyz[w.d.npts + t] = b.y[q]; xyz[2 * (size_t)w.d.npts + t] = b.z[q];
        }
        TRY(dalloc(ctx, &w.xyz, xyz.size()));
        TRY(copy_sync(ctx, w.xyz, xyz.data(), 8 * xyz.size(), hipMemcpyHostToDevice));
        maxw = std::max(maxw, (size_t)w.d.npts);
        maxq = std::max(maxq, (size_t)force_parts(w.d));
    }
    TRY(dalloc(ctx, &ctx->d_force_pn, 3 * maxw));
    TRY(dalloc(ctx, &ctx->d_force_part, 6 * maxq));
    TRY(dalloc(ctx, &ctx->d_force_out, 6 * std::max<size_t>(ctx->walls.size(), 1)));
    TRY(dalloc(ctx, &ctx->d_force_last, 9));
    TRY(dalloc(ctx, &ctx->d_fhist, 9 * (size_t)ctx->fhist_cap));
    TRY(dalloc(ctx, &ctx->d_fhist_n, 1));
    CK(hipMemsetAsync(ctx->d_fhist_n, 0, sizeof(unsigned long long), ctx->stream));
    ctx->setup_done = true;
    return WCNS_OK;
}

static int field_ptr(wcns_ctx *ctx, int which, double **p, int *ncomp)
{
    *ncomp = 1;
    switch (which) {
    case 0: case 1: case 2: case 3: case 4: *p = ctx->prim[ctx->cur][which]; return 0;
    case 5: *p = ctx->t; return 0;
    case 7: *p = ctx->dtdt; return 0;
    case 8: case 9: case 10: *p = ctx->sr[which - 8]; return 0;
    case 11: case 12: case 13: *p = ctx->srv[which - 11]; return 0;
    case 14: *p = ctx->visl; return 0;
    case 15: if (!ctx->vist) return WCNS_E_ARG; *p = ctx->vist; return 0;
    case 16: *p = ctx->metric[9]; return 0;
    case 26: if (!ctx->qke) return WCNS_E_ARG; *p = ctx->qke; return 0;
    case 27: if (!ctx->dist) return WCNS_E_ARG; *p = ctx->dist; return 0;
    case 28: if (!ctx->qkw) return WCNS_E_ARG; *p = ctx->qkw; return 0;
    case 100:
        if (ctx->lhs_prec == WCNS_LHS_FP32) return WCNS_E_UNSUPPORTED;
        *p = ctx->dq; *ncomp = 5; return 0;
    case 102: *p = ctx->rhs; *ncomp = 5; return 0;
    default:
        if (which >= 17 && which <= 25) { *p = ctx->metric[which - 17]; return 0; }
        return WCNS_E_ARG;
    }
}

int wcns_set_prim(wcns_ctx *ctx, int nb, const double *r, const double *u, const double *v, const double *w,
                  const double *p)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (nb < 0 || nb >= (int)ctx->blocks.size()) return WCNS_E_ARG;
    ctx->pt_fresh = false;
    HBlock &b = ctx->blocks[nb];
    const double *src[5] = {r, u, v, w, p};
    for (int bb = 0; bb < 2; bb++)
        for (int m = 0; m < 5; m++)
            TRY(copy_sync(ctx, ctx->prim[bb][m] + b.db.off, src[m], 8 * b.g3, hipMemcpyHostToDevice));
    // initial_t (sub_tgh.f90:36-58): temperature on the interior for nvis == 1
    std::vector<double> t(b.g3, 0.0);
    if (ctx->prm.nvis == 1) {
        double moo2 = ctx->prm.moo * ctx->prm.moo;
        for (int k = 1; k <= b.nk; k++) for (int j = 1; j <= b.nj; j++) for (int i = 1; i <= b.ni; i++) {
            long q = gidx(b.db, i, j, k) - b.db.off;
            t[q] = moo2 * (ctx->prm.gama * p[q] / r[q]);
        }
    }
    TRY(copy_sync(ctx, ctx->t + b.db.off, t.data(), 8 * b.g3, hipMemcpyHostToDevice));
    return WCNS_OK;
}

int wcns_get_field(wcns_ctx *ctx, int nb, int which, double *out)
{
    if (!ctx || !ctx->setup_done || !out) return WCNS_E_STATE;
    if (nb < 0 || nb >= (int)ctx->blocks.size()) return WCNS_E_ARG;
    double *p;
    int nc;
    if (field_ptr(ctx, which, &p, &nc)) return fail(ctx, WCNS_E_ARG, "unknown field");
    HBlock &b = ctx->blocks[nb];
    CK(hipStreamSynchronize(ctx->stream));
    for (int m = 0; m < nc; m++)
        TRY(copy_sync(ctx, out + m * b.g3, p + m * ctx->total + b.db.off, 8 * b.g3, hipMemcpyDeviceToHost));
    return WCNS_OK;
}

int wcns_set_field(wcns_ctx *ctx, int nb, int which, const double *in)
{
    if (!ctx || !ctx->setup_done || !in) return WCNS_E_STATE;
    if (nb < 0 || nb >= (int)ctx->blocks.size()) return WCNS_E_ARG;
    ctx->pt_fresh = false;
    double *p;
    int nc;
    if (field_ptr(ctx, which, &p, &nc)) return fail(ctx, WCNS_E_ARG, "unknown field");
    HBlock &b = ctx->blocks[nb];
    CK(hipStreamSynchronize(ctx->stream));
    for (int m = 0; m < nc; m++)
        TRY(copy_sync(ctx, p + m * ctx->total + b.db.off, in + m * b.g3, 8 * b.g3, hipMemcpyHostToDevice));
    // a primitive goes to both buffers, as wcns_set_prim does: the update
    // writes only interior nodes into the other buffer
    if (which >= 0 && which <= 4)
        TRY(copy_sync(ctx, ctx->prim[ctx->cur ^ 1][which] + b.db.off, in, 8 * b.g3, hipMemcpyHostToDevice));
    return WCNS_OK;
}

// ---------------------------------------------------------------- stages
int wcns_boundary_all(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    ctx->pt_fresh = false;
    DFields F = ctx->fields();
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    Scope sc(ctx, WCNS_T_BC);
    if (ctx->surf_levels) {
        for (size_t l = 0; l + 1 < ctx->bc_lvl.size(); l++)
            if (launch_bc_level(ctx->d_regs, ctx->d_rwork + ctx->bc_lvl[l], ctx->bc_lvl[l + 1] - ctx->bc_lvl[l],
                                ctx->d_blocks, F, ctx->dp, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "bc launch");
        return WCNS_OK;
    }
    for (auto &b : ctx->blocks)
        for (int r : b.bcorder)
            if (launch_bc_region(b.regs[r].d, dbs.data(), F, ctx->dp, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "bc launch");
    return WCNS_OK;
}

int wcns_initial_q_dq_c_t(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    DFields F = ctx->fields();
    Scope sc(ctx, WCNS_T_INIT);
    if (launch_init_stage(ctx->grid[GRID_NODES], &ctx->grid[GRID_GHOST_T], F, ctx->dp, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "init launch");
    return WCNS_OK;
}

int wcns_timestep(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    DFields F = ctx->fields();
    Scope sc(ctx, WCNS_T_TIMESTEP);
    // with nvis = 1 and csrv > 0 the time-step kernel also forms the viscous radii
    if (launch_timestep(ctx->grid[GRID_NODES], F, ctx->dp, ctx->stream)) return fail(ctx, WCNS_E_HIP, "timestep launch");
    return WCNS_OK;
}

static int vis_meanflow(wcns_ctx *ctx, const DFields &F);

int wcns_r_h_s_vis(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (ctx->prm.nvis != 1) return WCNS_OK;
    DFields F = ctx->fields();
    if (ctx->dp.turb) {
        // RHS.f90:22: the turbulence model advances first; its update (nu~ on 2..n-1,
        // mu_t on 0..n) only touches the block's own nodes, so all blocks go together
        {
            Scope sc(ctx, WCNS_T_SA_RHS, 2);
            if (launch_sa_rhs(ctx->grid[GRID_NODES], ctx->grid[GRID_SA_RHS], F, ctx->sa, ctx->dp, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "SA rhs launch");
        }
        // lusgs_turbulence: forward then backward sweep per variable (independent
        // variables).  With the pipelined sweeps the operands go through the
        // plane-major pool's dQ slots, which hold nothing until the implicit
        // stage restages them: the unknown, the diagonal and un[0..2] in, the
        // unknown back.
        // Measured: on many blocks (64-block SA: 2.9 -> 1.15 ms per sweep) the
        // sweeps are bound by the lines the lanes fetch, and the staging pays;
        // on one block of one tile per CU they are latency-bound and it does
        // not, so the plane-major path is taken beyond four tiles per CU.
        const bool staged = ctx->sa_staged >= 0 ? ctx->sa_staged == 1 : ctx->plan.nwork > 4 * ctx->ncu;
        double *pm = ctx->tiled_sweeps && staged ? (double *)ctx->skew[0] : nullptr;
        for (int var = 0; var < ctx->dp.turb; var++) {
            if (pm) {
                Scope sc(ctx, WCNS_T_SWEEP_STAGE);
                const double *src[5] = {ctx->sa.dqke[var], ctx->sa.spec[var], ctx->sa.un[0], ctx->sa.un[1],
                                        ctx->sa.un[2]};
                const int slot[5] = {SK_A, SK_A + 1, SK_A + 2, SK_A + 3, SK_A + 4};
                if (launch_skew_stage_f64(ctx->grid[GRID_SKEW], ctx->d_skoff, src, slot, 5, pm, ctx->sktot,
                                          ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "staging the turbulence sweep operands");
            }
            for (int fwd = 1; fwd >= 0; fwd--) {
                Scope sc(ctx, fwd ? WCNS_T_SA_SWEEP_L : WCNS_T_SA_SWEEP_U);
                if (launch_sa_sweep(fwd, var, ctx->plan, ctx->sa, pm, ctx->sktot, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "turbulence sweep launch");
            }
            if (pm) {
                Scope sc(ctx, WCNS_T_SWEEP_STAGE);
                if (launch_unskew_f64(ctx->grid[GRID_SKEW], ctx->d_skoff, pm, ctx->sktot, SK_A, 1, ctx->sa.dqke[var],
                                      ctx->total, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "copying the turbulence sweep result");
            }
        }
        Scope sc(ctx, WCNS_T_SA_UPDATE);
        if (launch_sa_update(ctx->grid[GRID_SA_UPD], F, ctx->sa, ctx->dp, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "SA update launch");
    }
    return vis_meanflow(ctx, F);
}

// the mean-flow viscous terms of r_h_s_vis (mu_t as stored)
static int vis_meanflow(wcns_ctx *ctx, const DFields &F)
{
    {
        Scope sc(ctx, WCNS_T_VIS);
        if (launch_vis_node_deriv(ctx->grid[GRID_NODES], F, ctx->dn, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "vis deriv launch");
    }
    for (int dir = 0; dir < 3; dir++) {
        if (dir == 0 && ctx->idir_t) {
            {
                Scope st(ctx, WCNS_T_TRANSPOSE);
                if (stage_transposed(ctx, F, true)) return fail(ctx, WCNS_E_HIP, "transposing the viscous inputs");
            }
            const DFields FT = fields_t(ctx, F);
            {
                Scope sc(ctx, WCNS_T_VIS_I);
                if (launch_vis_faces(3, ctx->grid[GRID_VFACE], FT, ctx->dp, ctx->dn_t, ctx->fvb, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "vis launch");
            }
            {
                Scope sc(ctx, WCNS_T_VIS_ROWS);
                if (launch_vis_rows(3, ctx->grid[GRID_VROW], FT, ctx->dp, ctx->fvb, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "vis launch");
            }
            Scope st(ctx, WCNS_T_TRANSPOSE);
            if (add_transposed(ctx, F)) return fail(ctx, WCNS_E_HIP, "adding the i-direction viscous terms");
            continue;
        }
        if (dir > 0) {
            {
                Scope sc(ctx, WCNS_T_VIS_I + dir);
                if (launch_vis_faces(dir, ctx->grid[GRID_VFACE + dir], F, ctx->dp, ctx->dn, ctx->fvb, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "vis launch");
            }
            Scope sc(ctx, WCNS_T_VIS_ROWS);
            if (launch_vis_rows(dir, ctx->grid[GRID_VROW + dir], F, ctx->dp, ctx->fvb, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "vis launch");
            continue;
        }
        Scope sc(ctx, WCNS_T_VIS_I + dir);
        if (launch_rhs_vis_dir(dir, ctx->grid[GRID_RHS_VIS + dir], F, ctx->dp, ctx->dn, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "vis launch");
    }
    return WCNS_OK;
}

int wcns_r_h_s_invis(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    DFields F = ctx->fields();
    const bool var = inv_variant(ctx->dp);
    if (ctx->dp.nflux == 4) {
        Scope sc(ctx, WCNS_T_INV_I);
        for (auto &b : ctx->blocks)
            if (launch_pressure_ratios(b.db, F, ctx->stream)) return fail(ctx, WCNS_E_HIP, "pressure ratios");
    }
    for (int dir = 0; dir < 3; dir++) {
        if (dir == 0 && ctx->idir_t && !var) {
            {
                Scope st(ctx, WCNS_T_TRANSPOSE);
                if (stage_transposed(ctx, F, false)) return fail(ctx, WCNS_E_HIP, "transposing the inviscid inputs");
            }
            {
                Scope sc(ctx, WCNS_T_INV_I);
                if (launch_rhs_inv_dir(3, ctx->grid[GRID_RHS_INV_T], fields_t(ctx, F), ctx->dp, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "rhs launch");
            }
            Scope st(ctx, WCNS_T_TRANSPOSE);
            if (add_transposed(ctx, F)) return fail(ctx, WCNS_E_HIP, "adding the i-direction inviscid terms");
            continue;
        }
        Scope sc(ctx, WCNS_T_INV_I + dir);
        if (launch_rhs_inv_dir(dir, ctx->grid[GRID_RHS_INV + dir], F, ctx->dp, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "rhs launch");
    }
    return WCNS_OK;
}

static int communicate(wcns_ctx *ctx, int pv)
{
    Scope sc(ctx, pv ? WCNS_T_COMM_PV : WCNS_T_COMM_DQ);
    DFields F = ctx->fields();
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    int nvar = pv ? 6 : 5;
    if (launch_npp_gather(ctx->d_npp_pts, ctx->npp, ctx->d_blocks, F, ctx->d_npp_buf, nvar, !pv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp gather");
    if (ctx->surf_levels) {
        for (size_t l = 0; l + 1 < ctx->if_lvl.size(); l++)
            if (launch_avg2pm_level(ctx->d_regs, ctx->d_rwork + ctx->if_lvl[l], ctx->if_lvl[l + 1] - ctx->if_lvl[l],
                                    ctx->d_blocks, F, pv, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "avg2pm");
    } else {
        for (auto &g : ctx->iface)
            if (launch_avg2pm(g, dbs.data(), F, pv, ctx->stream)) return fail(ctx, WCNS_E_HIP, "avg2pm");
    }
    if (launch_npp_assign(ctx->d_grp_start, ctx->d_grp_n, ctx->d_npp_pts, ctx->d_write_ok, ctx->ngroups,
                          ctx->d_npp_buf, ctx->d_blocks, F, nvar, !pv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp assign");
    return WCNS_OK;
}

int wcns_communicate_dq(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    return communicate(ctx, 0);
}

int wcns_communicate_pv(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    ctx->pt_fresh = false;
    return communicate(ctx, 1);
}

int wcns_l_h_s(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    ctx->pt_fresh = false;
    DFields F = ctx->fields();
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    const int prec = F.dq32 ? 1 : 0;
    if (launch_dq_zero_faces(ctx->d_regs, ctx->d_rwork + ctx->zero_w0, ctx->zero_nw, ctx->d_blocks, F.rhs, nullptr,
                             F.total, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "zeroing the physical faces");
    if (ctx->tiled_sweeps) {
        {
            Scope sc(ctx, WCNS_T_SWEEP_STAGE);
            if (launch_skew_stage_all(SKEW_STEP, ctx->grid[GRID_SKEW], ctx->d_skoff, F, prec, ctx->skew[prec],
                                      ctx->sktot, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "staging the sweep operands");
        }
        for (int kind = 0; kind < 4; kind++) {
            Scope sc(ctx, WCNS_T_SWEEP_LU_L + kind);
            if (launch_sweep_batched(kind, ctx->plan, F, ctx->dp, prec, ctx->skew[prec], ctx->sktot, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "sweeps");
        }
        {
            Scope sc(ctx, WCNS_T_SWEEP_STAGE);
            if (launch_unskew_dq(ctx->grid[GRID_SKEW], ctx->d_skoff, prec, ctx->skew[prec], ctx->sktot, F.dq, F.dq32,
                                 F.total, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "copying the sweeps' dQ");
        }
    } else {
        for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
            const DBlock &db = ctx->blocks[nb].db;
            for (int kind = 0; kind < 4; kind++) {
                Scope sc(ctx, WCNS_T_SWEEP_LU_L + kind, db.ni + db.nj + db.nk - 2);
                if (launch_sweep(kind, db, F, ctx->dp, ctx->stream)) return fail(ctx, WCNS_E_HIP, "sweeps");
            }
        }
    }
    {
        Scope sc(ctx, WCNS_T_UPDATE);
        if (launch_dq_zero_faces(ctx->d_regs, ctx->d_rwork + ctx->zero_w0, ctx->zero_nw, ctx->d_blocks, F.dq, F.dq32,
                                 F.total, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "zeroing the physical faces");
        if (launch_fill_u64(ctx->d_nflag, (long)ctx->blocks.size(), 0ull, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "counter reset");
        if (launch_update(ctx->grid[GRID_NODES], F, ctx->dp, ctx->d_flag, ctx->d_nflag, ctx->flag_cap, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "update");
        if (launch_repair(ctx->d_blocks, (int)ctx->blocks.size(), F, ctx->dp, ctx->d_flag, ctx->d_nflag,
                          ctx->flag_cap, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "repair");
    }
    ctx->cur ^= 1;
    return WCNS_OK;
}

// residual_curve of the current dq on the stream; push appends it to the
// on-device history (the per-step monitor of Main.f90:109), no host wait
static int launch_monitor(wcns_ctx *ctx, bool push)
{
    Scope sc(ctx, WCNS_T_RESIDUAL, 3);
    if (launch_residual_all(ctx->grid[GRID_RESID], ctx->fields(), ctx->d_partial, ctx->d_res_wg0,
                            (int)ctx->blocks.size(), ctx->d_resblk, ctx->d_blocks, ctx->d_res_out,
                            push ? ctx->d_hist : nullptr, ctx->d_hist_n, ctx->hist_cap, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "residual monitor launch");
    return WCNS_OK;
}

static void entry_out(const double *e, double *rms, double *resmax, int *loc)
{
    if (rms) *rms = e[0];
    if (resmax) *resmax = e[1];
    if (loc) for (int m = 0; m < 5; m++) loc[m] = (int)e[2 + m];
}

int wcns_residual(wcns_ctx *ctx, double *rms, double *resmax, int loc[5])
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    TRY(launch_monitor(ctx, false));
    double e[8];
    CK(hipMemcpyAsync(e, ctx->d_res_out, sizeof e, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    TRY(check_counters(ctx));
    entry_out(e, rms, resmax, loc);
    return WCNS_OK;
}

int wcns_residual_noalloc(wcns_ctx *ctx, double *rms, double *resmax, int loc[5])
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (ctx->res_nwg > NOALLOC_WGS || (int)ctx->blocks.size() > NOALLOC_BLOCKS)
        return fail(ctx, WCNS_E_ARG, "wcns_residual_noalloc: the case needs " + std::to_string(ctx->res_nwg) +
                                         " monitor workgroups and " + std::to_string(ctx->blocks.size()) +
                                         " blocks; its scratch holds at most " + std::to_string(NOALLOC_WGS) +
                                         " and " + std::to_string(NOALLOC_BLOCKS));
    CK(hipSetDevice(ctx->device));
    double *partial = ctx->noalloc_scratch;
    double *resblk = partial + 3 * (size_t)NOALLOC_WGS;
    double *res_out = resblk + 3 * (size_t)NOALLOC_BLOCKS;
    {
        Scope sc(ctx, WCNS_T_RESIDUAL, 3);
        if (launch_residual_all(ctx->grid[GRID_RESID], ctx->fields(), partial, ctx->d_res_wg0, (int)ctx->blocks.size(),
                                resblk, ctx->d_blocks, res_out, nullptr, nullptr, 0, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "residual monitor launch");
    }
    double e[8];
    CK(hipMemcpyAsync(e, res_out, sizeof e, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    TRY(check_counters(ctx));
    entry_out(e, rms, resmax, loc);
    return WCNS_OK;
}

int wcns_residual_history(wcns_ctx *ctx, int max, double *rms, double *resmax, int *loc, long long *count)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (max < 0 || (max > 0 && !rms && !resmax && !loc)) return WCNS_E_ARG;
    unsigned long long n = 0;
    CK(hipMemcpyAsync(&n, ctx->d_hist_n, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (count) *count = (long long)n;
    const long want = std::min<long>(std::min<long>(max, (long)n), ctx->hist_cap);
    if (want <= 0) return WCNS_OK;
    std::vector<double> h(8 * (size_t)ctx->hist_cap);
    CK(hipMemcpyAsync(h.data(), ctx->d_hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    for (long q = 0; q < want; q++) {
        const unsigned long long idx = n - (unsigned long long)want + (unsigned long long)q;
        const double *e = h.data() + 8 * (size_t)(idx % (unsigned long long)ctx->hist_cap);
        entry_out(e, rms ? rms + q : nullptr, resmax ? resmax + q : nullptr, loc ? loc + 5 * q : nullptr);
    }
    return WCNS_OK;
}

}  // extern "C"

// The launches of one pseudo-time step (Main.f90:85-109), no host wait.
static int step_once(wcns_ctx *ctx)
{
    TRY(wcns_boundary_all(ctx));
    TRY(wcns_initial_q_dq_c_t(ctx));
    TRY(wcns_timestep(ctx));
    if (ctx->prm.nvis == 1) {
        TRY(wcns_r_h_s_vis(ctx));
        TRY(wcns_communicate_dq(ctx));
    }
    TRY(wcns_r_h_s_invis(ctx));
    TRY(wcns_communicate_dq(ctx));
    TRY(wcns_l_h_s(ctx));
    TRY(wcns_communicate_pv(ctx));
    return launch_monitor(ctx, true);
}

// Captures two consecutive steps (the primitive buffers alternate, so a pair
// returns to the buffer it started from) into an executable graph.
static int capture_pair(wcns_ctx *ctx, hipGraphExec_t *out)
{
    const bool tm = ctx->timing;
    const int c0 = ctx->cur;
    ctx->timing = false;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        ctx->timing = tm;
        return fail(ctx, WCNS_E_HIP, "hipStreamBeginCapture failed");
    }
    int e = step_once(ctx);
    if (!e) e = step_once(ctx);
    hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
    ctx->timing = tm;
    ctx->cur = c0;
    ctx->pt_fresh = false;
    if (e || ee != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return e ? e : fail(ctx, WCNS_E_HIP, std::string("stream capture: ") + hipGetErrorString(ee));
    }
    hipGraphExec_t x = nullptr;
    hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) return fail(ctx, WCNS_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
    *out = x;
    return WCNS_OK;
}

static void drop_graphs(wcns_ctx *ctx)
{
    for (int c = 0; c < 2; c++)
        if (ctx->graph[c]) { hipGraphExecDestroy(ctx->graph[c]); ctx->graph[c] = nullptr; }
}

extern "C" {

int wcns_step_graph(wcns_ctx *ctx, int nsteps)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (nsteps < 0) return WCNS_E_ARG;
    if (ctx->graph_prec != ctx->lhs_prec) { drop_graphs(ctx); ctx->graph_prec = ctx->lhs_prec; }
    int done = 0;
    for (; nsteps - done >= 2; done += 2) {
        hipGraphExec_t &ge = ctx->graph[ctx->cur];
        if (!ge) TRY(capture_pair(ctx, &ge));
        CK(hipGraphLaunch(ge, ctx->stream));
    }
    if (done < nsteps) TRY(step_once(ctx));
    CK(hipStreamSynchronize(ctx->stream));
    return check_counters(ctx);
}

int wcns_step(wcns_ctx *ctx, int nsteps, double *rms_hist)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (nsteps < 0) return WCNS_E_ARG;
    // the monitor of every step goes to the device history; the host reads
    // it once per hist_cap steps at most (before the ring wraps)
    for (int s0 = 0; s0 < nsteps; s0 += ctx->hist_cap) {
        const int ns = std::min(nsteps - s0, ctx->hist_cap);
        for (int s = 0; s < ns; s++) TRY(step_once(ctx));
        CK(hipStreamSynchronize(ctx->stream));
        TRY(check_counters(ctx));
        if (rms_hist) TRY(wcns_residual_history(ctx, ns, rms_hist + s0, nullptr, nullptr, nullptr));
    }
    return WCNS_OK;
}

// force + force_dif of every no-slip wall on the stream into d_force_last, and
// with push also into the force history; no host wait
static int launch_forces(wcns_ctx *ctx, bool push)
{
    if (ctx->prm.sref == 0.0 || ctx->prm.lfref == 0.0) return fail(ctx, WCNS_E_ARG, "sref and lfref must be non-zero");
    DFields F = ctx->fields();
    const double ref[3] = {ctx->prm.xref, ctx->prm.yref, ctx->prm.zref};
    for (size_t w = 0; w < ctx->walls.size(); w++) {
        const auto &wr = ctx->walls[w];
        if (launch_force_region(wr.d, ctx->blocks[wr.d.nb].db, F, ctx->dp, wr.xyz, ctx->d_force_pn,
                                ctx->d_force_part, ctx->d_force_out + 6 * w, ref, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "force launch failed");
    }
    // PostProcess.f90:1458-1477 (force): the angles as init_inflow converts them
    const wcns_params &p = ctx->prm;
    const double pai = 3.14159265;   // Globalvar.f90:236
    const double att = p.attack * pai / 180.0, ssl = p.sideslip * pai / 180.0;
    ForceCoef c;
    c.wf = (double)(2 - p.nwholefield);
    c.scf = 1.0 / p.sref;
    c.vcm = 1.0 / (p.sref * p.lfref);
    c.sina = sin(att); c.cosa = cos(att); c.sinb = sin(ssl); c.cosb = cos(ssl);
    if (launch_force_coef(ctx->d_force_out, (int)ctx->walls.size(), c, ctx->d_force_last, push ? ctx->d_fhist : nullptr,
                          ctx->d_fhist_n, ctx->fhist_cap, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "force coefficient launch failed");
    return WCNS_OK;
}

int wcns_force(wcns_ctx *ctx, double out[9])
{
    if (!ctx || !out) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    TRY(launch_forces(ctx, false));
    CK(hipMemcpyAsync(out, ctx->d_force_last, 9 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_force_record(wcns_ctx *ctx)
{
    if (!ctx) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    return launch_forces(ctx, true);
}

int wcns_force_history(wcns_ctx *ctx, int max, double *out, long long *count)
{
    if (!ctx || max < 0 || (max > 0 && !out)) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    unsigned long long n = 0;
    CK(hipMemcpyAsync(&n, ctx->d_fhist_n, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (count) *count = (long long)n;
    const long want = std::min<long>(std::min<long>(max, (long)n), ctx->fhist_cap);
    if (want <= 0) return WCNS_OK;
    std::vector<double> h(9 * (size_t)ctx->fhist_cap);
    CK(hipMemcpyAsync(h.data(), ctx->d_fhist, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    for (long q = 0; q < want; q++) {
        const unsigned long long idx = n - (unsigned long long)want + (unsigned long long)q;
        memcpy(out + 9 * q, h.data() + 9 * (size_t)(idx % (unsigned long long)ctx->fhist_cap), 9 * sizeof(double));
    }
    return WCNS_OK;
}

// ------------------------------------------------------------ restart files
// Fortran unformatted sequential records, as the reference's flang/gfortran
// build writes them: a 4-byte byte count, the data, the count again.
static bool rec_write(FILE *f, const void *d, size_t n)
{
    if (n > 0x7fffffffu) return false;
    const uint32_t m = (uint32_t)n;
    return fwrite(&m, 4, 1, f) == 1 && (n == 0 || fwrite(d, 1, n, f) == n) && fwrite(&m, 4, 1, f) == 1;
}

static bool rec_read(FILE *f, std::vector<unsigned char> &buf)
{
    uint32_t m = 0, m2 = 0;
    if (fread(&m, 4, 1, f) != 1) return false;
    buf.resize(m);
    if (m && fread(buf.data(), 1, m, f) != m) return false;
    return fread(&m2, 4, 1, f) == 1 && m2 == m;
}

// a whole G3 field of block nb into host memory
static int fetch(wcns_ctx *ctx, int nb, int which, std::vector<double> &a)
{
    a.resize(ctx->blocks[nb].g3);
    return wcns_get_field(ctx, nb, which, a.data());
}

int wcns_save_flow(wcns_ctx *ctx, const char *flowname, int nout, double wholetime)
{
    if (!ctx || !flowname) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    FILE *f = fopen(flowname, "wb");
    if (!f) return fail(ctx, WCNS_E_ARG, std::string("cannot open ") + flowname);
    unsigned char head[4 + 7 * 8];
    const int32_t no = nout;
    const double hv[7] = {wholetime, ctx->dp.roo, ctx->dp.uoo, ctx->dp.voo, ctx->dp.woo, ctx->dp.poo, ctx->prm.moo};
    memcpy(head, &no, 4);
    memcpy(head + 4, hv, sizeof hv);
    bool ok = rec_write(f, head, sizeof head);
    std::vector<double> fld[6], rec;
    for (size_t nb = 0; ok && nb < ctx->blocks.size(); nb++) {
        const HBlock &b = ctx->blocks[nb];
        for (int m = 0; m < 6; m++) {
            int e = fetch(ctx, (int)nb, m, fld[m]);
            if (e) { fclose(f); return e; }
        }
        rec.clear();
        for (int k = 1; k <= b.nk; k++) for (int j = 1; j <= b.nj; j++) for (int i = 1; i <= b.ni; i++) {
            const long q = gidx(b.db, i, j, k) - b.db.off;
            for (int m = 0; m < 6; m++) rec.push_back(fld[m][q]);
        }
        ok = rec_write(f, rec.data(), rec.size() * sizeof(double));
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(ctx, WCNS_E_ARG, std::string("writing ") + flowname + " failed");
    if (ctx->qke) {
        const std::string tn = std::string(flowname) + ".turb";
        FILE *g = fopen(tn.c_str(), "wb");
        if (!g) return fail(ctx, WCNS_E_ARG, "cannot open " + tn);
        const int nl = ctx->qkw ? 2 : 1;
        std::vector<double> qk, qw, vt;
        for (size_t nb = 0; ok && nb < ctx->blocks.size(); nb++) {
            const HBlock &b = ctx->blocks[nb];
            int e = fetch(ctx, (int)nb, 26, qk);
            if (!e && nl == 2) e = fetch(ctx, (int)nb, 28, qw);
            if (!e) e = fetch(ctx, (int)nb, 15, vt);
            if (e) { fclose(g); return e; }
            rec.clear();
            for (int k = 1; k <= b.nk; k++) for (int j = 1; j <= b.nj; j++) for (int i = 1; i <= b.ni; i++) {
                const long q = gidx(b.db, i, j, k) - b.db.off;
                rec.push_back(qk[q]);
                if (nl == 2) rec.push_back(qw[q]);
                rec.push_back(vt[q]);
            }
            ok = rec_write(g, rec.data(), rec.size() * sizeof(double));
        }
        ok = (fclose(g) == 0) && ok;
        if (!ok) return fail(ctx, WCNS_E_ARG, "writing " + tn + " failed");
    }
    return WCNS_OK;
}

// Tecplot binary (version 102) as the reference's mod_tecios writes it
// (parallel.f90:1-173): every character as an int32, strings 0-terminated,
// zone and end-of-header markers as float32, one ordered block-packed zone
// per block or boundary region, all values as doubles.
namespace {
struct TecFile {
    FILE *f;
    bool ok = true;
    void raw(const void *p, size_t n) { ok = ok && fwrite(p, 1, n, f) == n; }
    void i32(int32_t v) { raw(&v, 4); }
    void f32(float v) { raw(&v, 4); }
    void str(const std::string &t)
    {
        for (unsigned char ch : t) i32(ch == ',' ? 0 : ch);   // tecio_ini turns the commas into separators
        i32(0);
    }
};
}  // namespace

int wcns_tecplot(wcns_ctx *ctx, const char *path, int kind)
{
    if (!ctx || !path || kind < WCNS_TEC_SURFACE || kind > WCNS_TEC_VOLUME) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    const int nl = ctx->prm.nlamtur, nvis = ctx->prm.nvis;
    std::string names;
    int nv;
    if (kind == WCNS_TEC_WALL) {
        nv = nvis == 1 ? 9 : 4;
        names = nvis == 1 ? "X,Y,Z,cp,cfx,cfy,cfz,cf,qw" : "X,Y,Z,cp";
    } else {
        nv = 10;
        names = "X,Y,Z,R,U,V,W,P,T,M";
        if (nl >= 0) {
            nv += nl + 1;
            names += nl == 0 ? ",muT" : (nl == 1 ? ",vT1,muT" : ",vT1,vT2,muT");
        }
    }
    // zones: (block, region or -1 for the whole block)
    std::vector<std::pair<int, int>> zones;
    for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
        if (kind == WCNS_TEC_VOLUME) { zones.push_back({(int)nb, -1}); continue; }
        const auto &regs = ctx->blocks[nb].regs;
        for (size_t nr = 0; nr < regs.size(); nr++) {
            const int bt = regs[nr].bctype;
            const bool draw = kind == WCNS_TEC_WALL ? (bt == 2 || bt == 20 || bt == 21)
                                                   : (bt / 10 != 7 && bt != 4 && bt != 5 && bt > 0);
            if (draw) zones.push_back({(int)nb, (int)nr});
        }
    }
    auto box = [&](const std::pair<int, int> &z, int lo[3], int hi[3]) {
        const HBlock &b = ctx->blocks[z.first];
        if (z.second < 0) {
            lo[0] = lo[1] = lo[2] = 1;
            hi[0] = b.ni; hi[1] = b.nj; hi[2] = b.nk;
        } else {
            const DRegion &d = b.regs[z.second].d;
            for (int m = 0; m < 3; m++) { lo[m] = d.s_lo[m]; hi[m] = d.s_hi[m]; }
        }
    };
    TecFile T{fopen(path, "wb")};
    if (!T.f) return fail(ctx, WCNS_E_ARG, std::string("cannot open ") + path);
    T.raw("#!TDV102", 8);
    T.i32(1);
    T.str("WCNS_SOLVER");
    T.i32(nv);
    T.str(names);
    for (auto &z : zones) {
        int lo[3], hi[3];
        box(z, lo, hi);
        std::string zn = "BLK" + std::to_string(z.first + 1);
        if (z.second >= 0)
            zn += "BC" + std::to_string(z.second + 1) + "T" +
                  std::to_string(ctx->blocks[z.first].regs[z.second].bctype);
        T.f32(299.0f);
        T.str(zn);
        T.i32(-1);      // zone colour: Tecplot's choice
        T.i32(0);       // ordered
        T.i32(0);       // block packing
        T.i32(0);       // values at the nodes
        T.i32(0);       // no face connections
        T.i32(hi[0] - lo[0] + 1);
        T.i32(hi[1] - lo[1] + 1);
        T.i32(hi[2] - lo[2] + 1);
        T.i32(0);       // no auxiliary data
    }
    T.f32(357.0f);
    const double gama = ctx->prm.gama;
    std::vector<double> fl[6], qk, qw2, vt, wp, col;
    for (auto &z : zones) {
        const HBlock &b = ctx->blocks[z.first];
        int lo[3], hi[3];
        box(z, lo, hi);
        T.f32(299.0f);
        for (int m = 0; m < nv; m++) T.i32(2);   // doubles
        T.i32(0);       // no shared variables
        T.i32(-1);      // no shared connectivity
        auto each = [&](auto fn) {
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) fn(i, j, k);
        };
        auto grid = [&](const std::vector<double> &a) {
            col.clear();
            each([&](int i, int j, int k) {
                col.push_back(a[(size_t)(i - 1) + (size_t)b.ni * ((size_t)(j - 1) + (size_t)b.nj * (k - 1))]);
            });
            T.raw(col.data(), 8 * col.size());
        };
        grid(b.x);
        grid(b.y);
        grid(b.z);
        if (kind == WCNS_TEC_WALL) {
            const DRegion &d = b.regs[z.second].d;
            double *dw = nullptr;
            TRY(dalloc(ctx, &dw, 6 * (size_t)d.npts));
            const double cpk = 1.0 / ((gama - 1.0) * ctx->prm.moo * ctx->prm.moo);
            int e = launch_wall_plot(d, b.db, ctx->fields(), ctx->dp, cpk, dw, ctx->stream);
            wp.resize(6 * (size_t)d.npts);
            if (!e) e = copy_sync(ctx, wp.data(), dw, 8 * wp.size(), hipMemcpyDeviceToHost);
            hipFree(dw);
            if (e) { fclose(T.f); return fail(ctx, WCNS_E_HIP, "wall surface values"); }
            // region_point numbers the face nodes i fastest, as the file does
            for (int c = 0; c < (nvis == 1 ? 6 : 1); c++) T.raw(wp.data() + (size_t)c * d.npts, 8 * (size_t)d.npts);
            continue;
        }
        for (int m = 0; m < 6; m++) {
            int e = fetch(ctx, z.first, m, fl[m]);
            if (e) { fclose(T.f); return e; }
        }
        if (nl >= 1) {
            int e = fetch(ctx, z.first, 26, qk);
            if (!e && nl == 2) e = fetch(ctx, z.first, 28, qw2);
            if (!e) e = fetch(ctx, z.first, 15, vt);
            if (e) { fclose(T.f); return e; }
        }
        // velocities of a wall surface from the node inside it (PostProcess.f90:233-249)
        int ip[3] = {0, 0, 0};
        if (kind == WCNS_TEC_SURFACE) {
            const DRegion &d = b.regs[z.second].d;
            const int bt = b.regs[z.second].bctype;
            if (bt == 2 || bt == 20 || bt == 21) ip[d.s_nd - 1] = d.s_lr;
        }
        auto node = [&](const std::vector<double> &a, int i, int j, int k) { return a[gidx(b.db, i, j, k) - b.db.off]; };
        auto field = [&](const std::vector<double> &a, bool inner) {
            col.clear();
            each([&](int i, int j, int k) {
                col.push_back(inner ? node(a, i - ip[0], j - ip[1], k - ip[2]) : node(a, i, j, k));
            });
            T.raw(col.data(), 8 * col.size());
        };
        field(fl[0], false);
        field(fl[1], true);
        field(fl[2], true);
        field(fl[3], true);
        field(fl[4], false);
        field(fl[5], false);
        // M = sqrt(u2 / c2) of the node's own values (PostProcess.f90:686-689)
        col.clear();
        each([&](int i, int j, int k) {
            const double r = node(fl[0], i, j, k), u = node(fl[1], i, j, k), v = node(fl[2], i, j, k);
            const double w = node(fl[3], i, j, k), p = node(fl[4], i, j, k);
            const double c2 = gama * p / r, u2 = u * u + v * v + w * w;
            col.push_back(sqrt(u2 / c2));
        });
        T.raw(col.data(), 8 * col.size());
        if (nl >= 1) {
            field(qk, false);
            if (nl == 2) field(qw2, false);
            field(vt, false);
        }
    }
    const bool ok = (fclose(T.f) == 0) && T.ok;
    if (!ok) return fail(ctx, WCNS_E_ARG, std::string("writing ") + path + " failed");
    return WCNS_OK;
}

int wcns_load_flow(wcns_ctx *ctx, const char *flowname, int t_from_file, int *nout, double *wholetime)
{
    if (!ctx || !flowname) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    FILE *f = fopen(flowname, "rb");
    if (!f) return fail(ctx, WCNS_E_ARG, std::string("cannot open ") + flowname);
    std::vector<unsigned char> buf;
    if (!rec_read(f, buf) || buf.size() != 4 + 7 * 8) {
        fclose(f);
        return fail(ctx, WCNS_E_ARG, std::string(flowname) + ": not a flow file of this case (header)");
    }
    int32_t no;
    double hv[7];
    memcpy(&no, buf.data(), 4);
    memcpy(hv, buf.data() + 4, sizeof hv);
    std::vector<double> fld[6];
    for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
        const HBlock &b = ctx->blocks[nb];
        const size_t n = (size_t)b.ni * b.nj * b.nk;
        if (!rec_read(f, buf) || buf.size() != n * 6 * sizeof(double)) {
            fclose(f);
            return fail(ctx, WCNS_E_ARG, std::string(flowname) + ": block " + std::to_string(nb + 1) +
                                             " does not match the case");
        }
        const double *d = reinterpret_cast<const double *>(buf.data());
        for (int m = 0; m < 6; m++) {
            int e = fetch(ctx, (int)nb, m, fld[m]);
            if (e) { fclose(f); return e; }
        }
        size_t o = 0;
        for (int k = 1; k <= b.nk; k++) for (int j = 1; j <= b.nj; j++) for (int i = 1; i <= b.ni; i++, o++) {
            const long q = gidx(b.db, i, j, k) - b.db.off;
            for (int m = 0; m < 6; m++) fld[m][q] = d[6 * o + m];
        }
        // read_flow_file (PreProcess.f90:1542-1616) sets the interior; the
        // temperature then follows initial_t as at set-up (PreProcess.f90:62),
        // or with t_from_file stays as the file holds it
        int e = wcns_set_prim(ctx, (int)nb, fld[0].data(), fld[1].data(), fld[2].data(), fld[3].data(),
                              fld[4].data());
        if (!e && (t_from_file || ctx->prm.nvis != 1)) e = wcns_set_field(ctx, (int)nb, 5, fld[5].data());
        if (e) { fclose(f); return e; }
    }
    fclose(f);
    if (ctx->qke) {
        const std::string tn = std::string(flowname) + ".turb";
        FILE *g = fopen(tn.c_str(), "rb");
        if (!g) return fail(ctx, WCNS_E_ARG, "cannot open " + tn);
        const int nl = ctx->qkw ? 2 : 1;
        std::vector<double> qk, qw, vt;
        for (size_t nb = 0; nb < ctx->blocks.size(); nb++) {
            const HBlock &b = ctx->blocks[nb];
            const size_t n = (size_t)b.ni * b.nj * b.nk;
            if (!rec_read(g, buf) || buf.size() != n * (nl + 1) * sizeof(double)) {
                fclose(g);
                return fail(ctx, WCNS_E_ARG, tn + ": block " + std::to_string(nb + 1) + " does not match the case");
            }
            const double *d = reinterpret_cast<const double *>(buf.data());
            int e = fetch(ctx, (int)nb, 26, qk);
            if (!e && nl == 2) e = fetch(ctx, (int)nb, 28, qw);
            if (!e) e = fetch(ctx, (int)nb, 15, vt);
            if (e) { fclose(g); return e; }
            size_t o = 0;
            for (int k = 1; k <= b.nk; k++) for (int j = 1; j <= b.nj; j++) for (int i = 1; i <= b.ni; i++, o++) {
                const long q = gidx(b.db, i, j, k) - b.db.off;
                qk[q] = d[(nl + 1) * o];
                if (nl == 2) qw[q] = d[(nl + 1) * o + 1];
                vt[q] = d[(nl + 1) * o + nl];
            }
            e = wcns_set_field(ctx, (int)nb, 26, qk.data());
            if (!e && nl == 2) e = wcns_set_field(ctx, (int)nb, 28, qw.data());
            if (!e) e = wcns_set_field(ctx, (int)nb, 15, vt.data());
            if (e) { fclose(g); return e; }
        }
        fclose(g);
    }
    if (nout) *nout = no;
    if (wholetime) *wholetime = hv[0];
    return WCNS_OK;
}

int wcns_get_counter(wcns_ctx *ctx, int which, long long *value)
{
    if (!ctx || !value || which < 0 || which > 2) return WCNS_E_ARG;
    if (!ctx->setup_done) return WCNS_E_STATE;
    unsigned long long c[8];
    CK(hipStreamSynchronize(ctx->stream));
    TRY(copy_sync(ctx, c, ctx->counters, sizeof c, hipMemcpyDeviceToHost));
    *value = (long long)c[which];
    return WCNS_OK;
}

int wcns_sweep_layout(wcns_ctx *ctx, int *tj, int *tk, int *sgs_paired)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (tj) *tj = ctx->tile_shape.empty() ? 0 : ctx->tile_shape[0].first;
    if (tk) *tk = ctx->tile_shape.empty() ? 0 : ctx->tile_shape[0].second;
    if (sgs_paired) *sgs_paired = ctx->plan.sgs_paired;
    return WCNS_OK;
}

int wcns_autotune(wcns_ctx *ctx, int reps, int *tj, int *tk, double *ms)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (!ctx->tiled_sweeps) return fail(ctx, WCNS_E_UNSUPPORTED, "the autotuner tunes the pipelined sweeps");
    reps = std::max(1, std::min(reps, 20));
    static const int cand[][2] = {{16, 4}, {8, 8}, {32, 2}, {64, 1}, {4, 16}, {8, 4}, {16, 2}, {32, 1}};
    DFields F = ctx->fields();
    const int prec = F.dq32 ? 1 : 0;
    // the sweeps read the plane-major copy of the current step's operands
    if (launch_skew_stage_all(SKEW_STEP, ctx->grid[GRID_SKEW], ctx->d_skoff, F, prec, ctx->skew[prec], ctx->sktot,
                              ctx->stream))
        return fail(ctx, WCNS_E_HIP, "staging the sweep operands");
    const auto saved = ctx->tile_shape;
    const int saved_pair = ctx->plan.sgs_paired;
    double best = 1e300;
    int bj = 0, bk = 0, bpair = saved_pair;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    // ms per repetition of sweep kinds k0..k1
    auto time_kinds = [&](int k0, int k1, float &ms_out) -> int {
        CK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; r++)
            for (int kind = k0; kind <= k1; kind++)
                if (launch_sweep_batched(kind, ctx->plan, F, ctx->dp, prec, ctx->skew[prec], ctx->sktot, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "sweeps");
        CK(hipEventRecord(e1, ctx->stream));
        CK(hipEventSynchronize(e1));
        float t = 0.f;
        CK(hipEventElapsedTime(&t, e0, e1));
        ms_out = t / reps;
        return check_counters(ctx);
    };
    for (auto &c : cand) {
        ctx->tile_shape.clear();
        for (auto &b : ctx->blocks) ctx->tile_shape.push_back({std::min(c[0], b.nj), std::min(c[1], b.nk)});
        TRY(build_sweep_plan(ctx));
        for (int kind = 0; kind < 4; kind++)      // warm-up
            if (launch_sweep_batched(kind, ctx->plan, F, ctx->dp, prec, ctx->skew[prec], ctx->sktot, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "sweeps");
        float t_lu = 0.f;
        TRY(time_kinds(0, 1, t_lu));
        // the SGS pair with seven wavefronts per tile, and with four (two per CU)
        for (int pair = 0; pair < 2; pair++) {
            if (ctx->sgs_forced && pair != saved_pair) continue;
            ctx->plan.sgs_paired = pair;
            float t_sgs = 0.f;
            if (launch_sweep_batched(2, ctx->plan, F, ctx->dp, prec, ctx->skew[prec], ctx->sktot, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "sweeps");
            TRY(time_kinds(2, 3, t_sgs));
            if (t_lu + t_sgs < best) { best = t_lu + t_sgs; bj = c[0]; bk = c[1]; bpair = pair; }
        }
        ctx->plan.sgs_paired = saved_pair;
    }
    ctx->plan.sgs_paired = bpair;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->tile_shape.clear();
    if (bj) for (auto &b : ctx->blocks) ctx->tile_shape.push_back({std::min(bj, b.nj), std::min(bk, b.nk)});
    else ctx->tile_shape = saved;
    TRY(build_sweep_plan(ctx));
    drop_graphs(ctx);               // captured graphs hold the old plan
    if (tj) *tj = bj;
    if (tk) *tk = bk;
    if (ms) *ms = best;
    return WCNS_OK;
}

int wcns_timing(wcns_ctx *ctx, int on)
{
    if (!ctx) return WCNS_E_ARG;
    CK(hipStreamSynchronize(ctx->stream));
    ctx->recs.clear();
    ctx->pool_used = 0;
    ctx->timing = on != 0;
    return WCNS_OK;
}

int wcns_timing_get(wcns_ctx *ctx, int slot, double *total_ms, long long *launches)
{
    if (!ctx || !total_ms || !launches || slot < 0 || slot >= WCNS_T_NSLOT) return WCNS_E_ARG;
    CK(hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    long long n = 0;
    for (auto &r : ctx->recs) {
        if (r.slot != slot) continue;
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        n += r.launches;
    }
    *total_ms = t;
    *launches = n;
    return WCNS_OK;
}

int wcns_sync(wcns_ctx *ctx)
{
    if (!ctx) return WCNS_E_ARG;
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_region_images(wcns_ctx *ctx, int nb, int nr, int *img, int *jmg, int *kmg, int box[6])
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (nb < 0 || nb >= (int)ctx->blocks.size() || nr < 0 || nr >= (int)ctx->blocks[nb].regs.size())
        return WCNS_E_ARG;
    const DRegion &g = ctx->blocks[nb].regs[nr].d;
    for (int m = 0; m < 3; m++) { box[m] = g.s_lo[m]; box[3 + m] = g.s_hi[m]; }
    if (g.bctype >= 0) return 1;
    int o = 0;
    for (int k = g.s_lo[2]; k <= g.s_hi[2]; k++)
        for (int j = g.s_lo[1]; j <= g.s_hi[1]; j++)
            for (int i = g.s_lo[0]; i <= g.s_hi[0]; i++) {
                int it, jt, kt;
                region_image(g, i, j, k, it, jt, kt);
                long q = (i - g.s_lo[0]) + (long)(g.s_hi[0] - g.s_lo[0] + 1) *
                                               ((j - g.s_lo[1]) + (long)(g.s_hi[1] - g.s_lo[1] + 1) * (k - g.s_lo[2]));
                img[q] = it; jmg[q] = jt; kmg[q] = kt;
                o++;
            }
    return WCNS_OK;
}

int wcns_multipoint(wcns_ctx *ctx, int *ngroups, int *npts, int *pts, int cap)
{
    if (!ctx || !ctx->setup_done || !ngroups) return WCNS_E_STATE;
    *ngroups = ctx->ngroups;
    for (int g = 0; g < ctx->ngroups; g++) if (npts) npts[g] = ctx->h_grp_n[g];
    int n = (int)ctx->h_npp_pts.size();
    for (int a = 0; a < n && a < cap; a++) {
        int v = ctx->h_npp_pts[a];
        pts[a] = (a % 4 == 0) ? v + 1 : v;     // block 1-based like the reference
    }
    return n / 4;
}


// ------------------------------------------------ tangent and adjoint of the residual
// AD pool: P-direction / P-gradient 5, metrics 10, T 1, mu 1, node derivatives 12,
// residual tangent 5, adjoint weights 5, per-face viscous flux-input weights 28
enum { AD_P = 0, AD_M = 5, AD_T = 15, AD_MU = 16, AD_DN = 17, AD_R = 29, AD_W = 34, AD_G = 39, AD_N = 67 };

static int ad_alloc(wcns_ctx *ctx)
{
    if (ctx->ad_pool) return WCNS_OK;
    long g3max = 1, wmax = 1;
    for (auto &b : ctx->blocks) g3max = std::max(g3max, (long)b.g3);
    for (auto &w : ctx->walls) wmax = std::max(wmax, (long)w.d.npts);
    TRY(dalloc(ctx, &ctx->ad_pool, (size_t)AD_N * ctx->total));
    TRY(dalloc(ctx, &ctx->ad_work, (size_t)gcl_ad_slots() * g3max + 1));
    TRY(dalloc(ctx, &ctx->ad_force, (size_t)9 * wmax));
    return WCNS_OK;
}

static int ad_check(wcns_ctx *ctx, int nb)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (nb < 0 || nb >= (int)ctx->blocks.size()) return fail(ctx, WCNS_E_ARG, "block index");
    if (inv_variant(ctx->dp))
        return fail(ctx, WCNS_E_UNSUPPORTED, "tangent and adjoint: schemes 4, 41, 42 with fluxes 1-3");
    const HBlock &b = ctx->blocks[nb];
    if (b.ni <= 4 || b.nj <= 4 || b.nk <= 4)
        return fail(ctx, WCNS_E_UNSUPPORTED, "the residual adjoint needs more than 4 nodes per direction");
    return ad_alloc(ctx);
}

// the block's primitives <- P, metrics <- GCL(X) (and their transposed copies)
static int ad_state(wcns_ctx *ctx, int nb, const double *P, const double *X)
{
    ctx->pt_fresh = false;
    const HBlock &b = ctx->blocks[nb];
    DFields F = ctx->fields();
    double *dst[5] = {F.r, F.u, F.v, F.w, F.p};
    for (int m = 0; m < 5; m++)
        CK(hipMemcpyAsync(dst[m] + b.db.off, P + (size_t)m * b.g3, 8 * (size_t)b.g3, hipMemcpyDeviceToDevice,
                          ctx->stream));
    const long nn = (long)b.ni * b.nj * b.nk;
    if (launch_gcl_metrics(b.db, X, X + nn, X + 2 * nn, F, ctx->ad_work, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "metrics launch");
    if (ctx->idir_t) {
        TransList L{};
        for (int m = 0; m < 10; m++) { L.src[L.n] = ctx->metric[m]; L.dst[L.n++] = ctx->mt[m]; }
        if (launch_transpose(0, ctx->grid[GRID_TRANS], L, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "transposing the metrics");
    }
    // the implicit operator's staged copy of the metrics follows them
    for (int prec = 0; prec < 2; prec++)
        if (ctx->skew[prec] && launch_skew_stage(SKEW_METRICS, b.db, ctx->skoff[nb], F, prec, ctx->skew[prec],
                                                 ctx->sktot, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "staging the metrics");
    return WCNS_OK;
}

// the residual of the current state (no turbulence-model update), left in rhs
static int ad_rhs(wcns_ctx *ctx)
{
    int e = wcns_initial_q_dq_c_t(ctx);
    if (e) return e;
    if (ctx->prm.nvis == 1 && (e = vis_meanflow(ctx, ctx->fields()))) return e;
    return wcns_r_h_s_invis(ctx);
}

static int ad_zero(wcns_ctx *ctx, int first, int count)
{
    return launch_fill_u64(reinterpret_cast<unsigned long long *>(ctx->ad_pool + (size_t)first * ctx->total),
                           (long)count * ctx->total, 0ull, ctx->stream)
               ? fail(ctx, WCNS_E_HIP, "clearing the adjoint arrays")
               : WCNS_OK;
}

int wcns_adj_residual(wcns_ctx *ctx, int nb, const double *P, const double *X, double *R)
{
    int e = ad_check(ctx, nb);
    if (e) return e;
    if (!P || !X || !R) return fail(ctx, WCNS_E_ARG, "null array");
    if ((e = ad_state(ctx, nb, P, X)) || (e = ad_rhs(ctx))) return e;
    if (launch_interior(ctx->blocks[nb].db, ctx->rhs, ctx->total, R, 5, 0, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "residual copy");
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_adj_jvp(wcns_ctx *ctx, int nb, const double *P, const double *X, const double *dP, const double *dX,
                 double *R, double *dR)
{
    int e = ad_check(ctx, nb);
    if (e) return e;
    if (!P || !X || !dR) return fail(ctx, WCNS_E_ARG, "null array");
    const HBlock &b = ctx->blocks[nb];
    const long T = ctx->total;
    if ((e = ad_state(ctx, nb, P, X)) || (e = ad_rhs(ctx))) return e;
    if (R && launch_interior(b.db, ctx->rhs, T, R, 5, 0, ctx->stream)) return fail(ctx, WCNS_E_HIP, "residual copy");
    if ((e = ad_zero(ctx, 0, AD_W))) return e;
    double *A = ctx->ad_pool;
    if (dP)
        for (int m = 0; m < 5; m++)
            CK(hipMemcpyAsync(A + (AD_P + m) * T + b.db.off, dP + (size_t)m * b.g3, 8 * (size_t)b.g3,
                              hipMemcpyDeviceToDevice, ctx->stream));
    DFields F = ctx->fields();
    if (dX && launch_gcl_jvp(b.db, F, X, dX, A + AD_M * T, ctx->ad_work, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "metric tangent");
    AdTan D{A + AD_P * T, A + AD_M * T, A + AD_T * T, A + AD_MU * T, A + AD_DN * T, A + AD_R * T};
    if (ctx->prm.nvis == 1) {
        if (launch_tmu_jvp(b.db, F, ctx->dp, A + AD_T * T, A + AD_MU * T, A + AD_P * T, ctx->stream) ||
            launch_dn_jvp(b.db, F, D, A + AD_DN * T, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "viscous tangent");
        for (int dir = 0; dir < 3; dir++)
            if (launch_rhs_vis_jvp(dir, b.db, F, ctx->dp, ctx->dn, D, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "viscous tangent");
    }
    for (int dir = 0; dir < 3; dir++)
        if (launch_rhs_inv_jvp(dir, b.db, F, ctx->dp, D, ctx->stream)) return fail(ctx, WCNS_E_HIP, "inviscid tangent");
    if (launch_interior(b.db, A + AD_R * T, T, dR, 5, 0, ctx->stream)) return fail(ctx, WCNS_E_HIP, "tangent copy");
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_adj_vjp(wcns_ctx *ctx, int nb, const double *P, const double *X, const double *W, double *gP, double *gX)
{
    int e = ad_check(ctx, nb);
    if (e) return e;
    if (!P || !X || !W) return fail(ctx, WCNS_E_ARG, "null array");
    const HBlock &b = ctx->blocks[nb];
    const long T = ctx->total, nn = (long)b.ni * b.nj * b.nk;
    if ((e = ad_state(ctx, nb, P, X)) || (e = ad_rhs(ctx))) return e;
    if ((e = ad_zero(ctx, 0, AD_N))) return e;
    double *A = ctx->ad_pool;
    if (launch_interior(b.db, A + AD_W * T, T, const_cast<double *>(W), 5, 1, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "weight copy");
    DFields F = ctx->fields();
    AdAdj G{A + AD_W * T, A + AD_P * T, A + AD_M * T, A + AD_T * T, A + AD_MU * T, A + AD_DN * T, A + AD_G * T};
    for (int dir = 0; dir < 3; dir++)
        if (launch_rhs_inv_vjp(dir, b.db, F, ctx->dp, G, ctx->stream)) return fail(ctx, WCNS_E_HIP, "inviscid adjoint");
    if (ctx->prm.nvis == 1) {
        for (int dir = 0; dir < 3; dir++)
            if (launch_rhs_vis_vjp(dir, b.db, F, ctx->dp, ctx->dn, G, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "viscous adjoint");
        if (launch_dn_vjp(b.db, F, G, ctx->stream) ||
            launch_tmu_vjp(b.db, F, ctx->dp, A + AD_T * T, A + AD_MU * T, A + AD_P * T, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "viscous adjoint");
    }
    if (gX) {
        CK(hipMemsetAsync(gX, 0, 8 * 3 * (size_t)nn, ctx->stream));
        if (launch_gcl_vjp(b.db, F, X, A + AD_M * T, gX, ctx->ad_work, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "metric adjoint");
    }
    if (gP)
        for (int m = 0; m < 5; m++)
            CK(hipMemcpyAsync(gP + (size_t)m * b.g3, A + (AD_P + m) * T + b.db.off, 8 * (size_t)b.g3,
                              hipMemcpyDeviceToDevice, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_adj_force(wcns_ctx *ctx, int nb, const double *P, const double *X, const double omega[6], double *J,
                   double *gP, double *gX)
{
    int e = ad_check(ctx, nb);
    if (e) return e;
    if (!P || !X || !omega || !J) return fail(ctx, WCNS_E_ARG, "null array");
    if (ctx->prm.sref == 0.0 || ctx->prm.lfref == 0.0) return fail(ctx, WCNS_E_ARG, "sref and lfref must be non-zero");
    const HBlock &b = ctx->blocks[nb];
    const long T = ctx->total, nn = (long)b.ni * b.nj * b.nk;
    if ((e = ad_state(ctx, nb, P, X))) return e;
    DFields F = ctx->fields();
    DParams dpp = ctx->dp;
    dpp.nvis = 0;                      // the pressure part only
    const wcns_params &p = ctx->prm;
    const double ref[3] = {p.xref, p.yref, p.zref};
    // the coefficients' scaling of wcns_force (PostProcess.f90:1415-1477)
    const int wf = 2 - p.nwholefield;
    double om[6];
    for (int c = 0; c < 6; c++) om[c] = omega[c] * wf * (c < 3 ? 1.0 / p.sref : 1.0 / (p.sref * p.lfref));
    double *A = ctx->ad_pool;
    const bool grad = gP || gX;
    if (grad) {
        if ((e = ad_zero(ctx, AD_P, 15))) return e;
        if (gX) CK(hipMemsetAsync(gX, 0, 8 * 3 * (size_t)nn, ctx->stream));
    }
    double *lpn3 = ctx->ad_force, *lxyz = nullptr, *xyz = nullptr;
    double sum = 0.0;
    for (size_t w = 0; w < ctx->walls.size(); w++) {
        const auto &wr = ctx->walls[w];
        if (wr.d.nb != nb) continue;
        const int np = wr.d.npts;
        lxyz = lpn3 + 3 * (size_t)np;
        xyz = lxyz + 3 * (size_t)np;
        if (launch_wall_xyz(wr.d, b.db, const_cast<double *>(X), xyz, 0, ctx->stream) ||
            launch_force_region(wr.d, b.db, F, dpp, xyz, ctx->d_force_pn, ctx->d_force_part, ctx->d_force_out, ref,
                                ctx->stream))
            return fail(ctx, WCNS_E_HIP, "force launch");
        double part[6];
        CK(hipMemcpyAsync(part, ctx->d_force_out, sizeof part, hipMemcpyDeviceToHost, ctx->stream));
        if (grad) {
            if (launch_force_vjp(wr.d, b.db, F, dpp, ctx->d_force_pn, xyz, ref, om, lpn3, lxyz, A + AD_P * T,
                                 A + AD_M * T, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "force adjoint");
            if (gX && launch_wall_xyz(wr.d, b.db, gX, lxyz, 1, ctx->stream)) return fail(ctx, WCNS_E_HIP, "force adjoint");
        }
        CK(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 6; c++) sum = sum + om[c] * part[c];
    }
    *J = sum;
    if (grad) {
        if (gX && launch_gcl_vjp(b.db, F, X, A + AD_M * T, gX, ctx->ad_work, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "metric adjoint");
        if (gP)
            for (int m = 0; m < 5; m++)
                CK(hipMemcpyAsync(gP + (size_t)m * b.g3, A + (AD_P + m) * T + b.db.off, 8 * (size_t)b.g3,
                                  hipMemcpyDeviceToDevice, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
    }
    return WCNS_OK;
}

}  // extern "C"

// ------------------------------------------------ tangent and adjoint of a whole step
// Pool of the step's derivatives (AS_N arrays of total elements): the step's
// primitives after the boundary conditions (PF, 5), the four sweeps' dQ (X,
// 20), an all-zero array, the adjoint weights of the new state (LPN, 6), μ_t
// as the radii read it (VT), and a work region shared by the tangent and the
// adjoint passes.
enum { AS_PF = 0, AS_X = 5, AS_ZERO = 25, AS_LPN = 26, AS_VT = 32, AS_W = 33 };
// tangent work: state after the BCs 6 (r..p, T), residual 5, radii, viscous
// radii and coed 7, sweeps 20, result 6 (r..p, T)
enum { AT_P = 0, AT_R = 6, AT_S = 11, AT_X = 18, AT_PN = 38, AT_N = 44 };
// adjoint work: state after the BCs 6, dQ 5, sweep outputs 20, multipliers
// 20, R 5, coed 1, radii 3, viscous radii 3
enum { AA_P = 0, AA_D = 6, AA_XB = 11, AA_MU = 31, AA_RB = 51, AA_CB = 56, AA_SR = 57, AA_SRV = 60, AA_N = 63 };
enum { AS_N = AS_W + AA_N };

// the BC sequence of boundary_all, region by region
template <typename Fn>
static int for_bc_regions(wcns_ctx *ctx, bool reverse, Fn fn)
{
    std::vector<const DRegion *> seq;
    for (auto &b : ctx->blocks)
        for (int r : b.bcorder) seq.push_back(&b.regs[r].d);
    for (size_t a = 0; a < seq.size(); a++) {
        size_t q = reverse ? seq.size() - 1 - a : a;
        int e = fn(*seq[q], q);
        if (e) return e;
    }
    return WCNS_OK;
}

// Differentiated variables of the state (r, u, v, w, p, and T for a viscous
// case, which the no-slip wall reads one step late) and the whole state the
// step carries (with a turbulence model also nu~ or k, mu_t, and omega).
static int as_nvar(const wcns_ctx *ctx) { return ctx->prm.nvis == 1 ? 6 : 5; }
static std::vector<double *> state_ptrs(wcns_ctx *ctx)
{
    DFields F = ctx->fields();
    std::vector<double *> v = {F.r, F.u, F.v, F.w, F.p};
    if (ctx->prm.nvis == 1) v.push_back(F.t);
    if (ctx->dp.turb) { v.push_back(F.qke); v.push_back(F.vist); }
    if (ctx->dp.turb == 2) v.push_back(F.qkw);
    return v;
}

static int as_check(wcns_ctx *ctx)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (inv_variant(ctx->dp))
        return fail(ctx, WCNS_E_UNSUPPORTED, "tangent and adjoint: schemes 4, 41, 42 with fluxes 1-3");
    if (ctx->lhs_prec != WCNS_LHS_FP64)
        return fail(ctx, WCNS_E_UNSUPPORTED, "the step tangent and adjoint differentiate the FP64 implicit operator");
    for (auto &b : ctx->blocks)
        if (b.ni <= 4 || b.nj <= 4 || b.nk <= 4)
            return fail(ctx, WCNS_E_UNSUPPORTED, "the step adjoint needs more than 4 nodes per direction");
    TRY(ad_alloc(ctx));
    if (ctx->as_pool) return WCNS_OK;
    TRY(dalloc(ctx, &ctx->as_pool, (size_t)AS_N * ctx->total));
    long off = 0;
    ctx->as_tape_off.clear();
    for_bc_regions(ctx, false, [&](const DRegion &g, size_t) {
        ctx->as_tape_off.push_back(off);
        off += 18L * g.npts;
        return 0;
    });
    TRY(dalloc(ctx, &ctx->as_tape, (size_t)off));
    // the multi-point entries grouped by node, in entry order
    std::map<std::tuple<int, int, int, int>, std::vector<int>> by_node;
    const int np = (int)ctx->h_npp_pts.size() / 4;
    for (int e = 0; e < np; e++) {
        const int *p = &ctx->h_npp_pts[4 * e];
        by_node[{p[0], p[1], p[2], p[3]}].push_back(e);
    }
    std::vector<int> us{0}, ue;
    for (auto &kv : by_node) {
        for (int e : kv.second) ue.push_back(e);
        us.push_back((int)ue.size());
    }
    ctx->as_nuniq = (int)by_node.size();
    TRY(dalloc(ctx, &ctx->as_ustart, us.size()));
    TRY(dalloc(ctx, &ctx->as_uent, std::max<size_t>(ue.size(), 1)));
    TRY(copy_sync(ctx, ctx->as_ustart, us.data(), 4 * us.size(), hipMemcpyHostToDevice));
    if (!ue.empty()) TRY(copy_sync(ctx, ctx->as_uent, ue.data(), 4 * ue.size(), hipMemcpyHostToDevice));
    return WCNS_OK;
}

static int as_fill(wcns_ctx *ctx, double *a, long n)
{
    return launch_fill_u64(reinterpret_cast<unsigned long long *>(a), n, 0ull, ctx->stream)
               ? fail(ctx, WCNS_E_HIP, "clearing the step adjoint arrays")
               : WCNS_OK;
}

// n pool arrays from / to n consecutive arrays of a flat buffer
static int as_copy(wcns_ctx *ctx, double *const *dst, const double *src, int n)
{
    for (int m = 0; m < n; m++)
        CK(hipMemcpyAsync(dst[m], src + (size_t)m * ctx->total, 8 * (size_t)ctx->total, hipMemcpyDeviceToDevice,
                          ctx->stream));
    return WCNS_OK;
}

static int as_copy_out(wcns_ctx *ctx, double *dst, double *const *src, int n)
{
    for (int m = 0; m < n; m++)
        CK(hipMemcpyAsync(dst + (size_t)m * ctx->total, src[m], 8 * (size_t)ctx->total, hipMemcpyDeviceToDevice,
                          ctx->stream));
    return WCNS_OK;
}

static int zero_faces(wcns_ctx *ctx, double *arr)
{
    return launch_dq_zero_faces(ctx->d_regs, ctx->d_rwork + ctx->zero_w0, ctx->zero_nw, ctx->d_blocks, arr, nullptr,
                                ctx->total, ctx->stream)
               ? fail(ctx, WCNS_E_HIP, "zeroing the physical faces")
               : WCNS_OK;
}

// the interface averagings in their sequence order on the arrays of F
static int communicate_on(wcns_ctx *ctx, const DFields &F, int pv)
{
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    int nvar = pv ? 6 : 5;
    if (launch_npp_gather(ctx->d_npp_pts, ctx->npp, ctx->d_blocks, F, ctx->d_npp_buf, nvar, !pv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp gather");
    for (auto &g : ctx->iface)
        if (launch_avg2pm(g, dbs.data(), F, pv, ctx->stream)) return fail(ctx, WCNS_E_HIP, "avg2pm");
    if (launch_npp_assign(ctx->d_grp_start, ctx->d_grp_n, ctx->d_npp_pts, ctx->d_write_ok, ctx->ngroups,
                          ctx->d_npp_buf, ctx->d_blocks, F, nvar, !pv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp assign");
    return WCNS_OK;
}

// their transpose on nv arrays lam (dQ with use_vol, else the primitives and T)
static int communicate_t(wcns_ctx *ctx, double *lam, int use_vol, int nv)
{
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    const double *vol = ctx->metric[9];
    if (launch_npp_assign_t(ctx->d_grp_start, ctx->d_grp_n, ctx->d_npp_pts, ctx->d_write_ok, ctx->ngroups,
                            ctx->d_blocks, vol, lam, ctx->total, ctx->d_npp_buf, use_vol, nv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp assign transpose");
    for (size_t a = ctx->iface.size(); a-- > 0;)
        if (launch_avg2pm_t(ctx->iface[a], dbs.data(), vol, lam, ctx->total, use_vol, nv, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "avg2pm transpose");
    if (launch_npp_gather_t(ctx->d_npp_pts, ctx->as_ustart, ctx->as_uent, ctx->as_nuniq, ctx->d_blocks, vol, lam,
                            ctx->total, ctx->d_npp_buf, use_vol, nv, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "npp gather transpose");
    return WCNS_OK;
}

static DFields with_prims(DFields F, double *p)
{
    const long T = F.total;
    F.r = p; F.u = p + T; F.v = p + 2 * T; F.w = p + 3 * T; F.p = p + 4 * T;
    return F;
}

// The residual's tangent (viscous part, averaging, inviscid part, averaging
// again, as step_once forms R) into tr, from the tangent tp of the state
// after the BCs
static int as_rhs_jvp(wcns_ctx *ctx, const DFields &F, const double *tp, double *tr)
{
    const long T = ctx->total;
    double *A = ctx->ad_pool;
    TRY(ad_zero(ctx, AD_M, AD_R - AD_M));
    AdTan D{tp, A + AD_M * T, A + AD_T * T, A + AD_MU * T, A + AD_DN * T, tr};
    if (ctx->prm.nvis == 1) {
        for (auto &b : ctx->blocks) {
            if (launch_tmu_jvp(b.db, F, ctx->dp, A + AD_T * T, A + AD_MU * T, tp, ctx->stream) ||
                launch_dn_jvp(b.db, F, D, A + AD_DN * T, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "viscous tangent");
            for (int dir = 0; dir < 3; dir++)
                if (launch_rhs_vis_jvp(dir, b.db, F, ctx->dp, ctx->dn, D, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "viscous tangent");
        }
        DFields Ft = F;
        Ft.rhs = tr;
        TRY(communicate_on(ctx, Ft, 0));
    }
    for (auto &b : ctx->blocks)
        for (int dir = 0; dir < 3; dir++)
            if (launch_rhs_inv_jvp(dir, b.db, F, ctx->dp, D, ctx->stream)) return fail(ctx, WCNS_E_HIP, "inviscid tangent");
    DFields Ft = F;
    Ft.rhs = tr;
    TRY(communicate_on(ctx, Ft, 0));
    return zero_faces(ctx, tr);
}

// One step of the state in the current buffers (P, if given, is copied there
// first: state_ptrs' arrays), as step_once computes it but with the sweeps'
// outputs kept and the boundary values taped; with dP (as_nvar arrays) its
// tangent too.  Leaves PF, X, R, the radii and coed of the step, and the new
// state current.
static int as_forward(wcns_ctx *ctx, const double *P, const double *dP)
{
    const long T = ctx->total;
    const int nv = as_nvar(ctx);
    const bool vis = ctx->prm.nvis == 1;
    double *S = ctx->as_pool, *W = S + (size_t)AS_W * T;
    DFields F = ctx->fields();
    std::vector<double *> st = state_ptrs(ctx);
    if (P) TRY(as_copy(ctx, st.data(), P, (int)st.size()));
    double *cur[5] = {F.r, F.u, F.v, F.w, F.p};
    const bool tan = dP != nullptr;
    if (tan) {
        TRY(as_fill(ctx, W, (long)AA_N * T));
        double *tp[6];
        for (int m = 0; m < nv; m++) tp[m] = W + (AT_P + m) * T;
        TRY(as_copy(ctx, tp, dP, nv));
    }
    if (ctx->dp.turb) {
        // the turbulence models' boundary values by the solver's own pass, then
        // the primitives as they were for the taped pass below
        TRY(as_copy_out(ctx, S + AS_PF * T, cur, 5));
        TRY(wcns_boundary_all(ctx));
        TRY(as_copy(ctx, cur, S + AS_PF * T, 5));
    }
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    TRY(for_bc_regions(ctx, false, [&](const DRegion &g, size_t q) {
        return launch_bc_fwd(g, dbs.data(), F, ctx->dp, tan ? W + AT_P * T : nullptr, ctx->as_tape + ctx->as_tape_off[q],
                             ctx->stream)
                   ? fail(ctx, WCNS_E_HIP, "boundary tangent launch")
                   : WCNS_OK;
    }));
    TRY(as_copy_out(ctx, S + AS_PF * T, cur, 5));
    if (ctx->dp.turb) CK(hipMemcpyAsync(S + AS_VT * T, F.vist, 8 * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
    TRY(wcns_initial_q_dq_c_t(ctx));
    TRY(wcns_timestep(ctx));
    if (tan && launch_ts_jvp(ctx->grid[GRID_NODES], F, ctx->dp, W + AT_P * T, W + AT_S * T, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "radii tangent");
    if (vis) {
        TRY(wcns_r_h_s_vis(ctx));
        TRY(communicate_on(ctx, F, 0));
    }
    TRY(wcns_r_h_s_invis(ctx));
    TRY(communicate_on(ctx, F, 0));
    TRY(zero_faces(ctx, F.rhs));
    if (tan) TRY(as_rhs_jvp(ctx, F, W + AT_P * T, W + AT_R * T));
    AdSweep A{};
    A.x = S + AS_X * T;
    if (tan) {
        A.tx = W + AT_X * T;
        A.tp = W + AT_P * T;
        A.tsr = W + AT_S * T;
        A.tsrv = W + (AT_S + 3) * T;
        A.tcoed = W + (AT_S + 6) * T;
        A.trhs = W + AT_R * T;
        A.zero = S + AS_ZERO * T;
    }
    if (!tan && ctx->tiled_sweeps) {
        // the solver's pipelined sweeps, each sweep's dQ copied out of the
        // plane-major pair (lusgs_l/u and gs_pr_u leave it in A, gs_pr_l in B)
        void *pool = ctx->skew[0];
        if (launch_skew_stage_all(SKEW_STEP, ctx->grid[GRID_SKEW], ctx->d_skoff, F, 0, pool, ctx->sktot, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "staging the sweep operands");
        for (int kind = 0; kind < 4; kind++)
            if (launch_sweep_batched(kind, ctx->plan, F, ctx->dp, 0, pool, ctx->sktot, ctx->stream) ||
                launch_unskew_f64(ctx->grid[GRID_SKEW], ctx->d_skoff, (const double *)pool, ctx->sktot,
                                  kind == 2 ? SK_B : SK_A, 5, A.x + 5 * kind * T, T, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "sweeps");
    } else {
        for (auto &b : ctx->blocks)
            if (launch_ad_sweeps_fwd(b.db, F, ctx->dp, A, ctx->stream)) return fail(ctx, WCNS_E_HIP, "sweeps");
    }
    // the update's dQ: the last sweep's, zero on the physical faces (x3 itself
    // is kept: the adjoint needs the values the sweep used)
    const double *x3 = S + (AS_X + 15) * T;
    CK(hipMemcpyAsync(F.dq, x3, 8 * 5 * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
    TRY(zero_faces(ctx, F.dq));
    if (tan) TRY(zero_faces(ctx, W + (AT_X + 15) * T));
    if (launch_fill_u64(ctx->d_nflag, (long)ctx->blocks.size(), 0ull, ctx->stream) ||
        launch_update(ctx->grid[GRID_NODES], F, ctx->dp, ctx->d_flag, ctx->d_nflag, ctx->flag_cap, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "update");
    if (tan && launch_upd_jvp(ctx->grid[GRID_NODES], ctx->dp, S + AS_PF * T, F.dq, W + AT_P * T, W + (AT_X + 15) * T,
                              W + AT_PN * T, T, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "update tangent");
    if (tan && vis && launch_tout_jvp(ctx->grid[GRID_NODES], F, ctx->dp, S + AS_PF * T, W + AT_P * T,
                                      W + (AT_PN + 5) * T, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "temperature tangent");
    ctx->cur ^= 1;
    ctx->pt_fresh = false;
    TRY(communicate_on(ctx, ctx->fields(), 1));
    if (tan) {
        DFields Ft = with_prims(ctx->fields(), W + AT_PN * T);
        // the averaged T is part of a viscous state; an inviscid one averages zeros
        Ft.t = vis ? W + (AT_PN + 5) * T : S + AS_ZERO * T;
        TRY(communicate_on(ctx, Ft, 1));
    }
    std::vector<unsigned long long> nf(ctx->blocks.size());
    CK(hipMemcpyAsync(nf.data(), ctx->d_nflag, 8 * nf.size(), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    TRY(check_counters(ctx));
    for (auto n : nf)
        if (n) return fail(ctx, WCNS_E_UNSUPPORTED, "a node left the update limits: the repaired step is not differentiable");
    return WCNS_OK;
}

// The adjoint of the step from state P (current buffers if null) for the
// weights Wn of the new state's as_nvar arrays, into gP (interior nodes).
static int as_reverse(wcns_ctx *ctx, const double *P, const double *Wn, double *gP)
{
    const long T = ctx->total;
    const int nv = as_nvar(ctx);
    const bool vis = ctx->prm.nvis == 1;
    double *S = ctx->as_pool, *W = S + (size_t)AS_W * T;
    TRY(as_forward(ctx, P, nullptr));
    double *lpn = S + AS_LPN * T;
    if (Wn != lpn) CK(hipMemcpyAsync(lpn, Wn, 8 * nv * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
    TRY(as_fill(ctx, W, (long)AA_N * T));
    TRY(communicate_t(ctx, lpn, 0, nv));
    double *lp = W + AA_P * T, *ld = W + AA_D * T;
    if (launch_upd_vjp(ctx->grid[GRID_NODES], ctx->dp, S + AS_PF * T, ctx->dq, lpn, lp, ld, T, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "update adjoint");
    DFields Fs = with_prims(ctx->fields(), S + AS_PF * T);
    if (vis && launch_tout_vjp(ctx->grid[GRID_NODES], Fs, ctx->dp, S + AS_PF * T, lpn + 5 * T, lp, ctx->stream))
        return fail(ctx, WCNS_E_HIP, "temperature adjoint");
    TRY(zero_faces(ctx, ld));
    CK(hipMemcpyAsync(W + (AA_XB + 15) * T, ld, 8 * 5 * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
    AdSweep A{};
    A.x = S + AS_X * T;
    A.mu = W + AA_MU * T;
    A.xb = W + AA_XB * T;
    A.rb = W + AA_RB * T;
    A.cb = W + AA_CB * T;
    A.lp = lp;
    A.lsr = W + AA_SR * T;
    A.lsrv = W + AA_SRV * T;
    for (size_t nb = ctx->blocks.size(); nb-- > 0;)
        if (launch_ad_sweeps_rev(ctx->blocks[nb].db, Fs, ctx->dp, A, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "sweeps adjoint");
    {
        // the radii saw μ_t before the turbulence model's update of the step
        DFields Fv = Fs;
        if (ctx->dp.turb) Fv.vist = S + AS_VT * T;
        if (launch_ts_vjp(ctx->grid[GRID_NODES], Fv, ctx->dp, A.lsr, A.lsrv, A.cb, lp, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "radii adjoint");
    }
    TRY(zero_faces(ctx, A.rb));
    TRY(communicate_t(ctx, A.rb, 1, 5));
    double *AD = ctx->ad_pool;
    TRY(ad_zero(ctx, AD_M, AD_N - AD_M));
    AdAdj G{A.rb, lp, AD + AD_M * T, AD + AD_T * T, AD + AD_MU * T, AD + AD_DN * T, AD + AD_G * T};
    for (auto &b : ctx->blocks)
        for (int dir = 0; dir < 3; dir++)
            if (launch_rhs_inv_vjp(dir, b.db, Fs, ctx->dp, G, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "inviscid adjoint");
    if (vis) {
        // the viscous part was averaged once more: its weights go through the
        // transpose again; T and mu of the state after the BCs, as the
        // residual used them (the primitive averaging has since moved T)
        TRY(communicate_t(ctx, A.rb, 1, 5));
        if (launch_init_stage(ctx->grid[GRID_NODES], &ctx->grid[GRID_GHOST_T], Fs, ctx->dp, ctx->stream))
            return fail(ctx, WCNS_E_HIP, "init launch");
        for (auto &b : ctx->blocks) {
            for (int dir = 0; dir < 3; dir++)
                if (launch_rhs_vis_vjp(dir, b.db, Fs, ctx->dp, ctx->dn, G, ctx->stream))
                    return fail(ctx, WCNS_E_HIP, "viscous adjoint");
            if (launch_dn_vjp(b.db, Fs, G, ctx->stream) ||
                launch_tmu_vjp(b.db, Fs, ctx->dp, AD + AD_T * T, AD + AD_MU * T, lp, ctx->stream))
                return fail(ctx, WCNS_E_HIP, "viscous adjoint");
        }
    }
    std::vector<DBlock> dbs;
    for (auto &b : ctx->blocks) dbs.push_back(b.db);
    TRY(for_bc_regions(ctx, true, [&](const DRegion &g, size_t q) {
        return launch_bc_rev(g, dbs.data(), Fs, ctx->dp, lp, ctx->as_tape + ctx->as_tape_off[q], ctx->stream)
                   ? fail(ctx, WCNS_E_HIP, "boundary adjoint launch")
                   : WCNS_OK;
    }));
    for (auto &b : ctx->blocks)
        if (launch_zero_ghosts(b.db, lp, T, nv, ctx->stream)) return fail(ctx, WCNS_E_HIP, "ghost clear");
    CK(hipMemcpyAsync(gP, lp, 8 * nv * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
    return WCNS_OK;
}

// J = sum_c omega_c C_c (pressure part of the wall force, scaled as
// wcns_force) of the current state, and dJ/dP into g (as_nvar pool arrays)
static int as_force(wcns_ctx *ctx, const double omega[6], double *J, double *g)
{
    if (ctx->prm.sref == 0.0 || ctx->prm.lfref == 0.0) return fail(ctx, WCNS_E_ARG, "sref and lfref must be non-zero");
    const long T = ctx->total;
    DFields F = ctx->fields();
    DParams dpp = ctx->dp;
    dpp.nvis = 0;
    const wcns_params &p = ctx->prm;
    const double ref[3] = {p.xref, p.yref, p.zref};
    const int wf = 2 - p.nwholefield;
    double om[6];
    for (int c = 0; c < 6; c++) om[c] = omega[c] * wf * (c < 3 ? 1.0 / p.sref : 1.0 / (p.sref * p.lfref));
    TRY(as_fill(ctx, g, (long)as_nvar(ctx) * T));
    TRY(ad_zero(ctx, AD_M, AD_R - AD_M));
    double *AD = ctx->ad_pool;
    double sum = 0.0;
    for (auto &wr : ctx->walls) {
        const DBlock &db = ctx->blocks[wr.d.nb].db;
        const int np = wr.d.npts;
        double *lpn3 = ctx->ad_force, *lxyz = lpn3 + 3 * (size_t)np;
        if (launch_force_region(wr.d, db, F, dpp, wr.xyz, ctx->d_force_pn, ctx->d_force_part, ctx->d_force_out, ref,
                                ctx->stream) ||
            launch_force_vjp(wr.d, db, F, dpp, ctx->d_force_pn, wr.xyz, ref, om, lpn3, lxyz, g, AD + AD_M * T,
                             ctx->stream))
            return fail(ctx, WCNS_E_HIP, "force adjoint");
        double part[6];
        CK(hipMemcpyAsync(part, ctx->d_force_out, sizeof part, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 6; c++) sum = sum + om[c] * part[c];
    }
    *J = sum;
    return WCNS_OK;
}

extern "C" {

int wcns_pool_layout(wcns_ctx *ctx, long *total, long *offsets, int *nstate, int *nvar)
{
    if (!ctx || !ctx->setup_done) return WCNS_E_STATE;
    if (total) *total = ctx->total;
    if (offsets)
        for (size_t nb = 0; nb < ctx->blocks.size(); nb++) offsets[nb] = ctx->blocks[nb].db.off;
    if (nstate) *nstate = (int)state_ptrs(ctx).size();
    if (nvar) *nvar = as_nvar(ctx);
    return WCNS_OK;
}

int wcns_adj_step_jvp(wcns_ctx *ctx, const double *P, const double *dP, double *Pn, double *dPn)
{
    TRY(as_check(ctx));
    if (!dP) return fail(ctx, WCNS_E_ARG, "null direction");
    TRY(as_forward(ctx, P, dP));
    const long T = ctx->total;
    std::vector<double *> st = state_ptrs(ctx);
    if (Pn) TRY(as_copy_out(ctx, Pn, st.data(), (int)st.size()));
    if (dPn) {
        const int nv = as_nvar(ctx);
        CK(hipMemcpyAsync(dPn, ctx->as_pool + (size_t)(AS_W + AT_PN) * T, 8 * nv * (size_t)T, hipMemcpyDeviceToDevice,
                          ctx->stream));
        for (auto &b : ctx->blocks)
            if (launch_zero_ghosts(b.db, dPn, T, nv, ctx->stream)) return fail(ctx, WCNS_E_HIP, "ghost clear");
    }
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_adj_step_vjp(wcns_ctx *ctx, const double *P, const double *Wn, double *gP)
{
    TRY(as_check(ctx));
    if (!Wn || !gP) return fail(ctx, WCNS_E_ARG, "null array");
    TRY(as_reverse(ctx, P, Wn, gP));
    CK(hipStreamSynchronize(ctx->stream));
    return WCNS_OK;
}

int wcns_adj_multistep(wcns_ctx *ctx, const double *P0, int nsteps, int every, const double omega[6],
                       double budget_gb, double *J, double *gP0, double *used_gb)
{
    TRY(as_check(ctx));
    if (!P0 || !omega || !J || nsteps < 1 || every < 1) return fail(ctx, WCNS_E_ARG, "bad arguments");
    const long T = ctx->total;
    const int ns = (int)state_ptrs(ctx).size(), nv = as_nvar(ctx);
    const int nck = (nsteps + every - 1) / every, seg = std::min(every, nsteps);
    const size_t state = (size_t)ns * T;
    const double store = 8.0 * (state * (nck + seg) + (size_t)nv * T);
    const double work = 8.0 * (double)T * (AS_N + AD_N);
    if (used_gb) *used_gb = (store + work) / 1e9;
    if (budget_gb > 0.0 && store + work > budget_gb * 1e9)
        return fail(ctx, WCNS_E_ARG, "checkpoints and work arrays need " + std::to_string((store + work) / 1e9) +
                                         " GB, above the budget");
    double *ck = nullptr;
    TRY(dalloc(ctx, &ck, state * (nck + seg) + (size_t)nv * T));
    double *sg = ck + state * nck, *lam = sg + state * seg;
    auto run = [&]() -> int {
        std::vector<double *> st = state_ptrs(ctx);
        TRY(as_copy(ctx, st.data(), P0, ns));
        // the states by the solver's own step (the derivative's primal equals it bit for bit)
        for (int n = 0; n < nsteps; n++) {
            if (n % every == 0) {
                st = state_ptrs(ctx);
                TRY(as_copy_out(ctx, ck + state * (n / every), st.data(), ns));
            }
            TRY(step_once(ctx));
        }
        TRY(check_counters(ctx));
        TRY(as_force(ctx, omega, J, lam));
        // each segment: its states replayed from the checkpoint, then reversed
        for (int c = nck - 1; c >= 0; c--) {
            const int s0 = c * every, s1 = std::min(nsteps, s0 + every);
            CK(hipMemcpyAsync(sg, ck + state * c, 8 * state, hipMemcpyDeviceToDevice, ctx->stream));
            st = state_ptrs(ctx);
            TRY(as_copy(ctx, st.data(), sg, ns));
            for (int n = s0; n + 1 < s1; n++) {
                TRY(step_once(ctx));
                st = state_ptrs(ctx);
                TRY(as_copy_out(ctx, sg + state * (n + 1 - s0), st.data(), ns));
            }
            for (int n = s1 - 1; n >= s0; n--) TRY(as_reverse(ctx, sg + state * (n - s0), lam, lam));
        }
        CK(hipMemcpyAsync(gP0, lam, 8 * nv * (size_t)T, hipMemcpyDeviceToDevice, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        return WCNS_OK;
    };
    int e = run();
    hipStreamSynchronize(ctx->stream);
    hipFree(ck);
    return e;
}

}  // extern "C"
