// wcns_kernels.hpp -- launchers of the gfx950 kernels (one translation unit
// per kernel family); all take plain structs by value and a stream.
#pragma once
#include "wcns_device.hpp"

namespace wcns {

// Boundary region as the kernels see it.  Face box s_lo..s_hi (1-based,
// inclusive), outward direction l3, and for interfaces the affine map
// image = t0 + M*(idx - s0) of src/parallel.f90:1185-1301.
struct DRegion {
    int bctype, nb, nbt, s_nd;
    int s_lo[3], s_hi[3], l3[3], t_l3[3];
    int s0[3], t0[3], M[9];
    int s_lr;
    int npts;                 // face points
    int serial;               // 1: reads overlap writes inside this region -> one thread
    double nrm[3];            // symmetry plane normal (BC 3, BC.f90:381-396)
};

__host__ __device__ inline void region_point(const DRegion &g, int t, int &i, int &j, int &k)
{
    int bi = g.s_hi[0] - g.s_lo[0] + 1, bj = g.s_hi[1] - g.s_lo[1] + 1;
    i = g.s_lo[0] + t % bi;
    j = g.s_lo[1] + (t / bi) % bj;
    k = g.s_lo[2] + t / (bi * bj);
}

__host__ __device__ inline void region_image(const DRegion &g, int i, int j, int k, int &it, int &jt, int &kt)
{
    int d0 = i - g.s0[0], d1 = j - g.s0[1], d2 = k - g.s0[2];
    it = g.t0[0] + g.M[0] * d0 + g.M[1] * d1 + g.M[2] * d2;
    jt = g.t0[1] + g.M[3] * d0 + g.M[4] * d1 + g.M[5] * d2;
    kt = g.t0[2] + g.M[6] * d0 + g.M[7] * d1 + g.M[8] * d2;
}

// All blocks in one launch: workgroup w does the work that workgroup wg[w].y
// of a launch over block wg[w].x alone would do.  One map per kind of launch
// (the number of threads per block differs); blocks with no work have no entry.
struct BlockGrid {
    const DBlock *blocks;     // device copy of every block
    const int2 *wg;           // per workgroup: (block, workgroup within the block)
    int nwg;
};
__device__ inline void grid_thread(const BlockGrid &G, DBlock &b, long &t)
{
    const int2 w = G.wg[blockIdx.x];
    b = G.blocks[w.x];
    t = (long)w.y * blockDim.x + threadIdx.x;
}
constexpr int PENCIL_SEG = 32;    // rows per thread of the pencil kernels
enum {
    GRID_NODES = 0,               // nodes 1..n, 256 threads per workgroup
    GRID_GHOST_T = 1,             // + dir: ghost temperatures of the i, j, k faces
    GRID_RHS_VIS = 4,             // + dir: viscous pencil segments
    GRID_RHS_INV = 7,             // + dir: inviscid pencil segments
    GRID_SA_RHS = 10,             // SA nodes 2..n-1
    GRID_SA_UPD = 11,             // SA nodes 0..n
    GRID_TRANS = 12,              // 32x32 tiles of every k plane (i<->j transposes)
    GRID_RHS_INV_T = 13,          // inviscid i pencils as j pencils of the transposed view
    GRID_RHS_VIS_T = 14,          // viscous i pencils, likewise
    GRID_VFACE = 15,              // + d (0: i on the transposed view, 1: j, 2: k): viscous half nodes
    GRID_VROW = 18,               // + d: viscous FLUX_DXYZ rows
    GRID_SKEW = 21,               // plane-major staging tiles
    GRID_RESID = 22,              // residual monitor: residual_parts(b) workgroups per block, in block order
    GRID_N = 23
};
long grid_threads(int kind, const DBlock &b);
int grid_block_size(int kind);

int launch_rhs_inv_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s);
// schemes 45 / 47 or flux 4: the walk of the direct pencils with ghost
// metrics / pressure ratios
bool inv_variant(const DParams &P);
int launch_pressure_ratios(const DBlock &b, const DFields &F, hipStream_t s);
int launch_ghost_metrics(const DRegion &g, const DBlock &b, const DBlock &bt, const DFields &F, double *gm, long base,
                         hipStream_t s);
// The pipelined sweeps of all blocks run in one persistent launch per sweep:
// every block's (j,k) columns are cut into TJ x TK tiles, one workgroup per
// tile, and the workgroups follow a work list that is a topological order of
// every block's tile graph.
struct SweepBlock {
    DBlock b;
    long soff;                    // offset of the block in the plane-major arrays
    int TJ, TK, nTJ, nTK;
    double *xj, *xk;              // exchange slots: [tile][TK][ni][5], [tile][TJ][ni][5]
};
struct SweepPlan {
    const SweepBlock *blocks;     // device copy, one per block
    const int2 *work_fwd, *work_bwd;  // (block, tile) per workgroup, forward / backward sweeps
    int nwork, lanes;             // workgroups; lanes per workgroup (largest TJ*TK)
    unsigned long long *xch;      // all exchange slots
    long nxch;
    int *err;                     // set when a bounded wait expires
    long spin_limit;
    int sgs_paired;               // 1: SGS sweeps with three helper wavefronts of two parts each
};
long sweep_exchange_words(int TJ, int TK, int nTJ, int nTK, int ni);
// An exchange word nobody has written yet: a signalling-NaN pattern that no
// arithmetic result has.  Every slot is reset to it before each sweep.
constexpr unsigned long long XCH_EMPTY = 0x7ff4deadbeef0001ull;
// Exchange words seen as global memory.  The slot pointers come from structs
// in memory, whose address space the compiler cannot infer: through them the
// loads and stores would be flat instructions, which also count in the LDS
// counter that every workgroup barrier drains, so each step would wait for
// its own exchange traffic to reach memory.
typedef __attribute__((address_space(1))) unsigned long long xch_word;
__device__ inline xch_word *xch_ptr(double *slot) { return (xch_word *)(slot); }
int launch_fill_xch(const SweepPlan &S, hipStream_t s);
// fills n words with v by a kernel (no memset node, so it may sit in a captured graph)
int launch_fill_u64(unsigned long long *a, long n, unsigned long long v, hipStream_t s);
// Plane-major staging copy of the sweep operands: arrays of skew_size(b)
// elements per block (element (i,j,k), i,j,k in 0..n+1, at (i+j+k)*(nj+2)*(nk+2)
// + j + (nj+2)*k), SK_N arrays of sktot elements each, double or float (prec).
enum { SK_P = 0, SK_M = 5, SK_SR = 14, SK_SRV = 17, SK_RHS = 20, SK_COED = 25, SK_A = 26, SK_B = 31, SK_N = 36 };
enum { SKEW_STEP = 0, SKEW_METRICS = 1 };
long skew_size(const DBlock &b);
// Plane-major index (s = i+j+k, j, k) of nodes 0..n+1.  Row j sits at j +
// SK_JSHIFT of a row pitch that is a multiple of 16, so that the first row of
// every tile (j = 1 + multiple of 16, or of 8, 4) starts on a 128-byte line,
// and a 16-row tile reads one line per k instead of two.
constexpr int SK_JSHIFT = 15;
struct SkIndex {
    long off;
    int pl, nj2;
    __device__ long operator()(int i, int j, int k) const
    {
        return off + (long)(i + j + k) * pl + (j + SK_JSHIFT) + (long)nj2 * k;
    }
};
__host__ __device__ inline int skew_pitch(const DBlock &b) { return (b.nj + 2 + SK_JSHIFT + 15) / 16 * 16; }
__host__ __device__ inline SkIndex skew_index(const DBlock &b, long soff)
{
    return SkIndex{soff, skew_pitch(b) * (b.nk + 2), skew_pitch(b)};
}
long skew_stage_wgs(const DBlock &b);
// n FP64 field-pool arrays src[a] into plane-major arrays slot[a] of every
// block (GRID_SKEW map), and ncomp plane-major arrays from slot0 on back into
// the field pool (interior nodes, components total apart)
int launch_skew_stage_f64(const BlockGrid &G, const long *skoff, const double *const *src, const int *slot, int n,
                          double *pool, long sktot, hipStream_t s);
int launch_unskew_f64(const BlockGrid &G, const long *skoff, const double *pool, long sktot, int slot0, int ncomp,
                      double *dst, long total, hipStream_t s);
// the pipelined sweeps' result (plane-major dQ) into the field pool's dq
// (prec 0) or dq32 (prec 1), interior nodes of every block, over GRID_SKEW
int launch_unskew_dq(const BlockGrid &G, const long *skoff, int prec, void *pool, long sktot, double *dq, float *dq32,
                     long total, hipStream_t s);
// the staging of every block in one launch over G = the GRID_SKEW map;
// skoff: each block's offset in the plane-major arrays (device)
int launch_skew_stage_all(int what, const BlockGrid &G, const long *skoff, const DFields &F, int prec, void *pool,
                          long sktot, hipStream_t s);
int launch_skew_stage(int what, const DBlock &b, long soff, const DFields &F, int prec, void *pool, long sktot,
                      hipStream_t s);
int launch_sweep_batched(int kind, const SweepPlan &S, const DFields &F, const DParams &P, int prec, void *pool,
                         long sktot, hipStream_t s);
int launch_sweep(int kind, const DBlock &b, const DFields &F, const DParams &P, hipStream_t s);

// Tangent and adjoint of one step's four sweeps on one block (pool-layout
// arrays of F.total doubles; a group of 5 is component-major).  The forward
// writes the sweeps' dQ x (20: sweep q, component m at (5q+m)*total) and,
// with tx set, their tangents from those of the primitives (tp, 5), radii
// (tsr, 3), viscous radii (tsrv, 3, or null), coed (tcoed) and residual
// (trhs, 5); zero is an all-zero array.  The reverse takes the adjoint of x3
// in xb[15..19] and leaves the multipliers (mu, 20), the adjoints of x0..x2
// (xb), and accumulates those of R (rb, 5), coed (cb), the primitives (lp, 5),
// the radii (lsr, 3) and the viscous radii (lsrv, 3).
struct AdSweep {
    double *x, *tx;
    const double *tp, *tsr, *tsrv, *tcoed, *trhs, *zero;
    double *mu, *xb, *rb, *cb, *lp, *lsr, *lsrv;
};
int launch_ad_sweeps_fwd(const DBlock &b, const DFields &F, const DParams &P, const AdSweep &A, hipStream_t s);
int launch_ad_sweeps_rev(const DBlock &b, const DFields &F, const DParams &P, const AdSweep &A, hipStream_t s);

// Tangent and adjoint of the rest of an (inviscid) step, wcns_adstep.hip.
// Boundary regions one at a time, in the order of the solver's sequence: the
// forward sets the ghosts and face nodes as bc_point does, in D1 when tan is
// set (6 arrays: r, u, v, w, p, T; its tangent follows), and records the
// values each point reads (18 doubles per point) on the tape; the reverse
// moves the adjoint of every value a point wrote to the values it read.
int launch_bc_fwd(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, double *tan,
                  double *tape, hipStream_t s);
int launch_bc_rev(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, double *lam,
                  const double *tape, hipStream_t s);
// spectral radii, viscous radii and coed (k_timestep): tangent
// ts (7: sra, srb, src, srva, srvb, srvc, coed) from tp; adjoint lp += from
// ls (3), lv (3) and lc
int launch_ts_jvp(const BlockGrid &G, const DFields &F, const DParams &P, const double *tp, double *ts, hipStream_t s);
int launch_ts_vjp(const BlockGrid &G, const DFields &F, const DParams &P, const double *ls, const double *lv,
                  const double *lc, double *lp, hipStream_t s);
// the temperature initial_q_dq_c_t forms from primitives pf: tangent tt from
// tp; adjoint lp += from lt
int launch_tout_jvp(const BlockGrid &G, const DFields &F, const DParams &P, const double *pf, const double *tp,
                    double *tt, hipStream_t s);
int launch_tout_vjp(const BlockGrid &G, const DFields &F, const DParams &P, const double *pf, const double *lt,
                    double *lp, hipStream_t s);
// limited update (k_update) of primitives pf by dq: tangent tpn; adjoint
// lp += and ld = from lpn
int launch_upd_jvp(const BlockGrid &G, const DParams &P, const double *pf, const double *dq, const double *tpf,
                   const double *tdq, double *tpn, long total, hipStream_t s);
int launch_upd_vjp(const BlockGrid &G, const DParams &P, const double *pf, const double *dq, const double *lpn,
                   double *lp, double *ld, long total, hipStream_t s);
// transposed interface averagings (k_npp_gather / k_avg2pm / k_npp_assign),
// on nv (5 or 6) arrays lam; use_vol as for dQ; bufb holds 6 per entry
int launch_avg2pm_t(const DRegion &g, const DBlock *blocks, const double *vol, double *lam, long total, int use_vol,
                    int nv, hipStream_t s);
int launch_npp_assign_t(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok, int ngroups,
                        const DBlock *blocks, const double *vol, double *lam, long total, double *bufb, int use_vol,
                        int nv, hipStream_t s);
// (entries of the same node, which may sit in several groups, summed per node
// in order: node u owns ent[ustart[u] .. ustart[u+1]-1])
int launch_npp_gather_t(const int *pts, const int *ustart, const int *ent, int nuniq, const DBlock *blocks,
                        const double *vol, double *lam, long total, const double *bufb, int use_vol, int nv,
                        hipStream_t s);
// ghost nodes of nv arrays to zero
int launch_zero_ghosts(const DBlock &b, double *arr, long total, int nv, hipStream_t s);

int launch_vis_node_deriv(const BlockGrid &G, const DFields &F, double *dn, hipStream_t s);
int launch_rhs_vis_dir(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *dn, hipStream_t s);
// the viscous residual of one direction in two passes (dir 1: j, 2: k, 3: i on the
// transposed view); fvb: 5 pool arrays of half-node fluxes
int launch_vis_faces(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *dn, double *fvb,
                     hipStream_t s);
int launch_vis_rows(int dir, const BlockGrid &G, const DFields &F, const DParams &P, const double *fvb, hipStream_t s);

int launch_bc_region(const DRegion &g, const DBlock *blocks, const DFields &F, const DParams &P, hipStream_t s);
int launch_init_stage(const BlockGrid &nodes, const BlockGrid *ghost_t, const DFields &F, const DParams &P,
                      hipStream_t s);
int launch_timestep(const BlockGrid &G, const DFields &F, const DParams &P, hipStream_t s);
int launch_zero_rhs(const DBlock &b, const DFields &F, hipStream_t s);
int launch_npp_gather(const int *pts, int npts, const DBlock *dblocks, const DFields &F, double *buf,
                      int nvar, int use_vol, hipStream_t s);
int launch_avg2pm(const DRegion &g, const DBlock *blocks, const DFields &F, int pv, hipStream_t s);
int launch_npp_assign(const int *grp_start, const int *grp_n, const int *pts, const int *write_ok, int ngroups,
                      const double *buf, const DBlock *dblocks, const DFields &F, int nvar, int use_vol,
                      hipStream_t s);
// the limited update of every block, then the ordered repair of each block's
// flagged nodes (flag_list holds cap entries per block, nflag one count each)
int launch_update(const BlockGrid &nodes, const DFields &F, const DParams &P, int *flag_list,
                  unsigned long long *nflag, int cap, hipStream_t s);
int launch_repair(const DBlock *blocks, int nblocks, const DFields &F, const DParams &P, int *flag_list,
                  const unsigned long long *nflag, int cap, hipStream_t s);
// One dependency level of the boundary conditions / of the interface
// averaging, and the dQ = 0 of every physical face: work[w] = (region, first
// point), or (region, -1) for a region run in order on one thread; REGION_CHUNK
// points per workgroup.
constexpr int REGION_CHUNK = 128;
int launch_bc_level(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, const DFields &F,
                    const DParams &P, hipStream_t s);
int launch_avg2pm_level(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, const DFields &F,
                        int pv, hipStream_t s);
int launch_dq_zero_faces(const DRegion *regs, const int2 *work, int nwork, const DBlock *blocks, double *arr,
                         float *arr32, long total, hipStream_t s);
int residual_parts(const DBlock &b);
// residual_curve (PostProcess.f90:1195-1293) of every block over G = the
// GRID_RESID list (wg0: first workgroup of each block, nblocks + 1 entries):
// partial (3 per workgroup) -> resblk (sum, max, argmax per block) -> out[8] =
// rms, max, i, j, k, m, nb, entry number; with hist != nullptr the entry is also
// appended to the ring hist[cap][8] at *count % cap and *count is advanced.
int launch_residual_all(const BlockGrid &G, const DFields &F, double *partial, const int *wg0, int nblocks,
                        double *resblk, const DBlock *blocks, double *out, double *hist, unsigned long long *count,
                        int cap, hipStream_t s);
int launch_copy_interior(const DBlock &b, const double *src, double *dst, hipStream_t s);

// Turbulence model scratch of one step (field-pool frame, total elements each);
// the second variable's arrays (SST omega) are null for SA.
struct SaWork {
    double *grad;          // 12 arrays: d(u, v, w, nu~ or k)/d(x, y, z), variable-major
    double *gradw;         // 3 arrays: d(omega)/d(x, y, z)
    double *viseq[2];      // equivalent viscosity per variable
    double *un[3];         // contravariant velocities along i, j, k
    double *dqke[2];       // right-hand side / sweep unknown per variable
    double *spec[2];       // diagonal per variable
    double *f1;            // SST blending function F1
};
int launch_wall_dist(const DBlock &b, const double *x, const double *y, const double *z, const double *wall, int nw,
                     double *dist, hipStream_t s);
int launch_sa_init(const DBlock &b, const DFields &F, double q0, double w0, double vt, hipStream_t s);
int launch_sa_rhs(const BlockGrid &nodes, const BlockGrid &inner, const DFields &F, const SaWork &W,
                  const DParams &P, hipStream_t s);
// pm: the FP64 plane-major pool when the operands were staged there (slots
// SK_A .. SK_A+4: the unknown, the diagonal, un[0..2]), or null for the field pools
int launch_sa_sweep(int fwd, int var, const SweepPlan &S, const SaWork &W, double *pm, long sktot, hipStream_t s);
int launch_sa_update(const BlockGrid &G, const DFields &F, const SaWork &W, const DParams &P, hipStream_t s);

// i<->j transposed copies (wcns_transpose.hip): block b's copy has b's offset
// and plane size, node (i,j,k) at off + (j+2) + (nj+6)(i+2) + sy(k+2); the
// transposed view is the DBlock that addresses it with (j, i, k).
constexpr int TRANS_MAX = 16;
struct TransList {
    const double *src[TRANS_MAX];
    double *dst[TRANS_MAX];
    int n;
};
long transpose_tiles(const DBlock &b);
DBlock transposed_view(const DBlock &b);
// back = 0: dst (transposed) = src; back = 1: dst (normal) += src (transposed)
int launch_transpose(int back, const BlockGrid &G, const TransList &L, hipStream_t s);

// ------------------------------------------- tangent and adjoint of the residual
// Work arrays of one block in the field-pool frame (total elements per array).
// Metric arrays are in pool order kcx, kcy, kcz, etx, ety, etz, ctx, cty, ctz, vol.
struct AdTan {
    const double *dp;        // 5: direction of r, u, v, w, p
    const double *dm;        // 10: direction of the metrics and volume
    const double *dt;        // direction of T
    const double *dvisl;     // direction of mu
    const double *ddn;       // 12: direction of the node derivatives
    double *dr;              // 5: tangent of the residual (accumulated)
};
struct AdAdj {
    const double *w;         // 5: adjoint weights of the residual
    double *gp;              // 5: gradient w.r.t. r, u, v, w, p (accumulated)
    double *gm;              // 10: gradient w.r.t. the metrics and volume
    double *gt;              // gradient w.r.t. T
    double *gvisl;           // gradient w.r.t. mu
    double *gdn;             // 12: gradient w.r.t. the node derivatives
    double *gf;              // 28: per face, the viscous flux weight through the flux's inputs
};
int launch_rhs_inv_jvp(int dir, const DBlock &b, const DFields &F, const DParams &P, const AdTan &D, hipStream_t s);
int launch_rhs_inv_vjp(int dir, const DBlock &b, const DFields &F, const DParams &P, const AdAdj &A, hipStream_t s);
int launch_rhs_vis_jvp(int dir, const DBlock &b, const DFields &F, const DParams &P, const double *dn, const AdTan &D,
                       hipStream_t s);
int launch_rhs_vis_vjp(int dir, const DBlock &b, const DFields &F, const DParams &P, const double *dn, const AdAdj &A,
                       hipStream_t s);
int launch_dn_jvp(const DBlock &b, const DFields &F, const AdTan &D, double *ddn, hipStream_t s);
int launch_dn_vjp(const DBlock &b, const DFields &F, const AdAdj &A, hipStream_t s);
// T and mu of every node from p and rho: tangent (dt, dvisl from dp) and adjoint
// (gt, gvisl into gp)
int launch_tmu_jvp(const DBlock &b, const DFields &F, const DParams &P, double *dt, double *dvisl, const double *dp,
                   hipStream_t s);
int launch_tmu_vjp(const DBlock &b, const DFields &F, const DParams &P, const double *gt, const double *gvisl,
                   double *gp, hipStream_t s);
// compact arrays (nv of ni*nj*nk, i fastest) <-> pool arrays at the interior nodes
int launch_interior(const DBlock &b, double *pool, long total, double *compact, int nv, int to_pool, hipStream_t s);
// conservative metrics: tangent dM = (dM/dX) dX and adjoint gX += (dM/dX)^T gM;
// work: gcl_ad_slots() G3-shaped arrays; X, dX, gX in PLOT3D order (3 arrays)
int gcl_ad_slots();
// (F: the primal metrics of the block, already formed from X)
int launch_gcl_jvp(const DBlock &b, const DFields &F, const double *X, const double *dX, double *dm, double *work,
                   hipStream_t s);
int launch_gcl_vjp(const DBlock &b, const DFields &F, const double *X, const double *gm, double *gX, double *work,
                   hipStream_t s);

// wall force
long force_quads(const DRegion &g);
int force_parts(const DRegion &g);
// adjoint of the pressure force (k_force_pn run with nvis = 0 leaves pn3)
int launch_wall_xyz(const DRegion &g, const DBlock &b, double *X, double *xyz, int back, hipStream_t s);
// the force coefficients from the regions' sums (6 per region, in order):
// wf = 2 - nwholefield, scf = 1/sref, vcm = 1/(sref lfref), and the sines and
// cosines of the angle of attack and sideslip
struct ForceCoef { double wf, scf, vcm, sina, cosa, sinb, cosb; };
int launch_force_coef(const double *out6, int nwalls, const ForceCoef &c, double *out, double *hist,
                      unsigned long long *count, int cap, hipStream_t s);
// tecout_bin_wall's surface values of one wall region (six arrays of g.npts:
// cp, fnx, fny, fnz, cf, qw; cp only without viscosity)
int launch_wall_plot(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, double cpk, double *out,
                     hipStream_t s);
int launch_force_vjp(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, const double *pn3,
                     const double *xyz, const double ref[3], const double omega[6], double *lpn3, double *lxyz,
                     double *gp, double *gm, hipStream_t s);
int launch_force_region(const DRegion &g, const DBlock &b, const DFields &F, const DParams &P, const double *xyz,
                        double *pn3, double *partial, double *out6, const double ref[3], hipStream_t s);

// setup
int launch_gcl_metrics(const DBlock &b, const double *x, const double *y, const double *z, DFields F,
                       double *work, hipStream_t s);

}  // namespace wcns
