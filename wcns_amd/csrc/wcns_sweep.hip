// wcns_sweep.hip -- the four implicit sweeps of one pseudo-time step:
// lusgs_l (LHS.f90:381-478), lusgs_u (:480-577), gs_pr_l (:1777-1866) and
// gs_pr_u (:1956-2046).
//
// Node (i,j,k) of a sweep combines its three neighbours updated earlier in
// the same sweep (i-1,j-1,k-1 going forward, i+1,j+1,k+1 going backward) with,
// in the SGS sweeps, the three on the other side as the previous sweep left
// them.  Processing the hyperplanes i+j+k = s in sweep order gives every node
// exactly the operands of the reference's lexicographic loops, so both
// schedules below compute the reference's arithmetic node for node:
//
//   k_sweep_plane  one launch per hyperplane, dQ updated in place;
//   k_sweep_pipe   one persistent launch per sweep.  A workgroup owns a
//                  TJ x TK tile of (j,k) columns, one column per lane, and
//                  walks the hyperplanes of its tile; at step t lane (jl,kl)
//                  updates node i = t+1-jl-kl (forward).  Neighbours inside the
//                  tile come from the lane's own previous result (i) or from
//                  LDS (j, k); neighbours in the adjacent tile come through
//                  exchange slots, one per boundary node, initialised to a
//                  signalling-NaN pattern that no arithmetic result has and
//                  written with one relaxed device-scope atomic store per word:
//                  every word is its own ready flag, so no fence or progress
//                  counter is needed.  The SGS pair keeps the previous sweep's
//                  dQ in a second buffer (gs_pr_l writes B, reading A; gs_pr_u
//                  writes A, reading B), so no tile ever overwrites a value
//                  another tile has still to read.  Everything a node needs
//                  except the neighbours' new dQ -- primitives, metrics,
//                  spectral radii, the split-Jacobian coefficients, the
//                  complete terms of the previous-sweep neighbours -- is formed
//                  one step ahead, off the dependency chain.  The operands are
//                  read from a plane-major copy (index (i+j+k, j, k), one ghost
//                  layer) staged once per step by k_skew_stage, so the lanes of
//                  a tile read a few contiguous rows per array at every step
//                  instead of one cache line per lane.
#include <type_traits>
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// Five component arrays of a dQ field, passed by value (indexing a kernel
// argument through a pointer would copy it to scratch).
template <typename T>
struct TPtrs {
    T *c[5];
};

// Sub1.f90:676-717 getrkec_mml clamp + LHS.f90:647-678 getrarbrc clamp
__device__ inline int clampi(int v, int n) { return v < 1 ? 1 : (v > n ? n : v); }

// Where a sweep reads its operands: base pointers of the primitive, metric,
// spectral-radius, residual and diagonal arrays (element type S) and an index
// map IX from (i,j,k) to an element offset.  The per-hyperplane schedule reads
// the field pools directly (S = double, G3 index); the pipelined one reads the
// plane-major staging copy (S = T).
template <typename S>
struct Operands {
    const S *p[5];
    const S *m[3][3];     // [direction][x, y, z]: kcx kcy kcz / etx ety etz / ctx cty ctz
    const S *sr[3];       // sra, srb, src
    const S *srv[3];      // srva, srvb, srvc
    const S *rhs[5];
    const S *coed;
};

struct G3Index {
    DBlock b;
    __device__ long operator()(int i, int j, int k) const { return gidx(b, i, j, k); }
};


// Raw operands of the split Jacobian of neighbour (i,j,k)+off*e_dir: the
// neighbour's primitives, the metric at the neighbour with ghost indices
// clamped to the face, and the spectral radius clamped to 1..n.
template <typename T>
struct NbrRaw {
    T prim[5], gy[3], rad;
};

template <typename T, typename S, typename IX>
__device__ inline void nbr_load(const DBlock &b, const Operands<S> &O, const IX &ix, int i, int j, int k, int dir,
                                int off, NbrRaw<T> &r)
{
    int in = i + (dir == 0 ? off : 0), jn = j + (dir == 1 ? off : 0), kn = k + (dir == 2 ? off : 0);
    long gp = ix(in, jn, kn);
    int im = in, jm = jn, km = kn;
    if (im == 0) im = 1;
    if (im == b.ni + 1) im = b.ni;
    if (jm == 0) jm = 1;
    if (jm == b.nj + 1) jm = b.nj;
    if (km == 0) km = 1;
    if (km == b.nk + 1) km = b.nk;
    long gm = ix(im, jm, km);
    long gr;
    if (dir == 0) gr = ix(clampi(i + off, b.ni), j, k);
    else if (dir == 1) gr = ix(i, clampi(j + off, b.nj), k);
    else gr = ix(i, j, clampi(k + off, b.nk));
#pragma unroll
    for (int m = 0; m < 5; m++) r.prim[m] = (T)O.p[m][gp];
    // direction chosen by value, not by indexing the argument arrays (which
    // would copy them to scratch for a run-time dir)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const S *mc = dir == 0 ? O.m[0][c] : (dir == 1 ? O.m[1][c] : O.m[2][c]);
        r.gy[c] = (T)mc[gm];
    }
    const S *sr = dir == 0 ? O.sr[0] : (dir == 1 ? O.sr[1] : O.sr[2]);
    r.rad = (T)sr[gr];
}

template <int KIND>
constexpr int sweep_off() { return (KIND == 0 || KIND == 2) ? -1 : 1; }

// Every operand one node reads from memory, as loaded (no arithmetic yet).
template <int KIND, typename T>
struct NodeRaw {
    NbrRaw<T> nn[3];      // neighbours updated in this sweep, directions i, j, k
    NbrRaw<T> no[3];      // SGS: previous-sweep neighbours
    T dqo[3][5];          // SGS: their dQ (zero outside the block)
    T ghost[3][5];        // LU: dQ of a this-sweep neighbour outside the block
    T own[5];             // lusgs_u: the node's dQ from lusgs_l
    T rhs[5];
    T coed;
    T rv[3];              // LU, viscous: neighbours' viscous spectral radii
};

// Operands of one node that do not depend on this sweep's new dQ.
template <int KIND, typename T>
struct NodeIn {
    MxCoef<T> cn[3];      // neighbours updated in this sweep, directions i, j, k
    T fold[3][5];         // SGS: complete terms of the three previous-sweep neighbours
    T ghost[3][5];
    T own[5];
    T rhs[5];
    T coed;
    T rv[3];
};

// A: the dQ array the sweep pair works in; B: the SGS second buffer (A for the
// in-place schedule), both as five component pointers under the same index
// map.  gs_pr_l reads its previous-sweep neighbours from A, gs_pr_u from B.
template <int KIND, typename T, typename S, typename IX>
__device__ inline void node_load(const DBlock &b, const Operands<S> &O, const IX &ix, TPtrs<T> A, TPtrs<T> B,
                                 const DParams &P, int i, int j, int k, NodeRaw<KIND, T> &r)
{
    constexpr int off = sweep_off<KIND>();
    const long g = ix(i, j, k);
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) nbr_load<T>(b, O, ix, i, j, k, d, off, r.nn[d]);
#pragma unroll
    for (int m = 0; m < 5; m++) r.rhs[m] = (T)O.rhs[m][g];
    r.coed = (T)O.coed[g];
    if (KIND < 2) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int c = ijk[d] + off;
            if (c < 1 || c > nn[d]) {
                long gn = ix(i + (d == 0 ? off : 0), j + (d == 1 ? off : 0), k + (d == 2 ? off : 0));
#pragma unroll
                for (int m = 0; m < 5; m++) r.ghost[d][m] = A.c[m][gn];
            }
        }
        if (KIND == 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) r.own[m] = A.c[m][g];
        }
        if (P.nvis > 0) {
            r.rv[0] = (T)O.srv[0][ix(i + off, j, k)];
            r.rv[1] = (T)O.srv[1][ix(i, j + off, k)];
            r.rv[2] = (T)O.srv[2][ix(i, j, k + off)];
        }
    } else {
        const TPtrs<T> old = KIND == 2 ? A : B;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int c = ijk[d] - off;
            long gn = ix(i - (d == 0 ? off : 0), j - (d == 1 ? off : 0), k - (d == 2 ? off : 0));
#pragma unroll
            for (int m = 0; m < 5; m++) r.dqo[d][m] = (c < 1 || c > nn[d]) ? T(0.0) : old.c[m][gn];
            nbr_load<T>(b, O, ix, i, j, k, d, -off, r.no[d]);
        }
    }
}

// The dQ-independent arithmetic of a node: split-Jacobian coefficients of its
// this-sweep neighbours and, in the SGS sweeps, the complete terms of the
// previous-sweep ones.
template <int KIND, typename T>
__device__ inline void node_prep(const DParams &P, const NodeRaw<KIND, T> &r, NodeIn<KIND, T> &in)
{
    constexpr int off = sweep_off<KIND>();
    const T gama = (T)P.gama;
#pragma unroll
    for (int d = 0; d < 3; d++)
        mxdq_coef<T>(r.nn[d].prim, r.nn[d].gy, r.nn[d].rad, off < 0 ? T(1.0) : T(-1.0), gama, in.cn[d]);
#pragma unroll
    for (int m = 0; m < 5; m++) in.rhs[m] = r.rhs[m];
    in.coed = r.coed;
    if (KIND < 2) {
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int m = 0; m < 5; m++) in.ghost[d][m] = r.ghost[d][m];
#pragma unroll
        for (int m = 0; m < 5; m++) in.own[m] = r.own[m];
#pragma unroll
        for (int d = 0; d < 3; d++) in.rv[d] = r.rv[d];
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            MxCoef<T> co;
            mxdq_coef<T>(r.no[d].prim, r.no[d].gy, r.no[d].rad, -off < 0 ? T(1.0) : T(-1.0), gama, co);
            mxdq_apply<T>(co, r.dqo[d], in.fold[d]);
        }
    }
}

template <int KIND, typename T, typename S, typename IX>
__device__ inline void node_inputs(const DBlock &b, const Operands<S> &O, const IX &ix, TPtrs<T> A, TPtrs<T> B,
                                   const DParams &P, int i, int j, int k, NodeIn<KIND, T> &in)
{
    NodeRaw<KIND, T> r;
    node_load<KIND, T>(b, O, ix, A, B, P, i, j, k, r);
    node_prep<KIND, T>(P, r, in);
}

// The node's new dQ from its operands and the this-sweep neighbours' dQ nw.
template <int KIND, typename T>
__device__ inline void node_compute(const DParams &P, const NodeIn<KIND, T> &in, T (&nw)[3][5], T *out)
{
    T f[3][5];
#pragma unroll
    for (int d = 0; d < 3; d++) mxdq_apply<T>(in.cn[d], nw[d], f[d]);
    if (KIND < 2) {
        T rhs0[5];
#pragma unroll
        for (int m = 0; m < 5; m++) rhs0[m] = T(1.0) * (f[0][m] + f[1][m] + f[2][m]);
        if (P.nvis > 0) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                T v = T(0.5) * (in.rv[0] * nw[0][m] + in.rv[1] * nw[1][m] + in.rv[2] * nw[2][m]);
                rhs0[m] = KIND == 0 ? rhs0[m] + v : rhs0[m] - v;
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++)
            out[m] = KIND == 0 ? (-in.rhs[m] + rhs0[m]) * in.coed : in.own[m] - rhs0[m] * in.coed;
    } else {
        // gs_pr_l: lower = this sweep, upper = previous; gs_pr_u the other way round
#pragma unroll
        for (int m = 0; m < 5; m++) {
            T r0;
            if (KIND == 2)
                r0 = T(1.0) * (f[0][m] + f[1][m] + f[2][m] - in.fold[0][m] - in.fold[1][m] - in.fold[2][m]);
            else
                r0 = T(1.0) * (in.fold[0][m] + in.fold[1][m] + in.fold[2][m] - f[0][m] - f[1][m] - f[2][m]);
            out[m] = (r0 - in.rhs[m]) * in.coed;
        }
    }
}

// this-sweep neighbour outside the block: LU reads the dQ array there, SGS zero
template <int KIND, typename T>
__device__ inline void outside(const NodeIn<KIND, T> &in, int d, T *v)
{
#pragma unroll
    for (int m = 0; m < 5; m++) v[m] = KIND < 2 ? in.ghost[d][m] : T(0.0);
}

Operands<double> g3_operands(const DFields &F)
{
    Operands<double> O;
    const double *pr[5] = {F.r, F.u, F.v, F.w, F.p};
    const double *mt[3][3] = {{F.kcx, F.kcy, F.kcz}, {F.etx, F.ety, F.etz}, {F.ctx, F.cty, F.ctz}};
    for (int m = 0; m < 5; m++) { O.p[m] = pr[m]; O.rhs[m] = F.rhs + m * F.total; }
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < 3; c++) O.m[d][c] = mt[d][c];
    O.sr[0] = F.sra; O.sr[1] = F.srb; O.sr[2] = F.src;
    O.srv[0] = F.srva; O.srv[1] = F.srvb; O.srv[2] = F.srvc;
    O.coed = F.coed;
    return O;
}

using DqPtrs = TPtrs<double>;

template <int KIND>
__global__ void __launch_bounds__(256) k_sweep_plane(DBlock b, Operands<double> O, DqPtrs A, DParams P, int s)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int jlo = s - b.ni - b.nk, jhi = s - 2;
    jlo = jlo < 1 ? 1 : jlo;
    jhi = jhi > b.nj ? b.nj : jhi;
    if (jhi < jlo) return;
    long nj_ = jhi - jlo + 1;
    if (t >= nj_ * b.nk) return;
    int j = jlo + (int)(t % nj_);
    int k = (int)(t / nj_) + 1;
    int i = s - j - k;
    if (i < 1 || i > b.ni) return;
    constexpr int off = sweep_off<KIND>();
    G3Index ix{b};
    NodeIn<KIND, double> in;
    node_inputs<KIND, double>(b, O, ix, A, A, P, i, j, k, in);
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    double nw[3][5];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int c = ijk[d] + off;
        if (c < 1 || c > nn[d]) {
            outside<KIND, double>(in, d, nw[d]);
        } else {
            long gn = ix(i + (d == 0 ? off : 0), j + (d == 1 ? off : 0), k + (d == 2 ? off : 0));
#pragma unroll
            for (int m = 0; m < 5; m++) nw[d][m] = A.c[m][gn];
        }
    }
    double out[5];
    node_compute<KIND, double>(P, in, nw, out);
    long g = ix(i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) A.c[m][g] = out[m];
}

// resets exchange slots to XCH_EMPTY (arithmetic only ever yields quiet NaNs)
__global__ void k_fill_u64(unsigned long long *a, long n, unsigned long long v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

template <typename T>
__device__ inline void xch_put(double *slot, const T *v)
{
    xch_word *p = xch_ptr(slot);
#pragma unroll
    for (int m = 0; m < 5; m++)
        __hip_atomic_store(p + m, (unsigned long long)__double_as_longlong((double)v[m]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void xch_issue(double *slot, unsigned long long *raw)
{
    xch_word *p = xch_ptr(slot);
#pragma unroll
    for (int m = 0; m < 5; m++) raw[m] = __hip_atomic_load(p + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Complete a slot read issued earlier.  The common case -- all five words
// already written when they were read -- is one test on the registers, so that
// the only wait is for those five loads; the words still empty are re-read
// one by one.  The wait is bounded: after spin_limit re-reads, or once any
// lane of the launch has given up (err set), it stops and reports failure.
template <typename T>
__device__ inline bool xch_finish(double *slot, unsigned long long *raw, T *v, long spin_limit, int *err)
{
    bool ready = true;
#pragma unroll
    for (int m = 0; m < 5; m++) ready = ready && raw[m] != XCH_EMPTY;
    bool ok = true;
    if (!ready) {
        xch_word *p = xch_ptr(slot);
        long spins = 0;
#pragma unroll
        for (int m = 0; m < 5; m++) {
            while (raw[m] == XCH_EMPTY && ok) {
                if (++spins > spin_limit ||
                    ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                raw[m] = __hip_atomic_load(p + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 5; m++) v[m] = (T)__longlong_as_double((long long)raw[m]);
    return ok;
}

constexpr int PIPE_COLS = 64;                  // tile columns = lanes of the chain wavefront
constexpr int PIPE_MAX_THREADS = 8 * PIPE_COLS;  // chain, up to six helper wavefronts, exchange

// Operands of one tile's next hyperplane, formed by the helper wavefronts:
// two slots, structure of arrays over the tile's columns.  The 13 stored
// fields of a coefficient set are MxCoef without ae (= gama - 1).  Each sweep
// kind sizes only what it reads (a one-element stand-in otherwise): R for all
// but lusgs_u, which reads its own dQ instead; the dQ outside the block and
// the viscous radii for LU; the previous-sweep terms for SGS.  An array a
// kind does not read is a single element.  The exchange wavefront puts the
// upstream tiles' results for the tile's first row and column into xo.  In
// FP64 every kind then takes at most 79 KB, so that two workgroups fit in a
// CU's 160 KB.
template <bool USE, typename A, typename T>
using LdsArr = std::conditional_t<USE, A, T[1]>;
template <int KIND, typename T>
struct PipeLds {
    T cn[2][3][13][PIPE_COLS];
    LdsArr<(KIND >= 2), T[2][3][5][PIPE_COLS], T> fold;
    LdsArr<(KIND != 1), T[2][5][PIPE_COLS], T> rhs;
    T coed[2][PIPE_COLS];
    LdsArr<(KIND < 2), T[2][3][5][PIPE_COLS], T> ghost;
    LdsArr<(KIND == 1), T[2][5][PIPE_COLS], T> own;
    LdsArr<(KIND < 2), T[2][3][PIPE_COLS], T> rv;
    T sh[2][5][PIPE_COLS];       // this step's results, for the j and k neighbours in the tile
    T xo[2][2][5][PIPE_COLS];    // the j and k neighbours' results from the upstream tiles
};

template <int KIND, typename T>
__device__ inline void coef_put(PipeLds<KIND, T> &L, int slot, int d, int c, const MxCoef<T> &x)
{
    const T v[13] = {x.l1, x.x1, x.x2, x.nx, x.ny, x.nz, x.ct, x.af, x.c2, x.um, x.vm, x.wm, x.hm};
#pragma unroll
    for (int f = 0; f < 13; f++) L.cn[slot][d][f][c] = v[f];
}

template <int KIND, typename T>
__device__ inline void coef_get(const PipeLds<KIND, T> &L, int slot, int d, int c, T gama, MxCoef<T> &x)
{
    x.l1 = L.cn[slot][d][0][c]; x.x1 = L.cn[slot][d][1][c]; x.x2 = L.cn[slot][d][2][c];
    x.nx = L.cn[slot][d][3][c]; x.ny = L.cn[slot][d][4][c]; x.nz = L.cn[slot][d][5][c];
    x.ct = L.cn[slot][d][6][c]; x.af = L.cn[slot][d][7][c]; x.c2 = L.cn[slot][d][8][c];
    x.um = L.cn[slot][d][9][c]; x.vm = L.cn[slot][d][10][c]; x.wm = L.cn[slot][d][11][c];
    x.hm = L.cn[slot][d][12][c];
    x.ae = gama - T(1.0);
}

// One helper task: part p of the operands of node (i,j,k).  p = 0..2: the
// coefficients of the this-sweep neighbour in direction p, plus a share of
// the node's other operands (p = 0: R and the diagonal; p = 1: lusgs_u's own
// dQ; p = 2: the viscous radii; and, for LU, the dQ outside the block in
// direction p); p = 3..5 (SGS): the complete term of the previous-sweep
// neighbour in direction p-3.  Loaded into TaskRaw, formed into the LDS slot
// by task_prep.  The share sits at the same places of x for every p
// (x[0..4] a 5-vector, x[5] the diagonal, x[6..10] the outside dQ), so that
// no index depends on p.
template <typename T>
struct TaskRaw {
    NbrRaw<T> nb;
    T x[11];
};

template <int KIND, typename T, typename S, typename IX>
__device__ inline void task_load(int p, const DBlock &b, const Operands<S> &O, const IX &ix, TPtrs<T> A, TPtrs<T> B,
                                 const DParams &P, int i, int j, int k, TaskRaw<T> &r)
{
    constexpr int off = sweep_off<KIND>();
    const long g = ix(i, j, k);
    if (p < 3) {
        nbr_load<T>(b, O, ix, i, j, k, p, off, r.nb);
        if (p == 0) {
#pragma unroll
            for (int m = 0; m < 5; m++) r.x[m] = (T)O.rhs[m][g];
            r.x[5] = (T)O.coed[g];
        } else if (p == 1 && KIND == 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) r.x[m] = A.c[m][g];
        } else if (p == 2 && KIND < 2 && P.nvis > 0) {
            r.x[0] = (T)O.srv[0][ix(i + off, j, k)];
            r.x[1] = (T)O.srv[1][ix(i, j + off, k)];
            r.x[2] = (T)O.srv[2][ix(i, j, k + off)];
        }
        if (KIND < 2) {
            int cc = (p == 0 ? i : (p == 1 ? j : k)) + off;
            if (cc < 1 || cc > (p == 0 ? b.ni : (p == 1 ? b.nj : b.nk))) {
                long gn = ix(i + (p == 0 ? off : 0), j + (p == 1 ? off : 0), k + (p == 2 ? off : 0));
#pragma unroll
                for (int m = 0; m < 5; m++) r.x[6 + m] = A.c[m][gn];
            }
        }
    } else {
        const int d = p - 3;
        const TPtrs<T> old = KIND == 2 ? A : B;
        int cc = (d == 0 ? i : (d == 1 ? j : k)) - off;
        const int nd = d == 0 ? b.ni : (d == 1 ? b.nj : b.nk);
        long gn = ix(i - (d == 0 ? off : 0), j - (d == 1 ? off : 0), k - (d == 2 ? off : 0));
#pragma unroll
        for (int m = 0; m < 5; m++) r.x[m] = (cc < 1 || cc > nd) ? T(0.0) : old.c[m][gn];
        nbr_load<T>(b, O, ix, i, j, k, d, -off, r.nb);
    }
}

template <int KIND, typename T>
__device__ inline void task_prep(int p, const DParams &P, int i, int j, int k, const DBlock &b, const TaskRaw<T> &r,
                                 PipeLds<KIND, T> &L, int slot, int c)
{
    constexpr int off = sweep_off<KIND>();
    const T gama = (T)P.gama;
    if (p < 3) {
        MxCoef<T> x;
        mxdq_coef<T>(r.nb.prim, r.nb.gy, r.nb.rad, off < 0 ? T(1.0) : T(-1.0), gama, x);
        coef_put(L, slot, p, c, x);
        if (p == 0) {
            if constexpr (KIND != 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) L.rhs[slot][m][c] = r.x[m];
            }
            L.coed[slot][c] = r.x[5];
        }
        if constexpr (KIND == 1) {
            if (p == 1) {
#pragma unroll
                for (int m = 0; m < 5; m++) L.own[slot][m][c] = r.x[m];
            }
        }
        if constexpr (KIND < 2) {
            if (p == 2 && P.nvis > 0) {
#pragma unroll
                for (int d = 0; d < 3; d++) L.rv[slot][d][c] = r.x[d];
            }
            int cc = (p == 0 ? i : (p == 1 ? j : k)) + off;
            if (cc < 1 || cc > (p == 0 ? b.ni : (p == 1 ? b.nj : b.nk))) {
#pragma unroll
                for (int m = 0; m < 5; m++) L.ghost[slot][p][m][c] = r.x[6 + m];
            }
        }
    } else if constexpr (KIND >= 2) {
        MxCoef<T> x;
        mxdq_coef<T>(r.nb.prim, r.nb.gy, r.nb.rad, -off < 0 ? T(1.0) : T(-1.0), gama, x);
        T f[5];
        mxdq_apply<T>(x, r.x, f);
#pragma unroll
        for (int m = 0; m < 5; m++) L.fold[slot][p - 3][m][c] = f[m];
    }
}

// dQ of the neighbour outside the block: the staged ghost value for LU, 0 for SGS
template <int KIND, typename T>
__device__ inline T ghost_dq(const PipeLds<KIND, T> &L, int slot, int d, int m, int c)
{
    if constexpr (KIND < 2) return L.ghost[slot][d][m][c];
    else return T(0.0);
}

// The chain's operands of one node, read from the LDS slot all at once
// before any arithmetic, so that their latencies overlap.
template <int KIND, typename T>
struct ChainIn {
    MxCoef<T> x[3];
    T coed, rhs[5], own[5], rv[3], fold[3][5];
};

template <int KIND, typename T>
__device__ inline void chain_read(const DParams &P, const PipeLds<KIND, T> &L, int slot, int c, ChainIn<KIND, T> &q)
{
    const T gama = (T)P.gama;
#pragma unroll
    for (int d = 0; d < 3; d++) coef_get(L, slot, d, c, gama, q.x[d]);
    q.coed = L.coed[slot][c];
    if constexpr (KIND != 1) {
#pragma unroll
        for (int m = 0; m < 5; m++) q.rhs[m] = L.rhs[slot][m][c];
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) q.own[m] = L.own[slot][m][c];
    }
    if constexpr (KIND < 2) {
        if (P.nvis > 0) {
#pragma unroll
            for (int d = 0; d < 3; d++) q.rv[d] = L.rv[slot][d][c];
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int m = 0; m < 5; m++) q.fold[d][m] = L.fold[slot][d][m][c];
    }
}

// node_compute on the operands of chain_read (the same operations in the
// same order).
template <int KIND, typename T>
__device__ inline void chain_math(const DParams &P, const ChainIn<KIND, T> &q, T (&nw)[3][5], T *out)
{
    T f[3][5];
#pragma unroll
    for (int d = 0; d < 3; d++) mxdq_apply<T>(q.x[d], nw[d], f[d]);
    const T coed = q.coed;
    if constexpr (KIND < 2) {
        T rhs0[5];
#pragma unroll
        for (int m = 0; m < 5; m++) rhs0[m] = T(1.0) * (f[0][m] + f[1][m] + f[2][m]);
        if (P.nvis > 0) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                T v = T(0.5) * (q.rv[0] * nw[0][m] + q.rv[1] * nw[1][m] + q.rv[2] * nw[2][m]);
                rhs0[m] = KIND == 0 ? rhs0[m] + v : rhs0[m] - v;
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            if constexpr (KIND == 0) out[m] = (-q.rhs[m] + rhs0[m]) * coed;
            else out[m] = q.own[m] - rhs0[m] * coed;
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            const T fo0 = q.fold[0][m], fo1 = q.fold[1][m], fo2 = q.fold[2][m];
            T r0;
            if (KIND == 2)
                r0 = T(1.0) * (f[0][m] + f[1][m] + f[2][m] - fo0 - fo1 - fo2);
            else
                r0 = T(1.0) * (fo0 + fo1 + fo2 - f[0][m] - f[1][m] - f[2][m]);
            out[m] = (r0 - q.rhs[m]) * coed;
        }
    }
}

// Pipelined sweep, all blocks in one launch.  Workgroup w takes tile
// work[w].y of block work[w].x.  Wavefront 0 is the chain: lane c = (jl,kl)
// updates its column's node of the current hyperplane from the operands in
// LDS and the neighbours' new dQ.  Wavefronts 1..H are helpers: helper lane
// (c, p) forms part p of column c's operands for the next hyperplane (H = 3
// for LU, 6 for SGS) and, meanwhile, loads those of the hyperplane after.
// With PAIR (SGS only), three helpers form parts p and p + 3 each: four
// wavefronts instead of seven, so that two workgroups fit in a CU.
// O: the plane-major operands (every block at its own offset); A, B: the
// plane-major dQ pair.  lusgs_l/u write A, gs_pr_l reads A and writes B,
// gs_pr_u reads B and writes A (which it does not read), from where
// k_unskew_all copies the step's dQ to the field pool for the update.
template <int KIND, typename T, bool PAIR>
__global__ void __launch_bounds__(PIPE_MAX_THREADS) k_sweep_pipe(const SweepBlock *blocks, const int2 *work,
                                                                 Operands<T> O, TPtrs<T> A, TPtrs<T> B, DParams P,
                                                                 int *err, long spin_limit)
{
    static_assert(!PAIR || KIND >= 2, "paired helpers are for the SGS sweeps");
    constexpr bool fwd = (KIND == 0 || KIND == 2);
    constexpr int off = sweep_off<KIND>();
    constexpr int H = (KIND < 2 || PAIR) ? 3 : 6;
    const int2 wk = work[blockIdx.x];
    const SweepBlock sb = blocks[wk.x];
    const DBlock b = sb.b;
    const SkIndex ix = skew_index(b, sb.soff);
    const int TJ = sb.TJ, TK = sb.TK, nTJ = sb.nTJ, nTK = sb.nTK, ni = b.ni;
    const int tile = wk.y;
    const int J = tile % nTJ, K = tile / nTJ;
    // wave-uniform: kept in a scalar register, so that the helper's choice of
    // operand arrays by part is a scalar select, not a per-lane load of the
    // pointer from the kernel-argument block
    // Roles by wavefront: 0 the chain, 1..H the helpers, H+1 the exchange.
    // A workgroup's wavefronts go to the CU's four SIMDs in turn, so wave 4
    // shares the chain's SIMD: with eight wavefronts (SGS, six helpers) the
    // mostly sleeping exchange wavefront takes that place and the helpers
    // move up, so that no helper competes with the chain for issue slots.
    const int hw = __builtin_amdgcn_readfirstlane(threadIdx.x / PIPE_COLS);
    const int wave = H == 6 ? (hw == 4 ? H + 1 : (hw > 4 ? hw - 1 : hw)) : hw;
    const int c = threadIdx.x % PIPE_COLS;          // column of this lane
    const int p = wave - 1;                         // helper part (chain: -1)
    const int jl = c % TJ, kl = c / TJ;
    const int jlo = 1 + J * TJ, klo = 1 + K * TK;
    const int jn = min(TJ, b.nj - jlo + 1), kn = min(TK, b.nk - klo + 1);
    const bool col = kl < kn && jl < jn;
    const int j = jlo + jl, k = klo + kl;
    const int nsteps = ni + jn + kn - 2;
    const int i0 = fwd ? 1 - jl - kl : ni + (jn - 1 - jl) + (kn - 1 - kl);
    auto i_at = [&](int st) { return fwd ? i0 + st : i0 - st; };
    auto live = [&](int st) { int ii = i_at(st); return col && st < nsteps && ii >= 1 && ii <= ni; };
    const T gama = (T)P.gama;

    __shared__ PipeLds<KIND, T> L;

    // chain-wavefront state
    const int jsrc = fwd ? (jl > 0 ? 0 : (J > 0 ? 1 : 2)) : (jl < jn - 1 ? 0 : (J < nTJ - 1 ? 1 : 2));
    const int ksrc = fwd ? (kl > 0 ? 0 : (K > 0 ? 1 : 2)) : (kl < kn - 1 ? 0 : (K < nTK - 1 ? 1 : 2));
    const bool putj = col && (fwd ? (jl == jn - 1 && J < nTJ - 1) : (jl == 0 && J > 0));
    const bool putk = col && (fwd ? (kl == kn - 1 && K < nTK - 1) : (kl == 0 && K > 0));
    const long rowlen = (long)ni * 5;
    double *xj_in = sb.xj + (jsrc == 1 ? ((long)(tile + off) * TK + kl) * rowlen : 0);
    double *xk_in = sb.xk + (ksrc == 1 ? ((long)(tile + off * nTJ) * TJ + jl) * rowlen : 0);
    double *xj_out = sb.xj + ((long)tile * TK + kl) * rowlen;
    double *xk_out = sb.xk + ((long)tile * TJ + jl) * rowlen;
    const TPtrs<T> W = KIND == 2 ? B : A;

    // Roles, each a whole wavefront with a loop of its own (so that one
    // role's registers are free in the others'); every wavefront meets
    // 1 + nsteps barriers.
    //   chain     (p < 0): updates the tile's node of each hyperplane from
    //             LDS only -- it issues no global load, so it never waits for
    //             memory (on gfx950 a wait for a load also waits for every
    //             older store, the chain's exchange stores included);
    //   helpers   (0 <= p < H): form the operands of the next hyperplane;
    //   exchange  (p == H): completes the upstream tiles' words for the next
    //             hyperplane into xo, reading them two steps ahead.
    // Everything read from memory is issued two steps before it is used.  The
    // loops are unrolled by two so that each of two register sets keeps its
    // own registers (a copy from one set to the other would wait for the
    // loads in flight).
    if (p < 0) {
        T prev[5] = {T(0), T(0), T(0), T(0), T(0)};
        // where the j and k neighbours' dQ come from, fixed for the lane: LDS
        // index of the in-tile neighbour (or the lane's own, when it reads xo
        // or the block's ghost instead), so that every read below is
        // unconditional and only the choice is made per lane
        const int cj = jsrc == 0 ? c + off : c, ck = ksrc == 0 ? c + off * TJ : c;
        __syncthreads();
        for (int step = 0; step < nsteps; step++) {
            const int slot = step & 1;
            if (live(step)) {
                const int i = i_at(step);
                ChainIn<KIND, T> q;
                chain_read<KIND, T>(P, L, slot, c, q);
                const bool iin = i + off >= 1 && i + off <= ni;
                T nw[3][5];
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    const T g0 = ghost_dq(L, slot, 0, m, c), g1 = ghost_dq(L, slot, 1, m, c);
                    const T g2 = ghost_dq(L, slot, 2, m, c);
                    const T sj = L.sh[slot ^ 1][m][cj], sk = L.sh[slot ^ 1][m][ck];
                    const T xj = L.xo[slot][0][m][c], xk = L.xo[slot][1][m][c];
                    nw[0][m] = iin ? prev[m] : g0;
                    nw[1][m] = jsrc == 0 ? sj : (jsrc == 1 ? xj : g1);
                    nw[2][m] = ksrc == 0 ? sk : (ksrc == 1 ? xk : g2);
                }
                T out[5];
                chain_math<KIND, T>(P, q, nw, out);
                if (putj) xch_put<T>(xj_out + (long)(i - 1) * 5, out);
                if (putk) xch_put<T>(xk_out + (long)(i - 1) * 5, out);
                const long g = ix(i, j, k);
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    L.sh[slot][m][c] = out[m];
                    prev[m] = out[m];
                    W.c[m][g] = out[m];
                }
            }
            __syncthreads();
        }
    } else if (p == H) {
        bool bad = false;             // a bounded wait of this lane expired: stop waiting
        unsigned long long rjA[5], rkA[5], rjB[5], rkB[5];
        auto xissue = [&](int st, unsigned long long (&rj)[5], unsigned long long (&rk)[5]) {
            if (live(st)) {
                const int i = i_at(st);
                if (jsrc == 1) xch_issue(xj_in + (long)(i - 1) * 5, rj);
                if (ksrc == 1) xch_issue(xk_in + (long)(i - 1) * 5, rk);
            }
        };
        // complete the words of hyperplane st into LDS slot st & 1
        auto xfinish = [&](int st, unsigned long long (&rj)[5], unsigned long long (&rk)[5]) {
            if (live(st)) {
                const int i = i_at(st), sl = st & 1;
                T v[5];
                if (jsrc == 1) {
                    if (!xch_finish<T>(xj_in + (long)(i - 1) * 5, rj, v, bad ? 0 : spin_limit, err)) {
                        if (!bad) atomicExch(err, 1);
                        bad = true;
                    }
#pragma unroll
                    for (int m = 0; m < 5; m++) L.xo[sl][0][m][c] = v[m];
                }
                if (ksrc == 1) {
                    if (!xch_finish<T>(xk_in + (long)(i - 1) * 5, rk, v, bad ? 0 : spin_limit, err)) {
                        if (!bad) atomicExch(err, 1);
                        bad = true;
                    }
#pragma unroll
                    for (int m = 0; m < 5; m++) L.xo[sl][1][m][c] = v[m];
                }
            }
        };
        xissue(0, rjA, rkA);
        xfinish(0, rjA, rkA);
        xissue(1, rjB, rkB);
        xissue(2, rjA, rkA);
        __syncthreads();
        auto xstep = [&](int step, unsigned long long (&rj)[5], unsigned long long (&rk)[5]) {
            xfinish(step + 1, rj, rk);
            xissue(step + 3, rj, rk);
            __syncthreads();
        };
        for (int step = 0; step < nsteps; step += 2) {
            xstep(step, rjB, rkB);
            if (step + 1 < nsteps) xstep(step + 1, rjA, rkA);
        }
    } else {
        // One loop per helper part, the part a compile-time constant: a part
        // chosen at run time made the compiler keep the loaded operands in a
        // private array behind a part-dependent offset, which waits for every
        // load as soon as it is issued.
        auto helper_loop = [&](auto PC) {
            constexpr int PP = decltype(PC)::value;
            TaskRaw<T> rawA, rawB, raw2A, raw2B;
            auto hload = [&](int st, TaskRaw<T> &r, TaskRaw<T> &r2) {
                if (live(st)) {
                    task_load<KIND, T>(PP, b, O, ix, A, B, P, i_at(st), j, k, r);
                    if constexpr (PAIR) task_load<KIND, T>(PP + 3, b, O, ix, A, B, P, i_at(st), j, k, r2);
                }
            };
            auto hprep = [&](int st, const TaskRaw<T> &r, const TaskRaw<T> &r2, int sl) {
                if (live(st)) {
                    task_prep<KIND, T>(PP, P, i_at(st), j, k, b, r, L, sl, c);
                    if constexpr (PAIR) task_prep<KIND, T>(PP + 3, P, i_at(st), j, k, b, r2, L, sl, c);
                }
            };
            hload(0, rawA, raw2A);
            hprep(0, rawA, raw2A, 0);
            hload(1, rawB, raw2B);
            hload(2, rawA, raw2A);
            __syncthreads();
            // form hyperplane step + 1 from r/r2, reload them with hyperplane step + 3
            auto helper_step = [&](int step, TaskRaw<T> &r, TaskRaw<T> &r2) {
                hprep(step + 1, r, r2, (step & 1) ^ 1);
                hload(step + 3, r, r2);
                __syncthreads();
            };
            for (int step = 0; step < nsteps; step += 2) {
                helper_step(step, rawB, raw2B);
                if (step + 1 < nsteps) helper_step(step + 1, rawA, raw2A);
            }
        };
        switch (p) {
        case 0: helper_loop(std::integral_constant<int, 0>{}); break;
        case 1: helper_loop(std::integral_constant<int, 1>{}); break;
        case 2: helper_loop(std::integral_constant<int, 2>{}); break;
        case 3: if constexpr (H > 3) helper_loop(std::integral_constant<int, 3>{}); break;
        case 4: if constexpr (H > 4) helper_loop(std::integral_constant<int, 4>{}); break;
        case 5: if constexpr (H > 5) helper_loop(std::integral_constant<int, 5>{}); break;
        default:
            for (int step = 0; step <= nsteps; step++) __syncthreads();
            break;
        }
    }
}

// ------------------------------------------------------- plane-major staging
// Copies field-pool arrays (i fastest) into the plane-major copy: a workgroup
// takes rows j0..j0+15 of plane k and 64 consecutive hyperplanes s0..s0+63,
// reads each row's 64 i-values contiguously into LDS and writes each
// hyperplane's 16 j-values contiguously.  Nodes i,j,k in 0..n+1.
constexpr int STG_J = 16, STG_S = 64, STG_MAX = 24;

struct StageList {
    const void *src[STG_MAX];   // field-pool array (double, or float when f32 is set)
    int dst[STG_MAX];           // plane-major array index
    unsigned char f32[STG_MAX];
    unsigned char shell[STG_MAX];  // 1: only nodes on the ghost shell (i, j or k = 0 or n+1)
    int n;
};

// STG_G arrays per pass: every thread issues its 4 * STG_G loads before the
// first LDS write
constexpr int STG_G = 4;
// One STG_S x STG_J tile of hyperplanes x rows in plane k of block b.
template <typename T>
__device__ void skew_tile(const DBlock &b, const SkIndex &ix, const StageList &L, T *pool, long sktot, int k, int j0,
                          int s0, T (&tile)[STG_G][STG_J][STG_S + 1])
{
    const int t = threadIdx.x;
    for (int a0 = 0; a0 < L.n; a0 += STG_G) {
        // load: row jr = t / 16 handles 4 hyperplanes of each array of the pass
        T v[STG_G][4];
#pragma unroll
        for (int e = 0; e < STG_G; e++) {
            const int a = a0 + e;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int jr = t / 16, sc = (t % 16) + 16 * q;
                int j = j0 + jr, sv = s0 + sc, i = sv - j - k;
                v[e][q] = T(0);
                if (a < L.n && j >= 0 && j <= b.nj + 1 && i >= 0 && i <= b.ni + 1) {
                    long g = gidx(b, i, j, k);
                    v[e][q] = L.f32[a] ? (T)((const float *)L.src[a])[g] : (T)((const double *)L.src[a])[g];
                }
            }
        }
#pragma unroll
        for (int e = 0; e < STG_G; e++)
#pragma unroll
            for (int q = 0; q < 4; q++) tile[e][t / 16][(t % 16) + 16 * q] = v[e][q];
        __syncthreads();
        // store: hyperplane sc, rows jr contiguous
#pragma unroll
        for (int e = 0; e < STG_G; e++) {
            const int a = a0 + e;
            if (a >= L.n) break;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int jr = t % 16, sc = t / 16 + 16 * q;
                int j = j0 + jr, sv = s0 + sc, i = sv - j - k;
                if (j >= 0 && j <= b.nj + 1 && i >= 0 && i <= b.ni + 1) {
                    bool on_shell = i == 0 || i == b.ni + 1 || j == 0 || j == b.nj + 1 || k == 0 || k == b.nk + 1;
                    if (!L.shell[a] || on_shell) pool[(long)L.dst[a] * sktot + ix(i, j, k)] = tile[e][jr][sc];
                }
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_skew_stage(DBlock b, SkIndex ix, StageList L, T *pool, long sktot)
{
    __shared__ T tile[STG_G][STG_J][STG_S + 1];
    skew_tile(b, ix, L, pool, sktot, blockIdx.z, (int)blockIdx.y * STG_J - SK_JSHIFT, blockIdx.x * STG_S, tile);
}

__host__ __device__ inline int skew_tiles_s(const DBlock &b) { return (b.ni + b.nj + b.nk + 4 + STG_S - 1) / STG_S; }
// row tiles j0 = STG_J * t - SK_JSHIFT, so that each stores whole lines
__host__ __device__ inline int skew_tiles_j(const DBlock &b) { return (b.nj + 2 + SK_JSHIFT + STG_J - 1) / STG_J; }
// hyperplane tiles that cover rows j0..j0+STG_J-1 of one plane: nodes i = 0..ni+1
// lie on s = j + k + i, so s0 = j0 + k onwards spans ni + 1 + STG_J planes
__host__ __device__ inline int skew_tiles_row(const DBlock &b) { return (b.ni + 1 + STG_J + STG_S - 1) / STG_S; }

// All blocks in one launch: workgroup w stages tile G.wg[w].y of block
// G.wg[w].x, numbered hyperplane tile fastest, then row tile, then plane k.
// The hyperplane tiles of a row tile start at its first plane with a node.
template <typename T>
__global__ void __launch_bounds__(256) k_skew_stage_all(BlockGrid G, const long *skoff, StageList L, T *pool,
                                                        long sktot)
{
    __shared__ T tile[STG_G][STG_J][STG_S + 1];
    const int2 w = G.wg[blockIdx.x];
    const DBlock b = G.blocks[w.x];
    const int gs = skew_tiles_row(b), gj = skew_tiles_j(b);
    const int bs = w.y % gs, bj = (w.y / gs) % gj, k = w.y / (gs * gj);
    const SkIndex ix = skew_index(b, skoff[w.x]);
    const int j0 = bj * STG_J - SK_JSHIFT;
    skew_tile(b, ix, L, pool, sktot, k, j0, j0 + k + bs * STG_S, tile);
}

// The reverse copy for the sweeps' result: the plane-major dQ (src, five
// arrays sktot apart) of interior nodes into the field pool (dst, five arrays
// total apart).  Loads take whole lines of a hyperplane's rows, stores whole
// runs of i, through the same LDS tile; the same work-list as the staging.
template <typename T>
__global__ void __launch_bounds__(256) k_unskew_all(BlockGrid G, const long *skoff, const T *src, long sktot,
                                                    int ncomp, T *dst, long total)
{
    __shared__ T tile[STG_G][STG_J][STG_S + 1];
    const int2 w = G.wg[blockIdx.x];
    const DBlock b = G.blocks[w.x];
    const int gs = skew_tiles_row(b), gj = skew_tiles_j(b);
    const int bs = w.y % gs, bj = (w.y / gs) % gj, k = w.y / (gs * gj);
    if (k < 1 || k > b.nk) return;                 // the whole workgroup
    const SkIndex ix = skew_index(b, skoff[w.x]);
    const int j0 = bj * STG_J - SK_JSHIFT, s0 = j0 + k + bs * STG_S;
    const int t = threadIdx.x;
    for (int a0 = 0; a0 < ncomp; a0 += STG_G) {
        T v[STG_G][4];
#pragma unroll
        for (int e = 0; e < STG_G; e++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int a = a0 + e, jr = t % 16, sc = t / 16 + 16 * q;
                const int j = j0 + jr, i = s0 + sc - j - k;
                v[e][q] = T(0);
                if (a < ncomp && j >= 1 && j <= b.nj && i >= 1 && i <= b.ni) v[e][q] = src[a * sktot + ix(i, j, k)];
            }
#pragma unroll
        for (int e = 0; e < STG_G; e++)
#pragma unroll
            for (int q = 0; q < 4; q++) tile[e][t % 16][t / 16 + 16 * q] = v[e][q];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < STG_G; e++) {
            const int a = a0 + e;
            if (a >= ncomp) break;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int jr = t / 16, sc = (t % 16) + 16 * q;
                const int j = j0 + jr, i = s0 + sc - j - k;
                if (j >= 1 && j <= b.nj && i >= 1 && i <= b.ni) dst[a * total + gidx(b, i, j, k)] = tile[e][jr][sc];
            }
        }
        __syncthreads();
    }
}

}  // namespace

int launch_unskew_dq(const BlockGrid &G, const long *skoff, int prec, void *pool, long sktot, double *dq, float *dq32,
                     long total, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    if (prec == 1)
        hipLaunchKernelGGL(k_unskew_all<float>, dim3(G.nwg), dim3(256), 0, s, G, skoff,
                           (const float *)pool + (long)SK_A * sktot, sktot, 5, dq32, total);
    else
        hipLaunchKernelGGL(k_unskew_all<double>, dim3(G.nwg), dim3(256), 0, s, G, skoff,
                           (const double *)pool + (long)SK_A * sktot, sktot, 5, dq, total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_unskew_f64(const BlockGrid &G, const long *skoff, const double *pool, long sktot, int slot0, int ncomp,
                      double *dst, long total, hipStream_t s)
{
    if (G.nwg == 0) return 0;
    if (ncomp < 1 || ncomp > 5) return -2;
    hipLaunchKernelGGL(k_unskew_all<double>, dim3(G.nwg), dim3(256), 0, s, G, skoff, pool + (long)slot0 * sktot,
                       sktot, ncomp, dst, total);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long skew_stage_wgs(const DBlock &b) { return (long)skew_tiles_row(b) * skew_tiles_j(b) * (b.nk + 2); }

int launch_sweep(int kind, const DBlock &b, const DFields &F, const DParams &P, hipStream_t s)
{
    int smin = 3, smax = b.ni + b.nj + b.nk;
    long nthr = (long)b.nj * b.nk;
    dim3 g = grid1(nthr, 256), bl(256);
    Operands<double> O = g3_operands(F);
    DqPtrs A;
    for (int m = 0; m < 5; m++) A.c[m] = F.dq + m * F.total;
    switch (kind) {
    case 0: for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_sweep_plane<0>, g, bl, 0, s, b, O, A, P, p); break;
    case 1: for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_sweep_plane<1>, g, bl, 0, s, b, O, A, P, p); break;
    case 2: for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_sweep_plane<2>, g, bl, 0, s, b, O, A, P, p); break;
    default: for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_sweep_plane<3>, g, bl, 0, s, b, O, A, P, p); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long sweep_exchange_words(int TJ, int TK, int nTJ, int nTK, int ni)
{
    return (long)nTJ * nTK * (TJ + TK) * ni * 5;
}

long skew_size(const DBlock &b)
{
    return (long)(b.ni + b.nj + b.nk + 4) * skew_pitch(b) * (b.nk + 2);
}

static StageList stage_list(int what, const DFields &F, int prec)
{
    StageList L{};
    auto add = [&](const void *src, int dst, int f32, int shell) {
        L.src[L.n] = src; L.dst[L.n] = dst; L.f32[L.n] = (unsigned char)f32; L.shell[L.n] = (unsigned char)shell;
        L.n++;
    };
    if (what == SKEW_METRICS) {
        const double *mt[9] = {F.kcx, F.kcy, F.kcz, F.etx, F.ety, F.etz, F.ctx, F.cty, F.ctz};
        for (int c = 0; c < 9; c++) add(mt[c], SK_M + c, 0, 0);
    } else {
        const double *pr[5] = {F.r, F.u, F.v, F.w, F.p};
        for (int m = 0; m < 5; m++) add(pr[m], SK_P + m, 0, 0);
        add(F.sra, SK_SR + 0, 0, 0); add(F.srb, SK_SR + 1, 0, 0); add(F.src, SK_SR + 2, 0, 0);
        if (F.srva) { add(F.srva, SK_SRV + 0, 0, 0); add(F.srvb, SK_SRV + 1, 0, 0); add(F.srvc, SK_SRV + 2, 0, 0); }
        for (int m = 0; m < 5; m++) add(F.rhs + m * F.total, SK_RHS + m, 0, 0);
        add(F.coed, SK_COED, 0, 0);
        // dQ outside the block (read by the LU sweeps) comes from the field pool's ghost layer
        for (int m = 0; m < 5; m++) {
            if (prec == 1) add(F.dq32 + m * F.total, SK_A + m, 1, 1);
            else add(F.dq + m * F.total, SK_A + m, 0, 1);
        }
    }
    return L;
}

int launch_skew_stage(int what, const DBlock &b, long soff, const DFields &F, int prec, void *pool, long sktot,
                      hipStream_t s)
{
    const StageList L = stage_list(what, F, prec);
    dim3 g(skew_tiles_s(b), skew_tiles_j(b), b.nk + 2);
    SkIndex ix = skew_index(b, soff);
    if (prec == 1) hipLaunchKernelGGL(k_skew_stage<float>, g, dim3(256), 0, s, b, ix, L, (float *)pool, sktot);
    else hipLaunchKernelGGL(k_skew_stage<double>, g, dim3(256), 0, s, b, ix, L, (double *)pool, sktot);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_skew_stage_f64(const BlockGrid &G, const long *skoff, const double *const *src, const int *slot, int n,
                          double *pool, long sktot, hipStream_t s)
{
    if (n < 1 || n > STG_MAX) return -2;
    StageList L{};
    for (int a = 0; a < n; a++) { L.src[a] = src[a]; L.dst[a] = slot[a]; L.f32[a] = 0; L.shell[a] = 0; }
    L.n = n;
    if (G.nwg == 0) return 0;
    hipLaunchKernelGGL(k_skew_stage_all<double>, dim3(G.nwg), dim3(256), 0, s, G, skoff, L, pool, sktot);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_skew_stage_all(int what, const BlockGrid &G, const long *skoff, const DFields &F, int prec, void *pool,
                          long sktot, hipStream_t s)
{
    const StageList L = stage_list(what, F, prec);
    if (G.nwg == 0) return 0;
    if (prec == 1)
        hipLaunchKernelGGL(k_skew_stage_all<float>, dim3(G.nwg), dim3(256), 0, s, G, skoff, L, (float *)pool, sktot);
    else hipLaunchKernelGGL(k_skew_stage_all<double>, dim3(G.nwg), dim3(256), 0, s, G, skoff, L, (double *)pool, sktot);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T>
static int launch_pipe(int kind, const SweepPlan &S, const DFields &F, const DParams &P, T *pool, long sktot,
                       hipStream_t s)
{
    Operands<T> O;
    for (int m = 0; m < 5; m++) { O.p[m] = pool + (SK_P + m) * sktot; O.rhs[m] = pool + (SK_RHS + m) * sktot; }
    for (int d = 0; d < 3; d++) {
        for (int c = 0; c < 3; c++) O.m[d][c] = pool + (SK_M + 3 * d + c) * sktot;
        O.sr[d] = pool + (SK_SR + d) * sktot;
        O.srv[d] = pool + (SK_SRV + d) * sktot;
    }
    O.coed = pool + SK_COED * sktot;
    TPtrs<T> A, B;
    for (int m = 0; m < 5; m++) {
        A.c[m] = pool + (SK_A + m) * sktot;
        B.c[m] = pool + (SK_B + m) * sktot;
    }
    const int2 *work = (kind == 0 || kind == 2) ? S.work_fwd : S.work_bwd;
    const bool pair = kind >= 2 && S.sgs_paired;
    dim3 g(S.nwork), bl(PIPE_COLS * ((kind < 2 || pair) ? 5 : 8));
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_sweep_pipe<0, T, false>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit); break;
    case 1: hipLaunchKernelGGL((k_sweep_pipe<1, T, false>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit); break;
    case 2:
        if (pair) hipLaunchKernelGGL((k_sweep_pipe<2, T, true>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit);
        else hipLaunchKernelGGL((k_sweep_pipe<2, T, false>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit);
        break;
    default:
        if (pair) hipLaunchKernelGGL((k_sweep_pipe<3, T, true>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit);
        else hipLaunchKernelGGL((k_sweep_pipe<3, T, false>), g, bl, 0, s, S.blocks, work, O, A, B, P, S.err, S.spin_limit);
        break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_fill_u64(unsigned long long *a, long n, unsigned long long v, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_u64, grid1(n, 256), dim3(256), 0, s, a, n, v);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_fill_xch(const SweepPlan &S, hipStream_t s) { return launch_fill_u64(S.xch, S.nxch, XCH_EMPTY, s); }

int launch_sweep_batched(int kind, const SweepPlan &S, const DFields &F, const DParams &P, int prec, void *pool,
                         long sktot, hipStream_t s)
{
    if (S.lanes > PIPE_COLS || S.lanes < 1 || S.nwork < 1) return -2;
    hipLaunchKernelGGL(k_fill_u64, grid1(S.nxch, 256), dim3(256), 0, s, S.xch, S.nxch, XCH_EMPTY);
    if (prec == 1) return launch_pipe<float>(kind, S, F, P, (float *)pool, sktot, s);
    return launch_pipe<double>(kind, S, F, P, (double *)pool, sktot, s);
}


// ------------------------------------------------ tangent and adjoint of the sweeps
// The four sweeps of one step, for the derivatives of a whole step.  Each
// sweep writes its own array (x_0..x_3): the values of the in-place schedule,
// since there lusgs_u reads lusgs_l's dQ at its own node and each SGS sweep
// reads the previous sweep's dQ at the nodes still ahead of it.  In those
// terms, with A+(n,d) / A-(n,d) the split Jacobian of node n in direction d
// with npn = +1 / -1 (mxdq_apply on mxdq_coef of n's primitives, metric row d
// and radius d) and c the diagonal coed:
//   x0_i = c_i (-R_i + sum_d A+(i-e_d) x0_{i-e_d})                  ascending
//   x1_i = x0_i - c_i sum_d A-(i+e_d) x1_{i+e_d}                     descending
//   x2_i = c_i (sum_d A+(i-e_d) x2_{i-e_d} - A-(i+e_d) x1_{i+e_d} - R_i)   ascending
//   x3_i = c_i (sum_d A+(i-e_d) x2_{i-e_d} - A-(i+e_d) x3_{i+e_d} - R_i)   descending
// with every term of a neighbour outside 1..n zero (its dQ is).  The tangent
// runs the solver's own node arithmetic in D1; the adjoint solves the
// transposed systems in the opposite order, each node applying its own
// transposed matrices to its neighbours' multipliers, then differentiates
// mu^T A(n,d) x_n with respect to n's primitives and radius.
namespace {

template <int KIND>
__device__ inline void raw_merge(const NodeRaw<KIND, double> &a, const NodeRaw<KIND, double> &t, NodeRaw<KIND, D1> &r)
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            r.nn[d].prim[m] = D1(a.nn[d].prim[m], t.nn[d].prim[m]);
            r.no[d].prim[m] = D1(a.no[d].prim[m], t.no[d].prim[m]);
            r.dqo[d][m] = D1(a.dqo[d][m], t.dqo[d][m]);
            r.ghost[d][m] = D1(a.ghost[d][m], t.ghost[d][m]);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            r.nn[d].gy[c] = D1(a.nn[d].gy[c], t.nn[d].gy[c]);
            r.no[d].gy[c] = D1(a.no[d].gy[c], t.no[d].gy[c]);
        }
        r.nn[d].rad = D1(a.nn[d].rad, t.nn[d].rad);
        r.no[d].rad = D1(a.no[d].rad, t.no[d].rad);
        r.rv[d] = D1(a.rv[d], t.rv[d]);
    }
#pragma unroll
    for (int m = 0; m < 5; m++) {
        r.own[m] = D1(a.own[m], t.own[m]);
        r.rhs[m] = D1(a.rhs[m], t.rhs[m]);
    }
    r.coed = D1(a.coed, t.coed);
}

__device__ inline bool plane_node(const DBlock &b, int s, int &i, int &j, int &k)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int jlo = s - b.ni - b.nk, jhi = s - 2;
    jlo = jlo < 1 ? 1 : jlo;
    jhi = jhi > b.nj ? b.nj : jhi;
    if (jhi < jlo) return false;
    long nj_ = jhi - jlo + 1;
    if (t >= nj_ * b.nk) return false;
    j = jlo + (int)(t % nj_);
    k = (int)(t / nj_) + 1;
    i = s - j - k;
    return i >= 1 && i <= b.ni;
}

// One hyperplane of sweep KIND into X (and its tangent Xt when TAN).  A is the
// array lusgs_l/u read outside the block and lusgs_u at its own node (x0),
// or gs_pr_l's previous sweep (x1); B gs_pr_u's previous sweep (x2).
template <int KIND, bool TAN>
__global__ void __launch_bounds__(256) k_ad_sweep(DBlock b, Operands<double> O, Operands<double> Ot, DqPtrs A,
                                                  DqPtrs B, DqPtrs At, DqPtrs Bt, DqPtrs X, DqPtrs Xt, DParams P,
                                                  int s)
{
    int i, j, k;
    if (!plane_node(b, s, i, j, k)) return;
    constexpr int off = sweep_off<KIND>();
    using T = typename std::conditional<TAN, D1, double>::type;
    G3Index ix{b};
    NodeIn<KIND, T> in;
    NodeRaw<KIND, double> rv{};
    node_load<KIND, double>(b, O, ix, A, B, P, i, j, k, rv);
    if constexpr (TAN) {
        NodeRaw<KIND, double> rt{};
        node_load<KIND, double>(b, Ot, ix, At, Bt, P, i, j, k, rt);
        NodeRaw<KIND, D1> r;
        raw_merge<KIND>(rv, rt, r);
        node_prep<KIND, D1>(P, r, in);
    } else {
        node_prep<KIND, double>(P, rv, in);
    }
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    T nw[3][5];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int c = ijk[d] + off;
        if (c < 1 || c > nn[d]) {
            outside<KIND, T>(in, d, nw[d]);
        } else {
            long gn = ix(i + (d == 0 ? off : 0), j + (d == 1 ? off : 0), k + (d == 2 ? off : 0));
#pragma unroll
            for (int m = 0; m < 5; m++) {
                if constexpr (TAN) nw[d][m] = D1(X.c[m][gn], Xt.c[m][gn]);
                else nw[d][m] = X.c[m][gn];
            }
        }
    }
    T out[5];
    node_compute<KIND, T>(P, in, nw, out);
    const long g = ix(i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) {
        X.c[m][g] = val(out[m]);
        if constexpr (TAN) Xt.c[m][g] = out[m].d;
    }
}

// g = M^T mu for the split Jacobian M of mxdq_apply: M = l1 I + x1 (c2 n w^T -
// u h^T) + x2 (n h^T - u w^T), w = (ct, -nx, -ny, -nz, 0), h = (af, -ae um,
// -ae vm, -ae wm, ae), n = (0, nx, ny, nz, ct), u = (1, um, vm, wm, hm).
__device__ inline void mxdq_apply_t(const MxCoef<double> &c, const double *mu, double *g)
{
    double nm = c.nx * mu[1] + c.ny * mu[2] + c.nz * mu[3] + c.ct * mu[4];
    double um = mu[0] + c.um * mu[1] + c.vm * mu[2] + c.wm * mu[3] + c.hm * mu[4];
    double aw = c.x1 * c.c2 * nm - c.x2 * um;
    double ah = c.x2 * nm - c.x1 * um;
    g[0] = c.l1 * mu[0] + aw * c.ct + ah * c.af;
    g[1] = c.l1 * mu[1] - aw * c.nx - ah * c.ae * c.um;
    g[2] = c.l1 * mu[2] - aw * c.ny - ah * c.ae * c.vm;
    g[3] = c.l1 * mu[3] - aw * c.nz - ah * c.ae * c.wm;
    g[4] = c.l1 * mu[4] + ah * c.ae;
}

// the operands of node g's own split Jacobian in direction D
template <int D, typename T>
__device__ inline void own_raw(const Operands<double> &O, long g, T *prim, T *gy, T &rad)
{
#pragma unroll
    for (int m = 0; m < 5; m++) prim[m] = O.p[m][g];
#pragma unroll
    for (int c = 0; c < 3; c++) gy[c] = O.m[D][c][g];
    rad = O.sr[D][g];
}

struct TSweep {
    const double *x[4][5];   // the four sweeps' dQ
    double *mu[4][5];        // their multipliers
    double *xb[4][5];        // adjoints of their outputs
    double *rb[5];           // adjoint of R (accumulated)
    double *cb;              // adjoint of coed (accumulated)
    double *lv[3];           // adjoint of the viscous radii (accumulated; LU, nvis = 1)
};

// One hyperplane of the transposed sweep KIND (processed in the opposite
// order to the sweep): z_i = xbar_i +- sum_d A(i,d)^T mu_{i-+e_d}, mu_i = c_i z_i,
// and the adjoints of R, coed and (lusgs_u) x0 that z carries.  With viscous
// radii rv the LU sweeps add 0.5 rv_d(n) x_n for each neighbour n (lusgs_l)
// or subtract it inside the product with c (lusgs_u): z also takes
// 0.5 rv_d(i) mu_{i-+e_d}, and rv_d(i) the adjoint 0.5 mu_{i-+e_d} . x_i.
template <int KIND>
__global__ void __launch_bounds__(256) k_ad_tsweep(DBlock b, Operands<double> O, TSweep S, DParams P, int s)
{
    int i, j, k;
    if (!plane_node(b, s, i, j, k)) return;
    G3Index ix{b};
    const long g = ix(i, j, k);
    constexpr bool up = (KIND == 0 || KIND == 2);   // multipliers of the nodes ahead, A+ of this node
    const double npn = up ? 1.0 : -1.0;
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    double z[5];
#pragma unroll
    for (int m = 0; m < 5; m++) z[m] = S.xb[KIND][m][g];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int c = ijk[d] + (up ? 1 : -1);
        if (c < 1 || c > nn[d]) continue;
        const long gn = ix(i + (d == 0 ? (up ? 1 : -1) : 0), j + (d == 1 ? (up ? 1 : -1) : 0),
                           k + (d == 2 ? (up ? 1 : -1) : 0));
        double prim[5], gy[3], rad;
        if (d == 0) own_raw<0, double>(O, g, prim, gy, rad);
        else if (d == 1) own_raw<1, double>(O, g, prim, gy, rad);
        else own_raw<2, double>(O, g, prim, gy, rad);
        MxCoef<double> cf;
        mxdq_coef<double>(prim, gy, rad, npn, P.gama, cf);
        double mu[5], tg[5];
#pragma unroll
        for (int m = 0; m < 5; m++) mu[m] = S.mu[KIND][m][gn];
        mxdq_apply_t(cf, mu, tg);
        if (KIND < 2 && P.nvis > 0) {
            const double rv = d == 0 ? O.srv[0][g] : (d == 1 ? O.srv[1][g] : O.srv[2][g]);
            double lv = 0.0;
#pragma unroll
            for (int m = 0; m < 5; m++) {
                tg[m] = KIND == 0 ? tg[m] + 0.5 * rv * mu[m] : tg[m] - 0.5 * rv * mu[m];
                lv = lv + mu[m] * S.x[KIND][m][g];
            }
            double *lvd = d == 0 ? S.lv[0] : (d == 1 ? S.lv[1] : S.lv[2]);
            lvd[g] = lvd[g] + 0.5 * lv;
        }
#pragma unroll
        for (int m = 0; m < 5; m++) z[m] = up ? z[m] + tg[m] : z[m] - tg[m];
    }
    const double coed = O.coed[g];
    double cb = 0.0;
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double mu = coed * z[m];
        S.mu[KIND][m][g] = mu;
        double sm;   // d x_i / d coed
        if (KIND == 1) {
            S.xb[0][m][g] = S.xb[0][m][g] + z[m];
            sm = (S.x[1][m][g] - S.x[0][m][g]) / coed;
        } else {
            S.rb[m][g] = S.rb[m][g] - mu;
            sm = S.x[KIND][m][g] / coed;
        }
        cb = cb + z[m] * sm;
    }
    S.cb[g] = S.cb[g] + cb;
}

// After the transposed gs_pr_u: xbar2_n += sum_d A+(n,d)^T mu3_{n+e_d}; after
// the transposed gs_pr_l: xbar1_n -= sum_d A-(n,d)^T mu2_{n-e_d}.
template <int KIND>
__global__ void __launch_bounds__(256) k_ad_tgather(DBlock b, Operands<double> O, TSweep S, DParams P)
{
    const long n = (long)b.ni * b.nj * b.nk;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    G3Index ix{b};
    const long g = ix(i, j, k);
    constexpr bool up = KIND == 3;
    constexpr int dst = KIND == 3 ? 2 : 1;
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int c = ijk[d] + (up ? 1 : -1);
        if (c < 1 || c > nn[d]) continue;
        const long gn = ix(i + (d == 0 ? (up ? 1 : -1) : 0), j + (d == 1 ? (up ? 1 : -1) : 0),
                           k + (d == 2 ? (up ? 1 : -1) : 0));
        double prim[5], gy[3], rad;
        if (d == 0) own_raw<0, double>(O, g, prim, gy, rad);
        else if (d == 1) own_raw<1, double>(O, g, prim, gy, rad);
        else own_raw<2, double>(O, g, prim, gy, rad);
        MxCoef<double> cf;
        mxdq_coef<double>(prim, gy, rad, up ? 1.0 : -1.0, P.gama, cf);
        double mu[5], tg[5];
#pragma unroll
        for (int m = 0; m < 5; m++) mu[m] = S.mu[KIND][m][gn];
        mxdq_apply_t(cf, mu, tg);
#pragma unroll
        for (int m = 0; m < 5; m++) acc[m] = acc[m] + tg[m];
    }
#pragma unroll
    for (int m = 0; m < 5; m++) S.xb[dst][m][g] = up ? S.xb[dst][m][g] + acc[m] : S.xb[dst][m][g] - acc[m];
}

// sum over pairs p of wt_p mu_p^T A(coef) x_p, in D1, along seed q of
// (primitives 0..4, radius 5)
__device__ inline double coef_pairs(const double *prim, const double *gy, double rad, double npn, double gama,
                                    int q, int np, const double (*mu)[5], const double (*x)[5])
{
    D1 pr[5], gd[3];
#pragma unroll
    for (int m = 0; m < 5; m++) pr[m] = D1(prim[m], q == m ? 1.0 : 0.0);
#pragma unroll
    for (int c = 0; c < 3; c++) gd[c] = D1(gy[c]);
    MxCoef<D1> cf;
    mxdq_coef<D1>(pr, gd, D1(rad, q == 5 ? 1.0 : 0.0), D1(npn), D1(gama), cf);
    double s = 0.0;
    for (int p = 0; p < np; p++) {
        D1 xx[5], f[5];
#pragma unroll
        for (int m = 0; m < 5; m++) xx[m] = D1(x[p][m]);
        mxdq_apply<D1>(cf, xx, f);
#pragma unroll
        for (int m = 0; m < 5; m++) s = s + mu[p][m] * f[m].d;
    }
    return s;
}

// Gradients of sum mu^T A(n,d) x_n over the sweeps' pairs w.r.t. node n's
// primitives (lp) and spectral radii (lsr), accumulated:
//   A+(n,d): (mu0, x0), (mu2, x2), (mu3, x2) at n+e_d;
//   A-(n,d): (-mu1, x1), (-mu2, x1), (-mu3, x3) at n-e_d.
__global__ void __launch_bounds__(256) k_ad_tcoef(DBlock b, Operands<double> O, TSweep S, DParams P, double *lp,
                                                  double *lsr, long total)
{
    const long n = (long)b.ni * b.nj * b.nk;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % b.ni) + 1, j = (int)((t / b.ni) % b.nj) + 1, k = (int)(t / ((long)b.ni * b.nj)) + 1;
    G3Index ix{b};
    const long g = ix(i, j, k);
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    double gp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double prim[5], gy[3], rad;
        if (d == 0) own_raw<0, double>(O, g, prim, gy, rad);
        else if (d == 1) own_raw<1, double>(O, g, prim, gy, rad);
        else own_raw<2, double>(O, g, prim, gy, rad);
        double gr = 0.0;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int c = ijk[d] + (side == 0 ? 1 : -1);
            if (c < 1 || c > nn[d]) continue;
            const long gn = ix(i + (d == 0 ? (side == 0 ? 1 : -1) : 0), j + (d == 1 ? (side == 0 ? 1 : -1) : 0),
                               k + (d == 2 ? (side == 0 ? 1 : -1) : 0));
            double mu[3][5], x[3][5];
            const int ka[3] = {0, 2, 3}, xa[3] = {0, 2, 2};      // side 0: A+
            const int kb[3] = {1, 2, 3}, xbb[3] = {1, 1, 3};     // side 1: A-
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const int km = side == 0 ? ka[p] : kb[p], kx = side == 0 ? xa[p] : xbb[p];
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    const double mv = km == 0 ? S.mu[0][m][gn] : km == 1 ? S.mu[1][m][gn] :
                                      km == 2 ? S.mu[2][m][gn] : S.mu[3][m][gn];
                    mu[p][m] = side == 0 ? mv : -mv;
                    x[p][m] = kx == 0 ? S.x[0][m][g] : kx == 1 ? S.x[1][m][g] : kx == 2 ? S.x[2][m][g] : S.x[3][m][g];
                }
            }
            const double npn = side == 0 ? 1.0 : -1.0;
            for (int q = 0; q < 6; q++) {
                const double v = coef_pairs(prim, gy, rad, npn, P.gama, q, 3, mu, x);
                if (q < 5) gp[q] = gp[q] + v;
                else gr = gr + v;
            }
        }
        lsr[d * total + g] = lsr[d * total + g] + gr;
    }
#pragma unroll
    for (int m = 0; m < 5; m++) lp[m * total + g] = lp[m * total + g] + gp[m];
}

Operands<double> tan_operands(const DFields &F, const AdSweep &A)
{
    Operands<double> O;
    const long T = F.total;
    for (int m = 0; m < 5; m++) { O.p[m] = A.tp + m * T; O.rhs[m] = A.trhs + m * T; }
    for (int d = 0; d < 3; d++) {
        for (int c = 0; c < 3; c++) O.m[d][c] = A.zero;
        O.sr[d] = A.tsr + d * T;
        O.srv[d] = A.tsrv ? A.tsrv + d * T : A.zero;
    }
    O.coed = A.tcoed;
    return O;
}

DqPtrs ptrs5(double *base, long total)
{
    DqPtrs p;
    for (int m = 0; m < 5; m++) p.c[m] = base + m * total;
    return p;
}

}  // namespace

int launch_ad_sweeps_fwd(const DBlock &b, const DFields &F, const DParams &P, const AdSweep &A, hipStream_t s)
{
    const int smin = 3, smax = b.ni + b.nj + b.nk;
    const long T = F.total;
    dim3 g = grid1((long)b.nj * b.nk, 256), bl(256);
    Operands<double> O = g3_operands(F);
    const bool tan = A.tx != nullptr;
    Operands<double> Ot = tan ? tan_operands(F, A) : O;
    DqPtrs X[4], Xt[4];
    for (int q = 0; q < 4; q++) {
        X[q] = ptrs5(A.x + 5 * q * T, T);
        Xt[q] = tan ? ptrs5(A.tx + 5 * q * T, T) : X[q];
    }
#define AD_SWEEP(KIND, AA, BB, AAt, BBt)                                                                             \
    if (tan) hipLaunchKernelGGL((k_ad_sweep<KIND, true>), g, bl, 0, s, b, O, Ot, AA, BB, AAt, BBt, X[KIND], Xt[KIND], \
                                P, p);                                                                               \
    else hipLaunchKernelGGL((k_ad_sweep<KIND, false>), g, bl, 0, s, b, O, Ot, AA, BB, AAt, BBt, X[KIND], Xt[KIND], P, p)
    for (int p = smin; p <= smax; p++) { AD_SWEEP(0, X[0], X[0], Xt[0], Xt[0]); }
    for (int p = smax; p >= smin; p--) { AD_SWEEP(1, X[0], X[0], Xt[0], Xt[0]); }
    for (int p = smin; p <= smax; p++) { AD_SWEEP(2, X[1], X[1], Xt[1], Xt[1]); }
    for (int p = smax; p >= smin; p--) { AD_SWEEP(3, X[2], X[2], Xt[2], Xt[2]); }
#undef AD_SWEEP
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_ad_sweeps_rev(const DBlock &b, const DFields &F, const DParams &P, const AdSweep &A, hipStream_t s)
{
    const int smin = 3, smax = b.ni + b.nj + b.nk;
    const long T = F.total;
    const long nn = (long)b.ni * b.nj * b.nk;
    dim3 g = grid1((long)b.nj * b.nk, 256), gn = grid1(nn, 256), bl(256);
    Operands<double> O = g3_operands(F);
    TSweep S;
    for (int q = 0; q < 4; q++)
        for (int m = 0; m < 5; m++) {
            S.x[q][m] = A.x + (5 * q + m) * T;
            S.mu[q][m] = A.mu + (5 * q + m) * T;
            S.xb[q][m] = A.xb + (5 * q + m) * T;
        }
    for (int m = 0; m < 5; m++) S.rb[m] = A.rb + m * T;
    S.cb = A.cb;
    for (int d = 0; d < 3; d++) S.lv[d] = A.lsrv + d * T;
    for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_ad_tsweep<3>, g, bl, 0, s, b, O, S, P, p);
    hipLaunchKernelGGL(k_ad_tgather<3>, gn, bl, 0, s, b, O, S, P);
    for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_ad_tsweep<2>, g, bl, 0, s, b, O, S, P, p);
    hipLaunchKernelGGL(k_ad_tgather<2>, gn, bl, 0, s, b, O, S, P);
    for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_ad_tsweep<1>, g, bl, 0, s, b, O, S, P, p);
    for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_ad_tsweep<0>, g, bl, 0, s, b, O, S, P, p);
    hipLaunchKernelGGL(k_ad_tcoef, gn, bl, 0, s, b, O, S, P, A.lp, A.lsr, T);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
