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
//                  one step ahead, off the dependency chain.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

inline dim3 grid1(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// Sub1.f90:676-717 getrkec_mml clamp + LHS.f90:647-678 getrarbrc clamp
__device__ inline int clampi(int v, int n) { return v < 1 ? 1 : (v > n ? n : v); }

// Coefficients of the split Jacobian of neighbour (i,j,k)+off*e_dir: the
// neighbour's primitives, the metric at the neighbour with ghost indices
// clamped to the face, and the spectral radius clamped to 1..n.
template <typename T>
__device__ inline void nbr_coef(const DBlock &b, const DFields &F, const DParams &P, int i, int j, int k, int dir,
                                int off, MxCoef<T> &c)
{
    int in = i + (dir == 0 ? off : 0), jn = j + (dir == 1 ? off : 0), kn = k + (dir == 2 ? off : 0);
    long gp = gidx(b, in, jn, kn);
    int im = in, jm = jn, km = kn;
    if (im == 0) im = 1;
    if (im == b.ni + 1) im = b.ni;
    if (jm == 0) jm = 1;
    if (jm == b.nj + 1) jm = b.nj;
    if (km == 0) km = 1;
    if (km == b.nk + 1) km = b.nk;
    long gm = gidx(b, im, jm, km);
    long gr;
    if (dir == 0) gr = gidx(b, clampi(i + off, b.ni), j, k);
    else if (dir == 1) gr = gidx(b, i, clampi(j + off, b.nj), k);
    else gr = gidx(b, i, j, clampi(k + off, b.nk));
    T prim[5] = {(T)F.r[gp], (T)F.u[gp], (T)F.v[gp], (T)F.w[gp], (T)F.p[gp]};
    const double *mx = dir == 0 ? F.kcx : (dir == 1 ? F.etx : F.ctx);
    const double *my = dir == 0 ? F.kcy : (dir == 1 ? F.ety : F.cty);
    const double *mz = dir == 0 ? F.kcz : (dir == 1 ? F.etz : F.ctz);
    T gy[3] = {(T)mx[gm], (T)my[gm], (T)mz[gm]};
    const double *sr = dir == 0 ? F.sra : (dir == 1 ? F.srb : F.src);
    mxdq_coef<T>(prim, gy, (T)sr[gr], off < 0 ? T(1.0) : T(-1.0), (T)P.gama, c);
}

// Operands of one node that do not depend on this sweep's new dQ.
template <int KIND, typename T>
struct NodeIn {
    MxCoef<T> cn[3];      // neighbours updated in this sweep, directions i, j, k
    T fold[3][5];         // SGS: complete terms of the three previous-sweep neighbours
    T ghost[3][5];        // LU: dQ of a this-sweep neighbour outside the block
    T own[5];             // lusgs_u: the node's dQ from lusgs_l
    T rhs[5];
    T coed;
    T rv[3];              // LU, viscous: neighbours' viscous spectral radii
};

template <int KIND>
constexpr int sweep_off() { return (KIND == 0 || KIND == 2) ? -1 : 1; }

// A: the dQ array the sweep pair works in; B: the SGS second buffer (A for the
// in-place schedule).  gs_pr_l reads its previous-sweep neighbours from A,
// gs_pr_u from B.
template <int KIND, typename T>
__device__ inline void node_inputs(const DBlock &b, const DFields &F, const DParams &P, const T *A, const T *B,
                                   int i, int j, int k, NodeIn<KIND, T> &in)
{
    constexpr int off = sweep_off<KIND>();
    const long T5 = F.total;
    const long g = gidx(b, i, j, k);
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) nbr_coef<T>(b, F, P, i, j, k, d, off, in.cn[d]);
#pragma unroll
    for (int m = 0; m < 5; m++) in.rhs[m] = (T)F.rhs[m * T5 + g];
    in.coed = (T)F.coed[g];
    if (KIND < 2) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int c = ijk[d] + off;
            if (c < 1 || c > nn[d]) {
                long gn = gidx(b, i + (d == 0 ? off : 0), j + (d == 1 ? off : 0), k + (d == 2 ? off : 0));
#pragma unroll
                for (int m = 0; m < 5; m++) in.ghost[d][m] = A[m * T5 + gn];
            }
        }
        if (KIND == 1) {
#pragma unroll
            for (int m = 0; m < 5; m++) in.own[m] = A[m * T5 + g];
        }
        if (P.nvis > 0) {
            in.rv[0] = (T)F.srva[gidx(b, i + off, j, k)];
            in.rv[1] = (T)F.srvb[gidx(b, i, j + off, k)];
            in.rv[2] = (T)F.srvc[gidx(b, i, j, k + off)];
        }
    } else {
        const T *old = KIND == 2 ? A : B;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int c = ijk[d] - off;
            T dqo[5];
            long gn = gidx(b, i - (d == 0 ? off : 0), j - (d == 1 ? off : 0), k - (d == 2 ? off : 0));
#pragma unroll
            for (int m = 0; m < 5; m++) dqo[m] = (c < 1 || c > nn[d]) ? T(0.0) : old[m * T5 + gn];
            MxCoef<T> co;
            nbr_coef<T>(b, F, P, i, j, k, d, -off, co);
            mxdq_apply<T>(co, dqo, in.fold[d]);
        }
    }
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

template <int KIND>
__global__ void __launch_bounds__(256) k_sweep_plane(DBlock b, DFields F, DParams P, int s)
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
    double *A = F.dq;
    const long T5 = F.total;
    NodeIn<KIND, double> in;
    node_inputs<KIND, double>(b, F, P, A, A, i, j, k, in);
    const int nn[3] = {b.ni, b.nj, b.nk};
    const int ijk[3] = {i, j, k};
    double nw[3][5];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int c = ijk[d] + off;
        if (c < 1 || c > nn[d]) {
            outside<KIND, double>(in, d, nw[d]);
        } else {
            long gn = gidx(b, i + (d == 0 ? off : 0), j + (d == 1 ? off : 0), k + (d == 2 ? off : 0));
#pragma unroll
            for (int m = 0; m < 5; m++) nw[d][m] = A[m * T5 + gn];
        }
    }
    double out[5];
    node_compute<KIND, double>(P, in, nw, out);
    long g = gidx(b, i, j, k);
#pragma unroll
    for (int m = 0; m < 5; m++) A[m * T5 + g] = out[m];
}

// Empty exchange slot: a signalling NaN; arithmetic only ever yields quiet NaNs.
constexpr unsigned long long XCH_EMPTY = 0x7ff4deadbeef0001ull;

__global__ void k_fill_u64(unsigned long long *a, long n, unsigned long long v)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

template <typename T>
__device__ inline void xch_put(double *slot, const T *v)
{
    unsigned long long *p = reinterpret_cast<unsigned long long *>(slot);
#pragma unroll
    for (int m = 0; m < 5; m++)
        __hip_atomic_store(p + m, (unsigned long long)__double_as_longlong((double)v[m]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void xch_issue(double *slot, unsigned long long *raw)
{
    unsigned long long *p = reinterpret_cast<unsigned long long *>(slot);
#pragma unroll
    for (int m = 0; m < 5; m++) raw[m] = __hip_atomic_load(p + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Complete a slot read issued earlier: re-read the words still empty.  The
// wait is bounded: after spin_limit re-reads, or once any lane of the launch
// has given up (err set), it stops and reports failure.
template <typename T>
__device__ inline bool xch_finish(double *slot, unsigned long long *raw, T *v, long spin_limit, int *err)
{
    unsigned long long *p = reinterpret_cast<unsigned long long *>(slot);
    long spins = 0;
    bool ok = true;
#pragma unroll
    for (int m = 0; m < 5; m++) {
        unsigned long long x = raw[m];
        while (x == XCH_EMPTY && ok) {
            if (++spins > spin_limit ||
                ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            x = __hip_atomic_load(p + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        v[m] = (T)__longlong_as_double((long long)x);
    }
    return ok;
}

constexpr int PIPE_MAX_LANES = 256;

template <int KIND, typename T>
__global__ void __launch_bounds__(PIPE_MAX_LANES) k_sweep_pipe(DBlock b, DFields F, DParams P, SweepTiles S,
                                                               T *A, T *B)
{
    constexpr bool fwd = (KIND == 0 || KIND == 2);
    constexpr int off = sweep_off<KIND>();
    const int TJ = S.TJ, TK = S.TK, ni = b.ni;
    const long T5 = F.total;
    const int tile = (fwd ? S.order_fwd : S.order_bwd)[blockIdx.x];
    const int J = tile % S.nTJ, K = tile / S.nTJ;
    const int lane = threadIdx.x, jl = lane % TJ, kl = lane / TJ;
    const int jlo = 1 + J * TJ, klo = 1 + K * TK;
    const int jn = min(TJ, b.nj - jlo + 1), kn = min(TK, b.nk - klo + 1);
    const bool col = kl < kn && jl < jn;
    const int j = jlo + jl, k = klo + kl;
    const int nsteps = ni + jn + kn - 2;
    // source of the this-sweep j and k neighbours: 0 this tile, 1 adjacent tile, 2 outside the block
    const int jsrc = fwd ? (jl > 0 ? 0 : (J > 0 ? 1 : 2)) : (jl < jn - 1 ? 0 : (J < S.nTJ - 1 ? 1 : 2));
    const int ksrc = fwd ? (kl > 0 ? 0 : (K > 0 ? 1 : 2)) : (kl < kn - 1 ? 0 : (K < S.nTK - 1 ? 1 : 2));
    // this lane's results are the adjacent tile's neighbours
    const bool putj = fwd ? (jl == jn - 1 && J < S.nTJ - 1) : (jl == 0 && J > 0);
    const bool putk = fwd ? (kl == kn - 1 && K < S.nTK - 1) : (kl == 0 && K > 0);
    const long rowlen = (long)ni * 5;
    double *xj_in = S.xj + (jsrc == 1 ? ((long)(tile + off) * TK + kl) * rowlen : 0);
    double *xk_in = S.xk + (ksrc == 1 ? ((long)(tile + off * S.nTJ) * TJ + jl) * rowlen : 0);
    double *xj_out = S.xj + ((long)tile * TK + kl) * rowlen;
    double *xk_out = S.xk + ((long)tile * TJ + jl) * rowlen;
    T *W = KIND == 2 ? B : A;

    __shared__ T sh[2][5][PIPE_MAX_LANES];
    T prev[5] = {T(0), T(0), T(0), T(0), T(0)};
    NodeIn<KIND, T> in;
    unsigned long long rj[5], rk[5];
    bool bad = false;                 // a bounded wait of this lane expired: stop waiting
    const int i0 = fwd ? 1 - jl - kl : ni + (jn - 1 - jl) + (kn - 1 - kl);
    int i = i0;
    bool act = col && i >= 1 && i <= ni;
    if (act) {
        node_inputs<KIND, T>(b, F, P, A, B, i, j, k, in);
        if (jsrc == 1) xch_issue(xj_in + (long)(i - 1) * 5, rj);
        if (ksrc == 1) xch_issue(xk_in + (long)(i - 1) * 5, rk);
    }
    for (int step = 0; step < nsteps; step++) {
        const int buf = step & 1;
        if (act) {
            T nw[3][5];
            if (i + off >= 1 && i + off <= ni) {
#pragma unroll
                for (int m = 0; m < 5; m++) nw[0][m] = prev[m];
            } else {
                outside<KIND, T>(in, 0, nw[0]);
            }
            if (jsrc == 0) {
#pragma unroll
                for (int m = 0; m < 5; m++) nw[1][m] = sh[buf ^ 1][m][lane + off];
            } else if (jsrc == 1) {
                if (!xch_finish<T>(xj_in + (long)(i - 1) * 5, rj, nw[1], bad ? 0 : S.spin_limit, S.err)) {
                    if (!bad) atomicExch(S.err, 1);
                    bad = true;
                }
            } else {
                outside<KIND, T>(in, 1, nw[1]);
            }
            if (ksrc == 0) {
#pragma unroll
                for (int m = 0; m < 5; m++) nw[2][m] = sh[buf ^ 1][m][lane + off * TJ];
            } else if (ksrc == 1) {
                if (!xch_finish<T>(xk_in + (long)(i - 1) * 5, rk, nw[2], bad ? 0 : S.spin_limit, S.err)) {
                    if (!bad) atomicExch(S.err, 1);
                    bad = true;
                }
            } else {
                outside<KIND, T>(in, 2, nw[2]);
            }
            T out[5];
            node_compute<KIND, T>(P, in, nw, out);
            const long g = gidx(b, i, j, k);
            if (putj) xch_put<T>(xj_out + (long)(i - 1) * 5, out);
            if (putk) xch_put<T>(xk_out + (long)(i - 1) * 5, out);
#pragma unroll
            for (int m = 0; m < 5; m++) {
                sh[buf][m][lane] = out[m];
                prev[m] = out[m];
                W[m * T5 + g] = out[m];
            }
        }
        // the next step's operands, formed before the barrier
        i = fwd ? i0 + step + 1 : i0 - step - 1;
        act = col && i >= 1 && i <= ni && step + 1 < nsteps;
        if (act) {
            node_inputs<KIND, T>(b, F, P, A, B, i, j, k, in);
            if (jsrc == 1) xch_issue(xj_in + (long)(i - 1) * 5, rj);
            if (ksrc == 1) xch_issue(xk_in + (long)(i - 1) * 5, rk);
        }
        __syncthreads();
    }
}

}  // namespace

int launch_sweep(int kind, const DBlock &b, const DFields &F, const DParams &P, hipStream_t s)
{
    int smin = 3, smax = b.ni + b.nj + b.nk;
    long nthr = (long)b.nj * b.nk;
    dim3 g = grid1(nthr, 256), bl(256);
    switch (kind) {
    case 0: for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_sweep_plane<0>, g, bl, 0, s, b, F, P, p); break;
    case 1: for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_sweep_plane<1>, g, bl, 0, s, b, F, P, p); break;
    case 2: for (int p = smin; p <= smax; p++) hipLaunchKernelGGL(k_sweep_plane<2>, g, bl, 0, s, b, F, P, p); break;
    default: for (int p = smax; p >= smin; p--) hipLaunchKernelGGL(k_sweep_plane<3>, g, bl, 0, s, b, F, P, p); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

long sweep_exchange_words(const SweepTiles &T, int ni)
{
    return (long)T.nTJ * T.nTK * (T.TJ + T.TK) * ni * 5;
}

int launch_sweep_tiled(int kind, const DBlock &b, const DFields &F, const DParams &P, const SweepTiles &T,
                       int prec, void *A, void *B, hipStream_t s)
{
    int ntiles = T.nTJ * T.nTK;
    if (T.TJ * T.TK > PIPE_MAX_LANES || T.TJ < 1 || T.TK < 1) return -2;
    long nx = sweep_exchange_words(T, b.ni);
    hipLaunchKernelGGL(k_fill_u64, grid1(nx, 256), dim3(256), 0, s, (unsigned long long *)T.xj, nx, XCH_EMPTY);
    dim3 g(ntiles), bl(T.TJ * T.TK);
#define WCNS_PIPE(K, TY) hipLaunchKernelGGL((k_sweep_pipe<K, TY>), g, bl, 0, s, b, F, P, T, (TY *)A, (TY *)B)
    if (prec == 1) {
        switch (kind) {
        case 0: WCNS_PIPE(0, float); break;
        case 1: WCNS_PIPE(1, float); break;
        case 2: WCNS_PIPE(2, float); break;
        default: WCNS_PIPE(3, float); break;
        }
    } else {
        switch (kind) {
        case 0: WCNS_PIPE(0, double); break;
        case 1: WCNS_PIPE(1, double); break;
        case 2: WCNS_PIPE(2, double); break;
        default: WCNS_PIPE(3, double); break;
        }
    }
#undef WCNS_PIPE
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
