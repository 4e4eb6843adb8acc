// wcns_transpose.hip -- i<->j transposed copies of block fields.
//
// The pencil kernels load coalesced only when adjacent lanes own pencils that
// are adjacent in memory, which holds for j and k pencils (lanes along i) but
// not for i pencils.  The i-direction residuals therefore run as j-direction
// kernels on transposed copies: node (i,j,k) of block b, stored at
// off + (i+2) + (ni+6)(j+2) + sy(k+2), has its copy at
// off + (j+2) + (nj+6)(i+2) + sy(k+2) -- the same block offset and plane size,
// with i and j exchanged.  Their contribution comes back through a transposed
// add into the residual.  Each workgroup moves one 32x32 tile of one k plane
// of one block through LDS, so both the read and the write are coalesced.
// Only what the view's pencils read is moved: every i, at interior j and k.
#include "wcns_kernels.hpp"

namespace wcns {

namespace {

constexpr int TT = 32;      // tile edge
constexpr int TROWS = 8;    // rows per pass: 256 threads = 32 x 8

// BACK = 0: dst (transposed) = src (normal); BACK = 1: dst (normal) += src (transposed).
// TG arrays per pass: every thread issues its loads of the pass before the
// first LDS write, so that each workgroup keeps 4 * TG loads per thread in flight.
constexpr int TG = 4;
template <int BACK>
__global__ void __launch_bounds__(256) k_transpose(BlockGrid G, TransList L)
{
    const int2 wk = G.wg[blockIdx.x];
    const DBlock b = G.blocks[wk.x];
    const int ei = b.ni + 6, ej = b.nj + 6;
    const int ti = (ei + TT - 1) / TT, tj = (b.nj + TT - 1) / TT;
    const int tile = wk.y;
    const int plane = tile / (ti * tj), rem = tile % (ti * tj);
    const int bi = rem % ti, bj = rem / ti;
    // every i (the pencil direction of the view, ghosts included), interior j
    // and k: G3 rows 3..nj+2 of planes 3..nk+2
    const long base = b.off + (long)(plane + 3) * b.sy;
    const int jlo = 3, jhi = b.nj + 3;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    __shared__ double s[TG][TT][TT + 1];
    for (int a0 = 0; a0 < L.n; a0 += TG) {
        double v[TG][TT / TROWS];
        // s[e][jl][il] holds node (bi*TT + il, bj*TT + jl) of array a0 + e
#pragma unroll
        for (int e = 0; e < TG; e++) {
            const int a = a0 + e;
#pragma unroll
            for (int r = 0; r < TT / TROWS; r++) {
                const int yl = ty + TROWS * r;
                v[e][r] = 0.0;
                if (a < L.n) {
                    if (BACK == 0) {
                        const int i = bi * TT + tx, j = jlo + bj * TT + yl;
                        if (i < ei && j < jhi) v[e][r] = L.src[a][base + i + (long)ei * j];
                    } else {
                        const int j = jlo + bj * TT + tx, i = bi * TT + yl;
                        if (i < ei && j < jhi) v[e][r] = L.src[a][base + j + (long)ej * i];
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < TG; e++)
#pragma unroll
            for (int r = 0; r < TT / TROWS; r++) {
                const int yl = ty + TROWS * r;
                if (BACK == 0) s[e][yl][tx] = v[e][r];
                else s[e][tx][yl] = v[e][r];
            }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < TG; e++) {
            const int a = a0 + e;
            if (a >= L.n) break;
            double *dst = L.dst[a];
#pragma unroll
            for (int r = 0; r < TT / TROWS; r++) {
                const int yl = ty + TROWS * r;
                if (BACK == 0) {
                    const int j = jlo + bj * TT + tx, i = bi * TT + yl;
                    if (i < ei && j < jhi) dst[base + j + (long)ej * i] = s[e][tx][yl];
                } else {
                    const int i = bi * TT + tx, j = jlo + bj * TT + yl;
                    if (i < ei && j < jhi) {
                        const long g = base + i + (long)ei * j;
                        dst[g] = dst[g] + s[e][yl][tx];
                    }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

long transpose_tiles(const DBlock &b)
{
    const long ti = (b.ni + 6 + TT - 1) / TT, tj = (b.nj + TT - 1) / TT;
    return ti * tj * b.nk;
}

DBlock transposed_view(const DBlock &b)
{
    DBlock t = b;
    t.ni = b.nj;
    t.nj = b.ni;
    t.sx = b.nj + 6;
    return t;
}

int launch_transpose(int back, const BlockGrid &G, const TransList &L, hipStream_t s)
{
    if (G.nwg == 0 || L.n == 0) return 0;
    if (L.n > TRANS_MAX) return -2;
    if (back) hipLaunchKernelGGL(k_transpose<1>, dim3(G.nwg), dim3(256), 0, s, G, L);
    else hipLaunchKernelGGL(k_transpose<0>, dim3(G.nwg), dim3(256), 0, s, G, L);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace wcns
