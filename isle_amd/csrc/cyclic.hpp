// Helpers of the structured elimination of the block-cyclic matrix
//
//        | I   -P_0                     |
//        |      I   -P_1                |
//   R =  |           .     .            |      det R = det(1 + P_0 P_1 ... P_{n-1})
//        |                 I  -P_{n-2}  |
//        | P_{n-1}              I       |
//
// (one block per group of time slices, P_j the product of the group's slice matrices).
// Step k of the elimination factorises the 2nx x 3nx matrix W_k = [[I, -P_k, 0], [B_k, 0, C_k]]
// (top: block row k, columns k, k+1, n-1; bottom: what is left of the last block row) over its
// first nx columns with partial pivoting; its Schur complement gives B_{k+1}, C_{k+1}.  The
// last step factorises [[I, -P_{n-2}], [B, C]] completely.  Everything here is bandwidth
// work: assembling the blocks of W_k, permuting stacked right-hand sides, and extracting
// (transposed, signed) blocks of the solution.
#pragma once

#include "common.hpp"

namespace isle {

// dst (rows x cols, row stride ldd, split re/im planes given separately) =
//   mode 0: 0,  1: identity,  2: src,  3: -src  (src row stride lds)
struct BlockItem {
    double* d_re;
    double* d_im;
    const double* s_re;
    const double* s_im;
    long ldd, lds;
    int rows, cols, mode;
};

__global__ __launch_bounds__(256) void block_fill_kernel(const BlockItem* __restrict__ items) {
    const BlockItem it = items[blockIdx.y];
    const long total = (long)it.rows * it.cols;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int r = (int)(e / it.cols), c = (int)(e % it.cols);
        double vr = 0, vi = 0;
        if (it.mode == 1) {
            vr = r == c ? 1.0 : 0.0;
        } else if (it.mode >= 2) {
            vr = it.s_re[r * it.lds + c];
            vi = it.s_im[r * it.lds + c];
            if (it.mode == 3) {
                vr = -vr;
                vi = -vi;
            }
        }
        it.d_re[r * it.ldd + c] = vr;
        it.d_im[r * it.ldd + c] = vi;
    }
}

// Row permutation of a stacked right-hand side of m = 2 nx rows after the row interchanges
// piv[0..npiv) (LAPACK order): sigma[i] = the row that ends at position i.
struct SigmaItem {
    const int* piv;
    int* sigma;
    int m, npiv;
};

__global__ __launch_bounds__(256) void sigma_kernel(const SigmaItem* __restrict__ items) {
    constexpr int PV_LDS = 1024;  // pivot lists up to this length are staged in LDS
    __shared__ int pv_lds[PV_LDS];
    const SigmaItem it = items[blockIdx.x];
    const bool staged = it.npiv <= PV_LDS;  // uniform over the workgroup
    if (staged) {
        for (int i = threadIdx.x; i < it.npiv; i += 256) pv_lds[i] = it.piv[i];
        __syncthreads();
    }
    const int* pv = staged ? pv_lds : it.piv;
    for (int r = threadIdx.x; r < it.m; r += 256) {
        int pos = r;
        for (int k = 0; k < it.npiv; ++k) {
            const int p = pv[k];
            pos = pos == k ? p : (pos == p ? k : pos);
        }
        it.sigma[pos] = r;
    }
}

// Permuted split of the stacked right-hand side [top; bot] (nx rows each, `cols` columns):
// row i of the permuted stack (sigma) goes to t (i < nx) or b (i >= nx).  The top block is
// the identity placed at columns [top_col, top_col + nx) (zero elsewhere); bot has row
// stride ldb, t and b row stride ldo.  b must not alias bot.
struct GatherItem {
    const int* sigma;
    const double* bot_re;
    const double* bot_im;
    double* t_re;
    double* t_im;
    double* b_re;
    double* b_im;
    long ldb, ldo;
    int nx, cols, top_col;
};

__global__ __launch_bounds__(256) void stack_gather_kernel(const GatherItem* __restrict__ items) {
    const GatherItem it = items[blockIdx.z];
    const int i = blockIdx.y;  // permuted row, 0 .. 2 nx - 1
    if (i >= 2 * it.nx) return;
    const int src = it.sigma[i];
    double* o_re = i < it.nx ? it.t_re + (long)i * it.ldo : it.b_re + (long)(i - it.nx) * it.ldo;
    double* o_im = i < it.nx ? it.t_im + (long)i * it.ldo : it.b_im + (long)(i - it.nx) * it.ldo;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < it.cols; c += gridDim.x * 256) {
        double vr = 0, vi = 0;
        if (src < it.nx) {
            vr = (c - it.top_col == src) ? 1.0 : 0.0;
        } else {
            vr = it.bot_re[(long)(src - it.nx) * it.ldb + c];
            vi = it.bot_im[(long)(src - it.nx) * it.ldb + c];
        }
        o_re[c] = vr;
        o_im[c] = vi;
    }
}

// d[i, :] = s[sigma[i], :] over the n rows of a split-plane block of `cols` columns (row
// strides lds and ldd): the row interchanges of an LU applied to its right-hand sides.
struct RowGatherItem {
    const int* sigma;
    const double* s_re;
    const double* s_im;
    double* d_re;
    double* d_im;
    long lds, ldd;
    int n, cols;
};

__global__ __launch_bounds__(256) void row_gather_kernel(const RowGatherItem* __restrict__ items) {
    const RowGatherItem it = items[blockIdx.z];
    const int i = blockIdx.y;
    if (i >= it.n) return;
    const long src = (long)it.sigma[i] * it.lds, dst = (long)i * it.ldd;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < it.cols; c += gridDim.x * 256) {
        it.d_re[dst + c] = it.s_re[src + c];
        it.d_im[dst + c] = it.s_im[src + c];
    }
}

// dst (n x n planes, re then im) = sign * src^T, src an n x n block with row stride lds.
struct TBlockItem {
    const double* s_re;
    const double* s_im;
    double* dst;
    long lds;
    int n;
    double sign;
};

__global__ __launch_bounds__(256) void tblock_kernel(const TBlockItem* __restrict__ items) {
    __shared__ double tile[32][33];
    const TBlockItem it = items[blockIdx.z];
    const int n = it.n;
    const int tiles = (n + 31) / 32;
    const int plane = blockIdx.y;
    const double* src = plane == 0 ? it.s_re : it.s_im;
    double* dst = it.dst + (long)plane * n * n;
    for (int tb = blockIdx.x; tb < tiles * tiles; tb += gridDim.x) {
        const int bi = tb / tiles, bj = tb % tiles;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
        for (int r = ty; r < 32; r += 8) {
            const int i = bi * 32 + r, j = bj * 32 + tx;
            tile[r][tx] = (i < n && j < n) ? src[(long)i * it.lds + j] : 0.0;
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int i = bj * 32 + r, j = bi * 32 + tx;  // dst[i][j] = src[j][i]
            if (i < n && j < n) dst[(long)i * n + j] = it.sign * tile[tx][r];
        }
        __syncthreads();
    }
}

}  // namespace isle
