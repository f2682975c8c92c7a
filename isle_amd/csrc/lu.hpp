// Batched complex FP64 LU with partial pivoting (one workgroup per matrix), its
// log-determinant epilogue, and a right-hand-side-parallel triangular solve.
//
// Pivoting follows LAPACK zgetrf: the pivot of column k is the first row of maximal
// |re| + |im| (izamax), full rows are swapped, and piv[k] records the swapped row.
// The log-determinant is sum_i log(U_ii) + i*pi*(number of k with piv[k] != k),
// with the imaginary part reduced by the IEEE remainder to [-pi, pi]
// (src/isle/cpp/math.hpp:265-293).
#pragma once

#include "common.hpp"

namespace isle {

struct LuItem {
    double* re;      // n x n, row-major, row stride ld (factorised in place)
    double* im;
    int* piv;        // length n (0-based row index swapped with row k)
    double* logdet;  // 2 doubles (re, im) or null
    int n;
    int ld;
};

constexpr int LU_THREADS = 512;

__device__ __forceinline__ double cabs1(double r, double i) { return fabs(r) + fabs(i); }

// Block-wide argmax (first index on ties).  scratch: 2 * (LU_THREADS/64) entries.
__device__ inline int block_argmax(double v, int idx, double* sval, int* sidx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    if (lane == 0) {
        sval[wave] = v;
        sidx[wave] = idx;
    }
    __syncthreads();
    if (wave == 0) {
        const int nw = blockDim.x >> 6;
        v = lane < nw ? sval[lane] : -1.0;
        idx = lane < nw ? sidx[lane] : 0x7fffffff;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off);
            const int oi = __shfl_xor(idx, off);
            if (ov > v || (ov == v && oi < idx)) {
                v = ov;
                idx = oi;
            }
        }
        if (lane == 0) sidx[0] = idx;
    }
    __syncthreads();
    const int r = sidx[0];
    __syncthreads();
    return r;
}

__device__ inline double block_sum(double v, double* s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) s[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s[w];
        s[0] = t;
    }
    __syncthreads();
    t = s[0];
    __syncthreads();
    return t;
}

// Blocked right-looking LU, one workgroup per matrix.  Per panel of NB columns:
//   1. the (n-k0) x NB panel is factorised with partial pivoting entirely in LDS;
//   2. its row interchanges are applied to all other columns (one column per thread);
//   3. U12 = L11^{-1} A12 by forward substitution (one column per thread), kept in LDS;
//   4. A22 -= L21 U12 with 4x4 complex register tiles fed from LDS.
// Dynamic LDS: panel [n][NB] + U12 [NB][n], complex (split planes).
template <int NB>
__global__ __launch_bounds__(LU_THREADS) void lu_kernel(const LuItem* __restrict__ items, int add_identity) {
    extern __shared__ double lds[];
    __shared__ double sval[LU_THREADS / 64];
    __shared__ int sidx[LU_THREADS / 64];
    __shared__ int perm_pos[2 * NB], perm_src[2 * NB], perm_cnt;
    const LuItem it = items[blockIdx.x];
    const int n = it.n, ld = it.ld, tid = threadIdx.x, nth = blockDim.x;
    double* re = it.re;
    double* im = it.im;
    double* p_re = lds;                     // panel [n][NB]
    double* p_im = lds + (size_t)n * NB;
    double* u_re = lds + 2 * (size_t)n * NB;  // U12 [NB][n]
    double* u_im = lds + 3 * (size_t)n * NB;

    if (add_identity) {
        for (int i = tid; i < n; i += nth) re[(long)i * ld + i] += 1.0;
        __syncthreads();
    }

    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int pr_rows = n - k0;
        // ---- 1. load and factorise the panel in LDS
        for (int e = tid; e < pr_rows * NB; e += nth) {
            const int r = e / NB, c = e % NB;
            double vr = 0, vi = 0;
            if (c < kb) {
                vr = re[(long)(k0 + r) * ld + k0 + c];
                vi = im[(long)(k0 + r) * ld + k0 + c];
            }
            p_re[r * NB + c] = vr;
            p_im[r * NB + c] = vi;
        }
        __syncthreads();
        for (int kk = 0; kk < kb; ++kk) {
            double best = -1.0;
            int bi = 0x7fffffff;
            for (int r = kk + tid; r < pr_rows; r += nth) {
                const double v = cabs1(p_re[r * NB + kk], p_im[r * NB + kk]);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
            const int p = block_argmax(best, bi, sval, sidx);
            if (tid == 0) it.piv[k0 + kk] = k0 + p;
            if (p != kk && tid < NB) {
                const double tr = p_re[kk * NB + tid], ti = p_im[kk * NB + tid];
                p_re[kk * NB + tid] = p_re[p * NB + tid];
                p_im[kk * NB + tid] = p_im[p * NB + tid];
                p_re[p * NB + tid] = tr;
                p_im[p * NB + tid] = ti;
            }
            __syncthreads();
            const double dr = p_re[kk * NB + kk], di = p_im[kk * NB + kk];
            const double den = dr * dr + di * di;
            const double ir = dr / den, ii = -di / den;
            const int rows = pr_rows - (kk + 1);
            const int cols = kb - (kk + 1);
            for (int r = kk + 1 + tid; r < pr_rows; r += nth) {
                const double xr = p_re[r * NB + kk], xi = p_im[r * NB + kk];
                p_re[r * NB + kk] = xr * ir - xi * ii;
                p_im[r * NB + kk] = xr * ii + xi * ir;
            }
            __syncthreads();
            if (cols > 0) {
                for (int e = tid; e < rows * cols; e += nth) {
                    const int r = kk + 1 + e / cols, c = kk + 1 + e % cols;
                    const double lr_ = p_re[r * NB + kk], li = p_im[r * NB + kk];
                    const double ur = p_re[kk * NB + c], ui = p_im[kk * NB + c];
                    p_re[r * NB + c] -= lr_ * ur - li * ui;
                    p_im[r * NB + c] -= lr_ * ui + li * ur;
                }
            }
            __syncthreads();
        }
        for (int e = tid; e < pr_rows * kb; e += nth) {
            const int r = e / kb, c = e % kb;
            re[(long)(k0 + r) * ld + k0 + c] = p_re[r * NB + c];
            im[(long)(k0 + r) * ld + k0 + c] = p_im[r * NB + c];
        }
        // ---- 2. row interchanges on the other columns: the panel's swaps are composed
        //         into one permutation of at most 2*NB rows, applied as a gather/scatter
        if (tid == 0) {
            int cnt = 0;
            for (int kk = 0; kk < kb; ++kk) {
                const int a = k0 + kk, p = it.piv[k0 + kk];
                int ia = -1, ip = -1;
                for (int q = 0; q < cnt; ++q) {
                    if (perm_pos[q] == a) ia = q;
                    if (perm_pos[q] == p) ip = q;
                }
                if (ia < 0) {
                    perm_pos[cnt] = a;
                    perm_src[cnt] = a;
                    ia = cnt++;
                }
                if (ip < 0) {
                    perm_pos[cnt] = p;
                    perm_src[cnt] = p;
                    ip = cnt++;
                }
                const int t = perm_src[ia];
                perm_src[ia] = perm_src[ip];
                perm_src[ip] = t;
            }
            perm_cnt = cnt;
        }
        __syncthreads();
        {
            const int cnt = perm_cnt;
            for (int c = tid; c < n; c += nth) {
                if (c >= k0 && c < k0 + kb) continue;
                double vr[2 * NB], vi[2 * NB];
#pragma unroll
                for (int q = 0; q < 2 * NB; ++q) {
                    if (q < cnt) {
                        vr[q] = re[(long)perm_src[q] * ld + c];
                        vi[q] = im[(long)perm_src[q] * ld + c];
                    }
                }
#pragma unroll
                for (int q = 0; q < 2 * NB; ++q) {
                    if (q < cnt) {
                        re[(long)perm_pos[q] * ld + c] = vr[q];
                        im[(long)perm_pos[q] * ld + c] = vi[q];
                    }
                }
            }
        }
        __syncthreads();
        const int c0 = k0 + kb;
        const int nt = n - c0;
        if (nt <= 0) break;
        // ---- 3. U12 = L11^{-1} A12, one column per thread, into global and LDS
        for (int j = tid; j < nt; j += nth) {
            double xr[NB], xi[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                if (r < kb) {
                    double sr = re[(long)(k0 + r) * ld + c0 + j], si = im[(long)(k0 + r) * ld + c0 + j];
#pragma unroll
                    for (int q = 0; q < NB; ++q) {
                        if (q < r) {
                            const double lr_ = p_re[r * NB + q], li = p_im[r * NB + q];
                            sr -= lr_ * xr[q] - li * xi[q];
                            si -= lr_ * xi[q] + li * xr[q];
                        }
                    }
                    xr[r] = sr;
                    xi[r] = si;
                    re[(long)(k0 + r) * ld + c0 + j] = sr;
                    im[(long)(k0 + r) * ld + c0 + j] = si;
                    u_re[r * n + j] = sr;
                    u_im[r * n + j] = si;
                } else {
                    xr[r] = xi[r] = 0;
                    u_re[r * n + j] = 0;
                    u_im[r * n + j] = 0;
                }
            }
        }
        __syncthreads();
        // ---- 4. A22 -= L21 U12 with 4x4 complex register tiles
        const int tiles_r = (nt + 3) / 4, tiles_c = (nt + 3) / 4;
        for (int tix = tid; tix < tiles_r * tiles_c; tix += nth) {
            const int tr_ = tix / tiles_c, tc = tix % tiles_c;
            double ar[4][4], ai[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) ar[a][b] = ai[a][b] = 0;
            int rows[4], cols[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                rows[a] = min(tr_ + a * tiles_r, nt - 1);  // strided rows for coalescing
                cols[a] = min(tc * 4 + a, nt - 1);
            }
#pragma unroll 2
            for (int q = 0; q < NB; ++q) {
                double lr_[4], li[4], ur[4], ui[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    lr_[a] = p_re[(kb + rows[a]) * NB + q];
                    li[a] = p_im[(kb + rows[a]) * NB + q];
                    ur[a] = u_re[q * n + cols[a]];
                    ui[a] = u_im[q * n + cols[a]];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        ar[a][b] += lr_[a] * ur[b] - li[a] * ui[b];
                        ai[a][b] += lr_[a] * ui[b] + li[a] * ur[b];
                    }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int r = tr_ + a * tiles_r;
                if (r >= nt) continue;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int c = tc * 4 + b;
                    if (c >= nt) continue;
                    const long o = (long)(c0 + r) * ld + c0 + c;
                    re[o] -= ar[a][b];
                    im[o] -= ai[a][b];
                }
            }
        }
        __syncthreads();
    }

    if (it.logdet) {
        double lr_ = 0, la = 0, cnt = 0;
        for (int i = tid; i < n; i += nth) {
            const double dr = re[(long)i * ld + i], di = im[(long)i * ld + i];
            lr_ += log(hypot(dr, di));
            la += atan2(di, dr);
            cnt += (it.piv[i] != i) ? 1.0 : 0.0;
        }
        lr_ = block_sum(lr_, sval);
        la = block_sum(la, sval);
        cnt = block_sum(cnt, sval);
        if (tid == 0) {
            if (((long)cnt) & 1) la += M_PI;
            it.logdet[0] = lr_;
            it.logdet[1] = remainder(la, 2.0 * M_PI);
        }
    }
}

inline void launch_lu(const LuItem* d_items, int nitems, int max_n, int add_identity, hipStream_t s) {
    const size_t lds = sizeof(double) * 4 * (size_t)max_n * 8;
    if (lds > 150 * 1024) throw std::invalid_argument("matrix too large for the LU kernel");
    ISLE_HIP(hipFuncSetAttribute((const void*)lu_kernel<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(lu_kernel<8>, dim3(nitems), dim3(LU_THREADS), lds, s, d_items, add_identity);
    ISLE_HIP(hipGetLastError());
}

// X = A^{-1} B for the LU factors of A.  Element (row r, rhs c) of B is read from
// b_re[c*bsc + r*bsr] / b_im[...] and X is written to x_re/x_im with its own strides
// (so interleaved complex and transposed layouts need no copy).  If b_re is null,
// B is the identity.  16 right-hand sides per workgroup, X staged in LDS.
struct SolveItem {
    const double* lu_re;
    const double* lu_im;
    const int* piv;
    const double* b_re;
    const double* b_im;
    long bsr, bsc;
    double* x_re;
    double* x_im;
    long xsr, xsc;
    int n, ld, nrhs;
};

constexpr int SOLVE_COLS = 16;
constexpr int SOLVE_THREADS = 256;

__global__ __launch_bounds__(SOLVE_THREADS) void lu_solve_kernel(const SolveItem* __restrict__ items) {
    extern __shared__ double lds[];
    const SolveItem it = items[blockIdx.y];
    const int n = it.n, ld = it.ld;
    const int c0 = blockIdx.x * SOLVE_COLS;
    if (c0 >= it.nrhs) return;
    double* xr = lds;                         // [n][16]
    double* xi = lds + (size_t)n * SOLVE_COLS;
    const int tid = threadIdx.x;
    const int c = tid & 15;   // rhs within block
    const int r = tid >> 4;   // row lane within a 16-row block

    for (int e = tid; e < n * SOLVE_COLS; e += SOLVE_THREADS) {
        const int row = e / SOLVE_COLS, cc = e % SOLVE_COLS;
        const int gc = c0 + cc;
        double vr = 0, vi = 0;
        if (gc < it.nrhs) {
            if (it.b_re) {
                vr = it.b_re[gc * it.bsc + row * it.bsr];
                vi = it.b_im[gc * it.bsc + row * it.bsr];
            } else {
                vr = (row == gc) ? 1.0 : 0.0;
            }
        }
        xr[row * SOLVE_COLS + cc] = vr;
        xi[row * SOLVE_COLS + cc] = vi;
    }
    __syncthreads();
    if (tid < SOLVE_COLS) {  // row interchanges, applied in order
        for (int k = 0; k < n; ++k) {
            const int p = it.piv[k];
            if (p != k) {
                const double tr = xr[k * SOLVE_COLS + tid], ti = xi[k * SOLVE_COLS + tid];
                xr[k * SOLVE_COLS + tid] = xr[p * SOLVE_COLS + tid];
                xi[k * SOLVE_COLS + tid] = xi[p * SOLVE_COLS + tid];
                xr[p * SOLVE_COLS + tid] = tr;
                xi[p * SOLVE_COLS + tid] = ti;
            }
        }
    }
    __syncthreads();
    __shared__ double dr_[16][17], di_[16][17];  // diagonal 16x16 block of L or U
    __shared__ double br_[2][SOLVE_COLS], bi_[2][SOLVE_COLS];
    const int nb = (n + 15) / 16;
    // forward substitution with unit lower L, blocks of 16 rows
    for (int b = 0; b < nb; ++b) {
        const int i = b * 16 + r;
        double xr_ = 0, xi_ = 0;
        if (i < n) {
            xr_ = xr[i * SOLVE_COLS + c];
            xi_ = xi[i * SOLVE_COLS + c];
            const double* lr = it.lu_re + (long)i * ld;
            const double* li = it.lu_im + (long)i * ld;
            for (int k = 0; k < b * 16; ++k) {
                const double ar = lr[k], ai = li[k];
                const double br = xr[k * SOLVE_COLS + c], bi = xi[k * SOLVE_COLS + c];
                xr_ -= ar * br - ai * bi;
                xi_ -= ar * bi + ai * br;
            }
            const int kk = b * 16 + c;
            dr_[r][c] = kk < n ? lr[kk] : 0.0;
            di_[r][c] = kk < n ? li[kk] : 0.0;
        }
        __syncthreads();
        for (int rr = 0; rr < 16 && b * 16 + rr < n; ++rr) {
            if (r == rr) {
                br_[rr & 1][c] = xr_;
                bi_[rr & 1][c] = xi_;
            }
            __syncthreads();
            if (r > rr && i < n) {
                const double ar = dr_[r][rr], ai = di_[r][rr];
                const double br = br_[rr & 1][c], bi = bi_[rr & 1][c];
                xr_ -= ar * br - ai * bi;
                xi_ -= ar * bi + ai * br;
            }
        }
        if (i < n) {
            xr[i * SOLVE_COLS + c] = xr_;
            xi[i * SOLVE_COLS + c] = xi_;
        }
        __syncthreads();
    }
    // back substitution with U, blocks of 16 rows from the bottom
    for (int b = nb - 1; b >= 0; --b) {
        const int i = b * 16 + r;
        const int kend = min(n, b * 16 + 16);
        double xr_ = 0, xi_ = 0;
        if (i < n) {
            xr_ = xr[i * SOLVE_COLS + c];
            xi_ = xi[i * SOLVE_COLS + c];
            const double* ur = it.lu_re + (long)i * ld;
            const double* ui = it.lu_im + (long)i * ld;
            for (int k = kend; k < n; ++k) {
                const double ar = ur[k], ai = ui[k];
                const double br = xr[k * SOLVE_COLS + c], bi = xi[k * SOLVE_COLS + c];
                xr_ -= ar * br - ai * bi;
                xi_ -= ar * bi + ai * br;
            }
            const int kk = b * 16 + c;
            dr_[r][c] = kk < n ? ur[kk] : 0.0;
            di_[r][c] = kk < n ? ui[kk] : 0.0;
        }
        __syncthreads();
        for (int rr = kend - 1 - b * 16; rr >= 0; --rr) {
            if (r == rr) {
                const double dr = dr_[rr][rr], di = di_[rr][rr];
                const double den = dr * dr + di * di;
                const double t = (xr_ * dr + xi_ * di) / den;
                xi_ = (xi_ * dr - xr_ * di) / den;
                xr_ = t;
                br_[rr & 1][c] = xr_;
                bi_[rr & 1][c] = xi_;
            }
            __syncthreads();
            if (r < rr) {
                const double ar = dr_[r][rr], ai = di_[r][rr];
                const double br = br_[rr & 1][c], bi = bi_[rr & 1][c];
                xr_ -= ar * br - ai * bi;
                xi_ -= ar * bi + ai * br;
            }
        }
        if (i < n) {
            xr[i * SOLVE_COLS + c] = xr_;
            xi[i * SOLVE_COLS + c] = xi_;
        }
        __syncthreads();
    }
    for (int e = tid; e < n * SOLVE_COLS; e += SOLVE_THREADS) {
        const int row = e / SOLVE_COLS, cc = e % SOLVE_COLS;
        const int gc = c0 + cc;
        if (gc < it.nrhs) {
            it.x_re[gc * it.xsc + row * it.xsr] = xr[row * SOLVE_COLS + cc];
            it.x_im[gc * it.xsc + row * it.xsr] = xi[row * SOLVE_COLS + cc];
        }
    }
}

inline void launch_lu_solve(const SolveItem* d_items, int nitems, int max_n, int max_rhs, hipStream_t s) {
    const size_t lds = sizeof(double) * 2 * (size_t)max_n * SOLVE_COLS;
    if (lds > 160 * 1024) throw std::invalid_argument("matrix too large for the LU solve kernel");
    ISLE_HIP(hipFuncSetAttribute((const void*)lu_solve_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid((max_rhs + SOLVE_COLS - 1) / SOLVE_COLS, nitems);
    hipLaunchKernelGGL(lu_solve_kernel, grid, dim3(SOLVE_THREADS), lds, s, d_items);
    ISLE_HIP(hipGetLastError());
}

}  // namespace isle
