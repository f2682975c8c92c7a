// Batched split-complex GEMM on CDNA4 MFMA (16x16x4, f64 or f32), LDS-tiled.
//
// Every time-slice product of the fermion-action chains is one of
//     C = diag(post) * A * diag(pre) * B      (complex operand pre-scaled along k)
//     C = A * diag(pre) * B * diag(post)
// with one operand the real (exp(-kappa) / K) matrix and the other a complex matrix
// stored as two real planes.  The diagonal phase factors exp(+-i phi_t) are applied
// while the complex tile is staged into LDS (pre) or in the epilogue (post), so a
// slice costs exactly two real GEMMs and no extra pass over HBM.
//
// Tiling: 256 threads (4 waves in 2x2), output tile BM x BN per plane, K staged in
// slabs of BK through double-buffered LDS; the next slab is prefetched into registers
// while the current one feeds the MFMAs, so a K-deep product costs K/BK global round
// trips.  Each group of 16 k is summed in a fresh accumulator before being added to the
// running sum (blocked summation).
#pragma once

#include "common.hpp"

namespace isle {

enum : int { SCALE_NONE = 0, SCALE_ROW = 1, SCALE_COL = 2, SUBTRACT = 3 };  // POST: C = C - A B

template <typename T>
struct GemmItem {
    const T* a_re;  // A: m x k, row stride lda
    const T* a_im;  // null when A is real
    const T* b_re;  // B: k x n, row stride ldb
    const T* b_im;  // null when B is real
    T* c_re;        // C: m x n, row stride ldc
    T* c_im;
    const T* pre_re;   // length k: complex factor applied to the complex operand along k
    const T* pre_im;
    const T* post_re;  // length m (row) or n (col)
    const T* post_im;
    int m, n, k;
    int lda, ldb, ldc;
};

constexpr int GEMM_THREADS = 256;

// One BM x BN output tile of item `it` (tile index `tile` in a ceil(m/BM) x tiles_n_max
// grid).  AC/BC: A/B complex.  PRE: scale the (single) complex operand along k.
// POST: SCALE_* or SUBTRACT.  Must be called by all 256 threads of the block.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int GEMM_BK>
__device__ __forceinline__ void gemm_tile(const GemmItem<T>& it, int tile, int tiles_n_max) {
    static_assert(!(PRE && AC && BC), "pre-scaling needs exactly one complex operand");
    static_assert(BM % 32 == 0 && BN % 32 == 0, "tile must be a multiple of 32");
    static_assert(GEMM_BK % 16 == 0, "K slab must be a multiple of 16");
    constexpr int MI = BM / 32;  // 16x16 MFMA blocks per wave along m
    constexpr int NJ = BN / 32;  // along n
    constexpr int AE = BM * GEMM_BK / GEMM_THREADS;  // A elements per thread per slab
    constexpr int BE = BN * GEMM_BK / GEMM_THREADS;  // B elements per thread per slab
    constexpr bool CPLX = AC || BC;

    const int tile_m = tile / tiles_n_max;
    const int tile_n = tile % tiles_n_max;
    const int m0 = tile_m * BM;
    const int n0 = tile_n * BN;
    if (m0 >= it.m || n0 >= it.n) return;

    __shared__ T As_re[2][GEMM_BK][BM];
    __shared__ T As_im[2][AC ? GEMM_BK : 1][AC ? BM : 1];
    __shared__ T Bs_re[2][GEMM_BK][BN];
    __shared__ T Bs_im[2][BC ? GEMM_BK : 1][BC ? BN : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave >> 1) * (BM / 2);
    const int wn = (wave & 1) * (BN / 2);
    const int lr = lane & 15;
    const int lk = lane >> 4;

    typedef typename Acc4<T>::type V4;
    V4 acc_re[MI][NJ], acc_im[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            acc_re[i][j] = V4{0, 0, 0, 0};
            acc_im[i][j] = V4{0, 0, 0, 0};
        }

    // A slab: BM rows x 16 k; thread covers row la_m, k = la_k .. la_k+AE-1
    constexpr int A_TPR = GEMM_BK / AE;  // threads per row
    const int la_m = tid / A_TPR;
    const int la_k = (tid % A_TPR) * AE;
    // B slab: 16 k x BN cols; thread covers k row lb_k, cols lb_n .. lb_n+BE-1
    constexpr int B_TPR = BN / BE;
    const int lb_k = tid / B_TPR;
    const int lb_n = (tid % B_TPR) * BE;

    T ra_re[AE], ra_im[AC ? AE : 1], rb_re[BE], rb_im[BC ? BE : 1];

    auto load_slab = [&](int k0) {
        const int gm = m0 + la_m;
#pragma unroll
        for (int q = 0; q < AE; ++q) {
            const int gk = k0 + la_k + q;
            T vr = 0, vi = 0;
            if (gm < it.m && gk < it.k) {
                const long off = (long)gm * it.lda + gk;
                vr = it.a_re[off];
                if (AC) vi = it.a_im[off];
                if (PRE && AC) {
                    const T pr = it.pre_re[gk], pi = it.pre_im[gk];
                    const T r = vr * pr - vi * pi;
                    vi = vr * pi + vi * pr;
                    vr = r;
                }
            }
            ra_re[q] = vr;
            if (AC) ra_im[q] = vi;
        }
        const int gk = k0 + lb_k;
        T pr = 1, pi = 0;
        if (PRE && BC && gk < it.k) {
            pr = it.pre_re[gk];
            pi = it.pre_im[gk];
        }
#pragma unroll
        for (int q = 0; q < BE; ++q) {
            const int gn = n0 + lb_n + q;
            T vr = 0, vi = 0;
            if (gk < it.k && gn < it.n) {
                const long off = (long)gk * it.ldb + gn;
                vr = it.b_re[off];
                if (BC) vi = it.b_im[off];
                if (PRE && BC) {
                    const T r = vr * pr - vi * pi;
                    vi = vr * pi + vi * pr;
                    vr = r;
                }
            }
            rb_re[q] = vr;
            if (BC) rb_im[q] = vi;
        }
    };
    auto store_slab = [&](int b) {
#pragma unroll
        for (int q = 0; q < AE; ++q) {
            As_re[b][la_k + q][la_m] = ra_re[q];
            if (AC) As_im[b][la_k + q][la_m] = ra_im[q];
        }
#pragma unroll
        for (int q = 0; q < BE; ++q) {
            Bs_re[b][lb_k][lb_n + q] = rb_re[q];
            if (BC) Bs_im[b][lb_k][lb_n + q] = rb_im[q];
        }
    };

    load_slab(0);
    store_slab(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < it.k; k0 += GEMM_BK) {
        const bool more = k0 + GEMM_BK < it.k;
        if (more) load_slab(k0 + GEMM_BK);

#pragma unroll
        for (int kg = 0; kg < GEMM_BK; kg += 16) {
            V4 part_re[MI][NJ], part_im[MI][NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    part_re[i][j] = V4{0, 0, 0, 0};
                    part_im[i][j] = V4{0, 0, 0, 0};
                }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const int kk = kg + ks * 4 + lk;
                T ar[MI], ai[MI], br[NJ], bi[NJ];
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    ar[i] = As_re[buf][kk][wm + i * 16 + lr];
                    ai[i] = AC ? As_im[buf][kk][wm + i * 16 + lr] : T(0);
                }
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    br[j] = Bs_re[buf][kk][wn + j * 16 + lr];
                    bi[j] = BC ? Bs_im[buf][kk][wn + j * 16 + lr] : T(0);
                }
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        if (AC && BC) {
                            part_re[i][j] = mfma16(ar[i], br[j], part_re[i][j]);
                            part_re[i][j] = mfma16(-ai[i], bi[j], part_re[i][j]);
                            part_im[i][j] = mfma16(ar[i], bi[j], part_im[i][j]);
                            part_im[i][j] = mfma16(ai[i], br[j], part_im[i][j]);
                        } else if (AC) {
                            part_re[i][j] = mfma16(ar[i], br[j], part_re[i][j]);
                            part_im[i][j] = mfma16(ai[i], br[j], part_im[i][j]);
                        } else if (BC) {
                            part_re[i][j] = mfma16(ar[i], br[j], part_re[i][j]);
                            part_im[i][j] = mfma16(ar[i], bi[j], part_im[i][j]);
                        } else {
                            part_re[i][j] = mfma16(ar[i], br[j], part_re[i][j]);
                        }
                    }
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    acc_re[i][j] += part_re[i][j];
                    if (CPLX) acc_im[i][j] += part_im[i][j];
                }
        }
        if (more) store_slab(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

    // ---- epilogue
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + acc_row<T>(lk, r);
                const int col = n0 + wn + j * 16 + lr;
                if (row < it.m && col < it.n) {
                    T vr = acc_re[i][j][r];
                    T vi = CPLX ? acc_im[i][j][r] : T(0);
                    const long off = (long)row * it.ldc + col;
                    if (POST == SUBTRACT) {
                        it.c_re[off] -= vr;
                        if (CPLX) it.c_im[off] -= vi;
                        continue;
                    }
                    if (POST != SCALE_NONE) {
                        const int di = POST == SCALE_ROW ? row : col;
                        const T pr = it.post_re[di], pi = it.post_im[di];
                        const T t = vr * pr - vi * pi;
                        vi = vr * pi + vi * pr;
                        vr = t;
                    }
                    it.c_re[off] = vr;
                    if (CPLX || POST != SCALE_NONE) it.c_im[off] = vi;
                }
            }
}

// grid: x = tile index over (ceil(m/BM) x ceil(n/BN)) of the largest item, y = item.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int GEMM_BK>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(const GemmItem<T>* __restrict__ items,
                                                            int tiles_n_max) {
    const GemmItem<T> it = items[blockIdx.y];
    gemm_tile<T, AC, BC, PRE, POST, BM, BN, GEMM_BK>(it, blockIdx.x, tiles_n_max);
}

// Tile shapes the launcher can pick (the plan's autotuner times them per step):
//   0: 64x64, K slab 16   1: 32x32, K slab 32   2: 32x32, K slab 64   3: 64x64, K slab 32
enum : int { N_TILES = 4 };

template <typename T, bool AC, bool BC, int BM, int BN, int BK>
constexpr size_t gemm_lds_bytes() {
    return sizeof(T) * 2 * BK * (BM * (AC ? 2 : 1) + BN * (BC ? 2 : 1));
}

// Launches the GEMM with tile shape `tile`; returns false if that shape does not fit
// this operand type (LDS budget), in which case nothing is launched.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int BK>
inline bool launch_gemm_tile(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s) {
    if constexpr (gemm_lds_bytes<T, AC, BC, BM, BN, BK>() > 96 * 1024) {
        return false;
    } else {
        const int tm = (max_m + BM - 1) / BM, tn = (max_n + BN - 1) / BN;
        hipLaunchKernelGGL((gemm_kernel<T, AC, BC, PRE, POST, BM, BN, BK>), dim3(tm * tn, nitems),
                           dim3(GEMM_THREADS), 0, s, d_items, tn);
        ISLE_HIP(hipGetLastError());
        return true;
    }
}

template <typename T, bool AC, bool BC, bool PRE, int POST>
inline bool launch_gemm(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s,
                        int tile = 0) {
    switch (tile) {
    case 1: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 32>(d_items, nitems, max_m, max_n, s);
    case 2: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 64>(d_items, nitems, max_m, max_n, s);
    case 3: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 32>(d_items, nitems, max_m, max_n, s);
    default: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 16>(d_items, nitems, max_m, max_n, s);
    }
}

}  // namespace isle
