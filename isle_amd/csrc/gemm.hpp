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

#include <type_traits>

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
    static_assert(AE * GEMM_THREADS == BM * GEMM_BK && BE * GEMM_THREADS == BN * GEMM_BK, "slab/threads");
    constexpr bool CPLX = AC || BC;
    // FP32 sums each group of 16 k in a fresh accumulator (blocked summation); FP64 does not need it
    constexpr bool BLOCKED = sizeof(T) == 4;

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
    constexpr int NP2 = (AC && BC) ? MI * NJ : 1;  // third product of the complex x complex case
    V4 acc_re[MI][NJ], acc_im[MI][NJ], acc_p2[NP2];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            acc_re[i][j] = V4{0, 0, 0, 0};
            acc_im[i][j] = V4{0, 0, 0, 0};
        }
#pragma unroll
    for (int q = 0; q < NP2; ++q) acc_p2[q] = V4{0, 0, 0, 0};

    // Slab staging: element e = tid + 256 q of the A slab is (row e / BK, k e % BK) and of
    // the B slab (k e / BN, col e % BN), so consecutive threads read consecutive addresses.
    T ra_re[AE], ra_im[AC ? AE : 1], rb_re[BE], rb_im[BC ? BE : 1];

    auto load_slab = [&](int k0) {
#pragma unroll
        for (int q = 0; q < AE; ++q) {
            const int e = tid + q * GEMM_THREADS;
            const int gm = m0 + e / GEMM_BK, gk = k0 + e % GEMM_BK;
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
#pragma unroll
        for (int q = 0; q < BE; ++q) {
            const int e = tid + q * GEMM_THREADS;
            const int gk = k0 + e / BN, gn = n0 + e % BN;
            T vr = 0, vi = 0;
            if (gk < it.k && gn < it.n) {
                const long off = (long)gk * it.ldb + gn;
                vr = it.b_re[off];
                if (BC) vi = it.b_im[off];
                if (PRE && BC) {
                    const T pr = it.pre_re[gk], pi = it.pre_im[gk];
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
            const int e = tid + q * GEMM_THREADS;
            As_re[b][e % GEMM_BK][e / GEMM_BK] = ra_re[q];
            if (AC) As_im[b][e % GEMM_BK][e / GEMM_BK] = ra_im[q];
        }
#pragma unroll
        for (int q = 0; q < BE; ++q) {
            const int e = tid + q * GEMM_THREADS;
            Bs_re[b][e / BN][e % BN] = rb_re[q];
            if (BC) Bs_im[b][e / BN][e % BN] = rb_im[q];
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
            // complex x complex uses three real products (Gauss):
            //   P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2
            constexpr int PB = BLOCKED ? MI * NJ : 1;
            V4 part_re[BLOCKED ? MI : 1][BLOCKED ? NJ : 1], part_im[BLOCKED ? MI : 1][BLOCKED ? NJ : 1];
            V4 part_p2[(BLOCKED && AC && BC) ? PB : 1];
            if constexpr (BLOCKED) {
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        part_re[i][j] = V4{0, 0, 0, 0};
                        part_im[i][j] = V4{0, 0, 0, 0};
                        if (AC && BC) part_p2[i * NJ + j] = V4{0, 0, 0, 0};
                    }
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
                        V4& cr = BLOCKED ? part_re[BLOCKED ? i : 0][BLOCKED ? j : 0] : acc_re[i][j];
                        V4& ci = BLOCKED ? part_im[BLOCKED ? i : 0][BLOCKED ? j : 0] : acc_im[i][j];
                        if (AC && BC) {
                            V4& c2 = BLOCKED ? part_p2[BLOCKED ? i * NJ + j : 0] : acc_p2[(AC && BC) ? i * NJ + j : 0];
                            cr = mfma16(ar[i], br[j], cr);
                            c2 = mfma16(ai[i], bi[j], c2);
                            ci = mfma16(ar[i] + ai[i], br[j] + bi[j], ci);
                        } else if (AC) {
                            cr = mfma16(ar[i], br[j], cr);
                            ci = mfma16(ai[i], br[j], ci);
                        } else if (BC) {
                            cr = mfma16(ar[i], br[j], cr);
                            ci = mfma16(ar[i], bi[j], ci);
                        } else {
                            cr = mfma16(ar[i], br[j], cr);
                        }
                    }
            }
            if constexpr (BLOCKED) {
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        if (AC && BC) {
                            acc_re[i][j] += part_re[i][j] - part_p2[i * NJ + j];
                            acc_im[i][j] += part_im[i][j] - part_re[i][j] - part_p2[i * NJ + j];
                        } else {
                            acc_re[i][j] += part_re[i][j];
                            if (CPLX) acc_im[i][j] += part_im[i][j];
                        }
                    }
            }
        }
        if (more) store_slab(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    if constexpr (!BLOCKED && AC && BC) {  // Gauss recombination of the three products
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                acc_im[i][j] = acc_im[i][j] - acc_re[i][j] - acc_p2[i * NJ + j];
                acc_re[i][j] = acc_re[i][j] - acc_p2[i * NJ + j];
            }
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

// Workgroups are dispatched round-robin over the 8 XCDs (linear id mod 8), each with its
// own L2.  xcd_map() re-labels the linear id so that all tiles of one item run on one
// XCD (items x, x+8, x+16, ... on XCD x): an item's operands then come from HBM into one
// L2 instead of up to eight.  Items beyond the last multiple of 8 keep the plain order.
// The map is a bijection of [0, tiles*nitems) onto (tile, item) pairs.
constexpr int N_XCD = 8;

__device__ __forceinline__ void xcd_map(int tiles, int nitems, int& tile, int& item) {
    const int lin = blockIdx.x + blockIdx.y * gridDim.x;
    const int nfull = (nitems / N_XCD) * N_XCD;
    if (lin < nfull * tiles) {
        const int x = lin % N_XCD, slot = lin / N_XCD;
        item = (slot / tiles) * N_XCD + x;
        tile = slot % tiles;
    } else {
        const int r = lin - nfull * tiles;
        item = nfull + r / tiles;
        tile = r % tiles;
    }
}

// grid: x = tile index over (ceil(m/BM) x ceil(n/BN)) of the largest item, y = item.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int GEMM_BK>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(const GemmItem<T>* __restrict__ items,
                                                            int tiles_n_max) {
    int tile, item;
    xcd_map(gridDim.x, gridDim.y, tile, item);
    const GemmItem<T> it = items[item];
    gemm_tile<T, AC, BC, PRE, POST, BM, BN, GEMM_BK>(it, tile, tiles_n_max);
}

// ---------------------------------------------------------------------------------------
// FP32 chain product with direct global->LDS loads (gfx950 global_load_lds_dwordx4):
//     C = POST(A * PRE(B)),  A real and SYMMETRIC (shared by all items), B complex planes.
// Because A = A^T, the k-rows of A are its columns, so both operand slabs are plain copies
// of row segments: slab rows k0..k0+31 of A (columns m0..) and of B (columns n0..).  The
// copies go straight into LDS (no staging registers), three slabs in flight.  In LDS a row
// of 4-float chunks is XOR-swizzled (chunk ^ 4 on odd rows) so the four k-rows read by one
// MFMA step fall into different banks; the swizzle is applied by choosing each lane's
// source chunk, since a wave's direct load writes 64 consecutive 16-byte chunks.
// Requirements (checked by the launcher): k % 32 == 0, m, n, lda, ldb, ldc % 4 == 0.
constexpr int RCD_BK = 32;

constexpr int RCD_STAGES = 3;

template <int BM, int BN>
struct RcdLayout {
    static constexpr int A = RCD_BK * BM;        // floats
    static constexpr int B = RCD_BK * BN;        // per plane
    static constexpr int PRE = 4 * 2 * RCD_BK;   // one copy of pre (re[32], im[32]) per wave
    static constexpr int STAGE = A + 2 * B + PRE;
    static constexpr size_t BYTES = sizeof(float) * STAGE * RCD_STAGES;
};

__device__ __forceinline__ int rcd_swz(int k, int m) { return ((((m >> 2) ^ ((k & 1) << 2))) << 2) | (m & 3); }

template <int POST, bool PRE, int BM, int BN>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_rcd_kernel(const GemmItem<float>* __restrict__ items,
                                                                int tiles_n_max) {
    typedef RcdLayout<BM, BN> LY;
    constexpr int MI = BM / 32, NJ = BN / 32;
    constexpr int ACH = RCD_BK * BM / 4, BCH = RCD_BK * BN / 4;  // 16-byte chunks per slab
    static_assert(ACH % 256 == 0 && BCH % 256 == 0, "slab chunks must cover whole waves");
    constexpr int AQ = ACH / 256, BQ = BCH / 256;                 // direct loads per wave per plane
    extern __shared__ float rcd_lds[];

    int tile, item;
    xcd_map(gridDim.x, gridDim.y, tile, item);
    const GemmItem<float> it = items[item];
    const int m0 = (tile / tiles_n_max) * BM, n0 = (tile % tiles_n_max) * BN;
    if (m0 >= it.m || n0 >= it.n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
    const int lr = lane & 15, lk = lane >> 4;
    const int nslab = it.k / RCD_BK;

    // (offsets as literal constants: the slab lambda must not name the layout's static members)
    constexpr int OFF_BR = RCD_BK * BM, OFF_BI = RCD_BK * (BM + BN), OFF_PRE = RCD_BK * (BM + 2 * BN);
    constexpr int STAGE_F = OFF_PRE + 4 * 2 * RCD_BK;
    static_assert(STAGE_F == LY::STAGE, "layout");
    auto issue = [&](int slab) {
        float* st = rcd_lds + (slab % RCD_STAGES) * STAGE_F;
        const int k0 = slab * RCD_BK;
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int c = (wave * AQ + q) * 64 + lane;           // chunk slot of this lane
            const int k = c / (BM / 4), slot = c % (BM / 4);
            int col = m0 + 4 * (slot ^ ((k & 1) << 2));
            if (col >= it.m) col = 0;                              // only feeds rows >= m of C
            __builtin_amdgcn_global_load_lds(it.a_re + (long)(k0 + k) * it.lda + col,
                                             st + (wave * AQ + q) * 256, 16, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int c = (wave * BQ + q) * 64 + lane;
            const int k = c / (BN / 4), slot = c % (BN / 4);
            int col = n0 + 4 * (slot ^ ((k & 1) << 2));
            if (col >= it.n) col = 0;                              // only feeds columns >= n of C
            const long off = (long)(k0 + k) * it.ldb + col;
            __builtin_amdgcn_global_load_lds(it.b_re + off, st + OFF_BR + (wave * BQ + q) * 256, 16, 0, 0);
            __builtin_amdgcn_global_load_lds(it.b_im + off, st + OFF_BI + (wave * BQ + q) * 256, 16, 0, 0);
        }
        if (PRE && lane < 16) {  // every wave copies pre[k0..k0+31] (re then im) into its own 64 floats
            const float* src = (lane < 8 ? it.pre_re : it.pre_im) + k0 + 4 * (lane & 7);
            __builtin_amdgcn_global_load_lds(src, st + OFF_PRE + wave * 64, 16, 0, 0);
        }
    };
    // direct loads issued per wave and slab (the pre copy is one instruction of 16 lanes)
    constexpr int PER_SLAB = AQ + 2 * BQ + (PRE ? 1 : 0);

    flt4 acc_re[MI][NJ], acc_im[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc_re[i][j] = acc_im[i][j] = flt4{0, 0, 0, 0};

    issue(0);
    if (nslab > 1) issue(1);
    for (int s = 0; s < nslab; ++s) {
        // wait for this wave's copies of slab s (slab s+1 may stay in flight), then for all
        // waves.  A plain s_barrier: __syncthreads() would also drain the copies in flight.
        if (s + 1 < nslab)
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(PER_SLAB) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        if (s + 2 < nslab) issue(s + 2);  // into the buffer every wave finished reading
        const float* st = rcd_lds + (s % RCD_STAGES) * LY::STAGE;
        const float* As = st;
        const float* Br = st + LY::A;
        const float* Bi = st + LY::A + LY::B;
        const float* pw = st + LY::A + 2 * LY::B + wave * 64;
#pragma unroll
        for (int kg = 0; kg < RCD_BK; kg += 16) {
            flt4 pr_[MI][NJ], pi_[MI][NJ];  // blocked summation: each 16 k in a fresh accumulator
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) pr_[i][j] = pi_[i][j] = flt4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const int kk = kg + ks * 4 + lk;
                float a[MI], br[NJ], bi[NJ];
#pragma unroll
                for (int i = 0; i < MI; ++i) a[i] = As[kk * BM + rcd_swz(kk, wm + i * 16 + lr)];
                float cr = 1.0f, ci = 0.0f;
                if (PRE) {
                    cr = pw[kk];
                    ci = pw[32 + kk];
                }
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    const int o = kk * BN + rcd_swz(kk, wn + j * 16 + lr);
                    float vr = Br[o], vi = Bi[o];
                    if (PRE) {
                        const float t = vr * cr - vi * ci;
                        vi = vr * ci + vi * cr;
                        vr = t;
                    }
                    br[j] = vr;
                    bi[j] = vi;
                }
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        pr_[i][j] = mfma16(a[i], br[j], pr_[i][j]);
                        pi_[i][j] = mfma16(a[i], bi[j], pi_[i][j]);
                    }
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    acc_re[i][j] += pr_[i][j];
                    acc_im[i][j] += pi_[i][j];
                }
        }
    }
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + acc_row<float>(lk, r);
                const int col = n0 + wn + j * 16 + lr;
                if (row < it.m && col < it.n) {
                    float vr = acc_re[i][j][r], vi = acc_im[i][j][r];
                    if (POST != SCALE_NONE) {
                        const int di = POST == SCALE_ROW ? row : col;
                        const float pr = it.post_re[di], pi = it.post_im[di];
                        const float t = vr * pr - vi * pi;
                        vi = vr * pi + vi * pr;
                        vr = t;
                    }
                    const long off = (long)row * it.ldc + col;
                    it.c_re[off] = vr;
                    it.c_im[off] = vi;
                }
            }
}

// Tile shapes the launcher can pick (the plan's autotuner times them per step):
//   0: 64x64, K slab 16   1: 32x32, K slab 32   2: 32x32, K slab 64   3: 64x64, K slab 32
//   4: 96x96, K slab 32   5: 96x96, K slab 16   (96 = 288 / 3: no partial tiles at Nx = 288)
//   6: 64x64 and 7: 96x96 direct-to-LDS (gemm_rcd_kernel): FP32, real symmetric A times a
//      complex B only; other shapes/types report "does not fit"
enum : int { N_TILES = 8 };

template <int POST, bool PRE, int BM, int BN>
inline bool launch_gemm_rcd(const GemmItem<float>* d_items, int nitems, int max_m, int max_n, hipStream_t s) {
    const size_t lds = RcdLayout<BM, BN>::BYTES;
    auto kern = gemm_rcd_kernel<POST, PRE, BM, BN>;
    ISLE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    const int tm = (max_m + BM - 1) / BM, tn = (max_n + BN - 1) / BN;
    hipLaunchKernelGGL(kern, dim3(tm * tn, nitems), dim3(GEMM_THREADS), lds, s, d_items, tn);
    ISLE_HIP(hipGetLastError());
    return true;
}

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

// direct tiles need: FP32, A real, square and symmetric, B complex, no A pre-scaling,
// k % 32 == 0, every dimension and leading dimension a multiple of 4 and 16-byte aligned
// operand / phase pointers (the caller vouches for all of it with `direct_ok`)
template <typename T, bool AC, bool BC, bool PRE, int POST>
inline bool launch_gemm(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s,
                        int tile = 0, bool direct_ok = false) {
    if (tile >= 6) {
        if constexpr (std::is_same<T, float>::value && !AC && BC && POST != SUBTRACT) {
            if (!direct_ok) return false;
            return tile == 6 ? launch_gemm_rcd<POST, PRE, 64, 64>(d_items, nitems, max_m, max_n, s)
                             : launch_gemm_rcd<POST, PRE, 96, 96>(d_items, nitems, max_m, max_n, s);
        } else {
            return false;
        }
    }
    switch (tile) {
    case 1: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 32>(d_items, nitems, max_m, max_n, s);
    case 2: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 64>(d_items, nitems, max_m, max_n, s);
    case 3: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 32>(d_items, nitems, max_m, max_n, s);
    case 4: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 96, 32>(d_items, nitems, max_m, max_n, s);
    case 5: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 96, 16>(d_items, nitems, max_m, max_n, s);
    default: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 16>(d_items, nitems, max_m, max_n, s);
    }
}

}  // namespace isle
