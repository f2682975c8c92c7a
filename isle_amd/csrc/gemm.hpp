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

// POST: SUBTRACT: C = C - A B; FUSED_DIAG (direct-to-LDS kernel only): with P = A B, the row
// sums of the elementwise product of P with a left factor L (or with the identity when L is
// null, i.e. diag P) are written per 16-column block: dpart[(s * m + i) * 2 + {0, 1}] =
// sum_{c in block s} L[i][c] P[i][c], s = 0 .. n/16 - 1 (FP64), and C = diag(post) P when
// post_re is set (the rows scaled by a complex vector after the sums), else C = P.
enum : int { SCALE_NONE = 0, SCALE_ROW = 1, SCALE_COL = 2, SUBTRACT = 3, FUSED_DIAG = 4 };

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
    const double* dl_re = nullptr;  // FUSED_DIAG: left factor L (m x n, row stride n), or null
    const double* dl_im = nullptr;
    double* dpart = nullptr;        // FUSED_DIAG: partial row sums [n/16][m] complex
};

constexpr int GEMM_THREADS = 256;

// global-memory access through an address-space-1 pointer (a generic access would also count
// against the LDS counter and wait for the slab copies into LDS); T without constructors
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T ld_global(const T* p) {
    return *(global_ptr<const T>)p;
}
template <typename T>
__device__ __forceinline__ void st_global(T* p, const T& v) {
    *(global_ptr<T>)p = v;
}


// Sum over the 16 lanes of a DPP row (the lanes that share lane / 16); lane 15 of the row ends
// with the total.  Both 32-bit halves of the double move with the same pattern.
template <int SHR>
__device__ __forceinline__ double dpp_shr_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x110 + SHR, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + SHR, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_shr_f64<1>(v);
    v += dpp_shr_f64<2>(v);
    v += dpp_shr_f64<4>(v);
    v += dpp_shr_f64<8>(v);
    return v;
}

// One BM x BN output tile of item `it` (tile index `tile` in a ceil(m/BM) x tiles_n_max
// grid).  AC/BC: A/B complex.  PRE: scale the (single) complex operand along k.
// POST: SCALE_* or SUBTRACT.  Must be called by all 256 threads of the block.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int GEMM_BK, bool G4 = false,
          bool SB = false>
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

    // A is stored transposed (k rows): consecutive threads write consecutive k, one LDS row
    // apart; one element of padding per row puts those writes in different banks
    // SB: one LDS buffer (the next slab waits in registers behind a second barrier): half the
    // LDS, so twice the workgroups per CU
    constexpr int NBUF = SB ? 1 : 2;
    __shared__ T As_re[NBUF][GEMM_BK][BM + 1];
    __shared__ T As_im[NBUF][AC ? GEMM_BK : 1][AC ? BM + 1 : 1];
    __shared__ T Bs_re[NBUF][GEMM_BK][BN];
    __shared__ T Bs_im[NBUF][BC ? GEMM_BK : 1][BC ? BN : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave >> 1) * (BM / 2);
    const int wn = (wave & 1) * (BN / 2);
    const int lr = lane & 15;
    const int lk = lane >> 4;

    typedef typename Acc4<T>::type V4;
    // complex x complex: three real products (Gauss) unless G4 asks for the four of the
    // textbook product (componentwise error bound: for recursions that amplify rounding)
    constexpr bool GAUSS = AC && BC && !G4;
    constexpr int NP2 = GAUSS ? MI * NJ : 1;  // third product of the Gauss form
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
            // Gauss: P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2
            constexpr int PB = BLOCKED ? MI * NJ : 1;
            V4 part_re[BLOCKED ? MI : 1][BLOCKED ? NJ : 1], part_im[BLOCKED ? MI : 1][BLOCKED ? NJ : 1];
            V4 part_p2[(BLOCKED && GAUSS) ? PB : 1];
            if constexpr (BLOCKED) {
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        part_re[i][j] = V4{0, 0, 0, 0};
                        part_im[i][j] = V4{0, 0, 0, 0};
                        if (GAUSS) part_p2[i * NJ + j] = V4{0, 0, 0, 0};
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
                        if (GAUSS) {
                            V4& c2 = BLOCKED ? part_p2[BLOCKED ? i * NJ + j : 0] : acc_p2[GAUSS ? i * NJ + j : 0];
                            cr = mfma16(ar[i], br[j], cr);
                            c2 = mfma16(ai[i], bi[j], c2);
                            ci = mfma16(ar[i] + ai[i], br[j] + bi[j], ci);
                        } else if (AC && BC) {
                            cr = mfma16(ar[i], br[j], cr);
                            cr = mfma16(-ai[i], bi[j], cr);
                            ci = mfma16(ar[i], bi[j], ci);
                            ci = mfma16(ai[i], br[j], ci);
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
                        if (GAUSS) {
                            acc_re[i][j] += part_re[i][j] - part_p2[i * NJ + j];
                            acc_im[i][j] += part_im[i][j] - part_re[i][j] - part_p2[i * NJ + j];
                        } else {
                            acc_re[i][j] += part_re[i][j];
                            if (CPLX) acc_im[i][j] += part_im[i][j];
                        }
                    }
            }
        }
        if constexpr (SB) {
            __syncthreads();  // every wave done reading the buffer
            if (more) store_slab(0);
            __syncthreads();
        } else {
            if (more) store_slab(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    if constexpr (!BLOCKED && GAUSS) {  // Gauss recombination of the three products
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
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int GEMM_BK, bool G4 = false,
          bool SB = false>
__global__ __launch_bounds__(GEMM_THREADS) void gemm_kernel(const GemmItem<T>* __restrict__ items,
                                                            int tiles_n_max) {
    int tile, item;
    xcd_map(gridDim.x, gridDim.y, tile, item);
    const GemmItem<T> it = items[item];
    gemm_tile<T, AC, BC, PRE, POST, BM, BN, GEMM_BK, G4, SB>(it, tile, tiles_n_max);
}

// ---------------------------------------------------------------------------------------
// Chain product with direct global->LDS loads (gfx950 global_load_lds_dwordx4), FP32 or FP64:
//     C = POST(A * PRE(B)),  A real and SYMMETRIC (shared by all items), B complex planes.
// Because A = A^T, the k-rows of A are its columns, so both operand slabs are plain copies
// of row segments: slab rows k0..k0+BK-1 of A (columns m0..) and of B (columns n0..).  The
// copies go straight into LDS (no staging registers), three slabs in flight.  In LDS a row
// of 16-byte chunks is XOR-swizzled (chunk ^ 4, or ^ 8 for FP64, on odd rows; rcd_xor) so
// the k-rows read by one MFMA step fall into different banks; the swizzle is applied by
// choosing each lane's source chunk, since a wave's direct load writes 64 consecutive
// 16-byte chunks.
// Requirements (the caller vouches): k % BK == 0; m, n, lda, ldb, ldc multiples of 4;
// 16-byte aligned operands and phase vectors.
template <typename T, int BM, int BN, int BK, int RCD_STAGES = 3, int NW = 4>
struct RcdLayout {
    static constexpr int EPC = 16 / sizeof(T);   // elements per 16-byte chunk
    static constexpr int A = BK * BM;            // elements
    static constexpr int B = BK * BN;            // per plane
    static constexpr int PRE = NW * 2 * BK;      // one copy of pre (re[BK], im[BK]) per wave
    static constexpr int STAGE = A + 2 * B + PRE;
    static constexpr int FAC = STAGE * RCD_STAGES;  // then the tile's row or column factors
    static constexpr int NFAC = BM > BN ? BM : BN;  // re[NFAC], im[NFAC]
    static constexpr size_t BYTES = sizeof(T) * (FAC + 2 * NFAC);
};

// XOR applied to the chunk index of odd k rows: 8 chunks (128 bytes) for FP64 rows of a
// multiple of 16 chunks, so that two k rows read in one LDS cycle fall in disjoint halves of
// a 64-bank line (4 chunks overlapped half of them: 27% of the FP64 kernel's LDS cycles were
// bank conflicts); 4 chunks otherwise.  ROW: elements per slab row (BM or BN).
template <int EPC, int ROW>
constexpr int rcd_xor() {
    return (EPC == 2 && (ROW / EPC) % 16 == 0) ? 8 : 4;
}

template <int EPC, int ROW>
__device__ __forceinline__ int rcd_swz(int k, int m) {
    return (((m / EPC) ^ ((k & 1) * rcd_xor<EPC, ROW>())) * EPC) + (m % EPC);
}

// Issues this wave's direct copies of slab `slab` (A rows, B re/im rows, and with PRE the
// phase factors) into stage slab % RCD_STAGES of the LDS ring.  POST is not used here: it
// gives kernels of different POST modes their own instantiation (the host compilation pass
// rejects one specialization of this device template shared by two of them).
template <typename T, bool PRE, int BM, int BN, int BK, int RCD_STAGES, int NW, int POST>
__device__ __forceinline__ void rcd_issue(const GemmItem<T>& it, T* lds, int slab, int m0, int n0, int wave,
                                          int lane) {
    constexpr int EPC = 16 / sizeof(T);
    constexpr int AQ = BK * BM / EPC / (64 * NW), BQ = BK * BN / EPC / (64 * NW);
    constexpr int OFF_BR = BK * BM, OFF_BI = BK * (BM + BN), OFF_PRE = BK * (BM + 2 * BN);
    constexpr int STAGE_E = OFF_PRE + NW * 2 * BK;
    constexpr int CH = 64 * EPC;  // elements written by one wave instruction
    constexpr int PC = BK / EPC;  // chunks of one phase plane
    T* st = lds + (slab % RCD_STAGES) * STAGE_E;
    const int k0 = slab * BK;
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
        const int c = (wave * AQ + q) * 64 + lane;  // chunk slot of this lane
        const int k = c / (BM / EPC), slot = c % (BM / EPC);
        int col = m0 + EPC * (slot ^ ((k & 1) * rcd_xor<EPC, BM>()));
        if (col >= it.m) col = 0;  // only feeds rows >= m of C
        __builtin_amdgcn_global_load_lds(it.a_re + (long)(k0 + k) * it.lda + col, st + (wave * AQ + q) * CH, 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
        const int c = (wave * BQ + q) * 64 + lane;
        const int k = c / (BN / EPC), slot = c % (BN / EPC);
        int col = n0 + EPC * (slot ^ ((k & 1) * rcd_xor<EPC, BN>()));
        if (col >= it.n) col = 0;  // only feeds columns >= n of C
        const long off = (long)(k0 + k) * it.ldb + col;
        __builtin_amdgcn_global_load_lds(it.b_re + off, st + OFF_BR + (wave * BQ + q) * CH, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(it.b_im + off, st + OFF_BI + (wave * BQ + q) * CH, 16, 0, 0);
    }
    if (PRE && lane < 2 * PC) {  // every wave copies pre[k0..k0+BK) (re then im) into its own 2*BK elements
        const T* src = (lane < PC ? it.pre_re : it.pre_im) + k0 + EPC * (lane % PC);
        __builtin_amdgcn_global_load_lds(src, st + OFF_PRE + wave * 2 * BK, 16, 0, 0);
    }
}

// Waves per workgroup: WM along m times WN along n, each owning a (BM/WM) x (BN/WN) block.
template <typename T, int BM, int BN, int BK, int WM, int WN>
constexpr bool rcd_shape_ok() {
    constexpr int EPC = 16 / sizeof(T), NW = WM * WN;
    return (BK * BM / EPC) % (64 * NW) == 0 && (BK * BN / EPC) % (64 * NW) == 0 && (BM / EPC) % 8 == 0 &&
           (BN / EPC) % 8 == 0 && 2 * BK / EPC <= 64 && BK % 16 == 0 && BM % (16 * WM) == 0 && BN % (16 * WN) == 0;
}

// The 12-wave tiles ask for 6 waves per SIMD (at most 80 VGPRs): two workgroups per CU, whose
// 2-stage LDS rings (74 KB in FP64, 80 KB in FP32) fit the 160 KB together.  The FP64
// kernels other than the row-scaled one would spill at that bound and keep the default.
template <typename T, int POST, bool PRE, int BM, int BN, int BK, int RCD_STAGES, int WM = 2, int WN = 2>
__global__ __launch_bounds__(64 * WM * WN, (WM * WN >= 12 && (sizeof(T) == 4 || POST == SCALE_ROW) ? 6 : 1))
void gemm_rcd_kernel(const GemmItem<T>* __restrict__ items, int tiles_n_max) {
    constexpr int NW = WM * WN;
    typedef RcdLayout<T, BM, BN, BK, RCD_STAGES, NW> LY;
    constexpr int EPC = 16 / sizeof(T);
    constexpr int MI = BM / (16 * WM), NJ = BN / (16 * WN);
    constexpr int ACH = BK * BM / EPC, BCH = BK * BN / EPC;  // 16-byte chunks per slab
    static_assert(rcd_shape_ok<T, BM, BN, BK, WM, WN>(), "tile shape");
    static_assert(ACH % (64 * NW) == 0 && BCH % (64 * NW) == 0, "slab chunks must cover whole waves");
    static_assert((BM / EPC) % 8 == 0 && (BN / EPC) % 8 == 0, "swizzle needs multiples of 8 chunks per row");
    static_assert(2 * BK / EPC <= 64 && BK % 16 == 0, "pre copy must fit one wave instruction");
    static_assert(RCD_STAGES >= 2 && RCD_STAGES <= 4, "one to three slabs in flight");
    constexpr int AQ = ACH / (64 * NW), BQ = BCH / (64 * NW);  // direct loads per wave per plane
    // FP32 sums each group of 16 k in a fresh accumulator (blocked summation: about half the
    // rounding error of one running sum over K = 288), except in the 12-wave tiles, where the
    // second set of accumulators would cost the second workgroup per CU
    constexpr bool BLOCKED = sizeof(T) == 4 && NW < 12;
    typedef typename Acc4<T>::type V4;
    extern __shared__ __align__(16) unsigned char rcd_raw[];
    T* rcd_lds = reinterpret_cast<T*>(rcd_raw);

    int tile, item;
    xcd_map(gridDim.x, gridDim.y, tile, item);
    const GemmItem<T> it = items[item];
    const int m0 = (tile / tiles_n_max) * BM, n0 = (tile % tiles_n_max) * BN;
    if (m0 >= it.m || n0 >= it.n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WN) * (BM / WM), wn = (wave % WN) * (BN / WN);
    const int lr = lane & 15, lk = lane >> 4;
    const int nslab = it.k / BK;

    constexpr int OFF_BR = BK * BM, OFF_BI = BK * (BM + BN), OFF_PRE = BK * (BM + 2 * BN);
    constexpr int STAGE_E = OFF_PRE + NW * 2 * BK;
    static_assert(STAGE_E == LY::STAGE, "layout");
    // direct loads issued per wave and slab (the pre copy is one instruction)
    constexpr int PER_SLAB = AQ + 2 * BQ + (PRE ? 1 : 0);

    V4 acc_re[MI][NJ], acc_im[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc_re[i][j] = acc_im[i][j] = V4{0, 0, 0, 0};

    // slabs are issued RCD_STAGES-1 ahead of the one being computed
    for (int j = 0; j < RCD_STAGES - 1 && j < nslab; ++j)
        rcd_issue<T, PRE, BM, BN, BK, RCD_STAGES, NW, POST>(it, rcd_lds, j, m0, n0, wave, lane);
    // the tile's row factors (column factors with SCALE_COL) into LDS, read by the epilogue
    // after the k loop's barriers (their loads in flight with the first slabs; 1 where absent)
    constexpr bool SCALED = POST == SCALE_ROW || POST == SCALE_COL;
    constexpr int NF = POST == SCALE_COL ? BN : BM;
    T* fac = rcd_lds + LY::FAC;
    if constexpr (SCALED || POST == FUSED_DIAG) {
        const int base = POST == SCALE_COL ? n0 : m0, lim = POST == SCALE_COL ? it.n : it.m;
        for (int t = tid; t < 2 * NF; t += 64 * NW) {
            const int e = base + t % NF;
            const T* src = t < NF ? it.post_re : it.post_im;
            T v = t < NF ? T(1) : T(0);
            if (it.post_re != nullptr && e < lim) v = ld_global(src + e);
            fac[(t < NF ? 0 : LY::NFAC) + t % NF] = v;
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }
    for (int s = 0; s < nslab; ++s) {
        // wait for this wave's copies of slab s (the later slabs already issued stay in
        // flight), then for all waves.  A plain s_barrier: __syncthreads() would also drain
        // the copies in flight.
        const int ahead = min(RCD_STAGES - 2, nslab - 1 - s);
        if (ahead >= 2)
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(2 * PER_SLAB) : "memory");
        else if (ahead == 1)
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(PER_SLAB) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        if (s + RCD_STAGES - 1 < nslab)  // into the buffer every wave finished reading
            rcd_issue<T, PRE, BM, BN, BK, RCD_STAGES, NW, POST>(it, rcd_lds, s + RCD_STAGES - 1, m0, n0, wave, lane);
        const T* st = rcd_lds + (s % RCD_STAGES) * STAGE_E;
        const T* As = st;
        const T* Br = st + OFF_BR;
        const T* Bi = st + OFF_BI;
        const T* pw = st + OFF_PRE + wave * 2 * BK;
#pragma unroll
        for (int kg = 0; kg < BK; kg += 16) {
            V4 pr_[BLOCKED ? MI : 1][BLOCKED ? NJ : 1], pi_[BLOCKED ? MI : 1][BLOCKED ? NJ : 1];
            if constexpr (BLOCKED) {
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) pr_[i][j] = pi_[i][j] = V4{0, 0, 0, 0};
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const int kk = kg + ks * 4 + lk;
                T a[MI], br[NJ], bi[NJ];
#pragma unroll
                for (int i = 0; i < MI; ++i) a[i] = As[kk * BM + rcd_swz<EPC, BM>(kk, wm + i * 16 + lr)];
                T cr = 1, ci = 0;
                if (PRE) {
                    cr = pw[kk];
                    ci = pw[BK + kk];
                }
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    const int o = kk * BN + rcd_swz<EPC, BN>(kk, wn + j * 16 + lr);
                    T vr = Br[o], vi = Bi[o];
                    if (PRE) {
                        const T t = vr * cr - vi * ci;
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
                        V4& xr = BLOCKED ? pr_[BLOCKED ? i : 0][BLOCKED ? j : 0] : acc_re[i][j];
                        V4& xi = BLOCKED ? pi_[BLOCKED ? i : 0][BLOCKED ? j : 0] : acc_im[i][j];
                        xr = mfma16(a[i], br[j], xr);
                        xi = mfma16(a[i], bi[j], xi);
                    }
            }
            if constexpr (BLOCKED) {
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        acc_re[i][j] += pr_[i][j];
                        acc_im[i][j] += pi_[i][j];
                    }
            }
        }
    }
    // Epilogue, one block of 16 rows at a time.  The memory operands of a block (row or column
    // factors, the left factor of the row sums) are all loaded, branch-free, before the first
    // of them is used: one memory latency per block instead of one per element.  Elements
    // outside the matrix load element 0 and discard it.
    // FUSED_DIAG with a row factor (uniform over the workgroup): the stored rows are scaled
    const bool diag_rows = POST == FUSED_DIAG && it.post_re != nullptr;
    if (nslab == 0) __syncthreads();  // the factors
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        T fr[NJ][4], fi[NJ][4];
        double lre[NJ][4], lim[NJ][4];
        if constexpr (SCALED || POST == FUSED_DIAG) {
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = POST == SCALE_COL ? wn + j * 16 + lr : wm + i * 16 + acc_row<T>(lk, r);
                    fr[j][r] = fac[e];
                    fi[j][r] = fac[LY::NFAC + e];
                }
        }
        if constexpr (POST == FUSED_DIAG) {
            const bool identity = it.dl_re == nullptr;  // uniform over the workgroup
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm + i * 16 + acc_row<T>(lk, r);
                    const int col = n0 + wn + j * 16 + lr;
                    const bool in = row < it.m && col < it.n;
                    lre[j][r] = in && row == col ? 1.0 : 0.0;
                    lim[j][r] = 0;
                }
            if (!identity) {
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = m0 + wm + i * 16 + acc_row<T>(lk, r);
                        const int col = n0 + wn + j * 16 + lr;
                        const bool in = row < it.m && col < it.n;
                        const long o = in ? (long)row * it.n + col : 0;
                        const double a = it.dl_re[o], b = it.dl_im[o];
                        lre[j][r] = in ? a : 0.0;
                        lim[j][r] = in ? b : 0.0;
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + acc_row<T>(lk, r);
                const int col = n0 + wn + j * 16 + lr;
                const bool in = row < it.m && col < it.n;
                T vr = acc_re[i][j][r], vi = acc_im[i][j][r];
                if constexpr (POST == FUSED_DIAG) {  // every lane takes part in the row sums
                    const double sr = row16_sum(lre[j][r] * (double)vr - lim[j][r] * (double)vi);
                    const double si = row16_sum(lre[j][r] * (double)vi + lim[j][r] * (double)vr);
                    const int slot = (n0 + wn + j * 16) / 16;
                    if (lr == 15 && row < it.m && slot < it.n / 16) {
                        it.dpart[((long)slot * it.m + row) * 2] = sr;
                        it.dpart[((long)slot * it.m + row) * 2 + 1] = si;
                    }
                }
                if (SCALED || diag_rows) {
                    const T t = vr * fr[j][r] - vi * fi[j][r];
                    vi = vr * fi[j][r] + vi * fr[j][r];
                    vr = t;
                }
                if (in) {
                    const long off = (long)row * it.ldc + col;
                    it.c_re[off] = vr;
                    it.c_im[off] = vi;
                }
            }
    }
}

// ---------------------------------------------------------------------------------------
// FP32 chain product with k-contiguous operand slabs (the fused right chains of the force):
//     C = A B,  A real (row-major, shared by all items), B complex given by its transpose
//     Bt (n x k, row-major split planes),
// so both slabs are row segments contiguous in k: rows m of A and rows n of Bt, columns
// k0 .. k0 + BK.  In LDS a slab row holds BK / 4 16-byte chunks; logical chunk c of row r sits
// at position c ^ rck_swz(r), which puts the 16 rows read by one MFMA step in distinct banks
// for 32- and for 64-bank LDS.  Lane (lr, lk) feeds k = kg + 4 lk + ks to MFMA step ks of a
// 16-k group, so one ds_read_b128 brings a lane's operand for four MFMAs (a quarter of the LDS
// instructions of gemm_rcd_kernel).  The product is stored transposed, Ct = C^T (n x m, the
// next chain step's Bt), one 16-byte store per lane and 16 x 16 block, with FUSED_DIAG (when
// dpart is set): the row sums of L .* C per 16-column block (as gemm_rcd_kernel) from the
// unscaled product, and C^T scaled by post[m] when post_re is set (the tile's row factors
// staged in LDS).
// Requirements (the caller vouches): k % BK == 0, m % 4 == 0, the row strides lda, ldb
// (of Bt) and ldc (of Ct) multiples of 4, 16-byte aligned operands, each operand plane below
// 4 GiB.
template <int BM, int BN, int BK, int STAGES>
struct RckLayout {
    static constexpr int A = BM * BK;  // floats per slab plane
    static constexpr int B = BN * BK;
    static constexpr int STAGE = A + 2 * B;
    static constexpr int POST = STAGES * STAGE;  // then the tile's row factors, 2 BM floats
    static constexpr size_t BYTES = sizeof(float) * (STAGE * STAGES + 2 * BM);
};

__device__ __forceinline__ int rck_swz(int r) { return (r & 7) ^ ((r >> 3) & 1); }

template <int BM, int BN, int BK, int WM, int WN>
constexpr bool rck_shape_ok() {
    constexpr int NW = WM * WN, RPI = 64 / (BK / 4);  // rows written by one wave instruction
    return BK == 32 && BM % RPI == 0 && BN % RPI == 0 && (BM / RPI) % NW == 0 && (BN / RPI) % NW == 0 &&
           BM % (16 * WM) == 0 && BN % (16 * WN) == 0 && 64 * NW >= 2 * BM;
}

// This wave's direct copies of k-slab `slab` (columns slab*BK ..) into stage `stage` of the
// LDS ring: uniform slab bases plus 32-bit per-lane byte offsets (one address register each).
template <int BM, int BN, int BK, int STAGES, int NW>
__device__ __forceinline__ void rck_issue(const GemmItem<float>& it, float* lds, int stage, int slab, int m0, int n0,
                                          int wave, int lane) {
    typedef RckLayout<BM, BN, BK, STAGES> LY;
    constexpr int CPR = BK / 4, RPI = 64 / CPR;
    constexpr int AQ = BM / RPI / NW, BQ = BN / RPI / NW;
    float* st = lds + stage * LY::STAGE;
    const int k0 = slab * BK;
    const int rr = lane / CPR, pos = lane % CPR;
    const char* ab = reinterpret_cast<const char*>(it.a_re + k0);
    const char* bbr = reinterpret_cast<const char*>(it.b_re + k0);
    const char* bbi = reinterpret_cast<const char*>(it.b_im + k0);
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
        const int ins = wave * AQ + q, r = ins * RPI + rr;
        int row = m0 + r;
        if (row >= it.m) row = 0;  // only feeds rows >= m of C
        const unsigned off = 4u * (unsigned)(row * it.lda + 4 * (pos ^ rck_swz(r)));
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const float*>(ab + off), st + ins * 256, 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
        const int ins = wave * BQ + q, r = ins * RPI + rr;
        int row = n0 + r;
        if (row >= it.n) row = 0;  // only feeds columns >= n of C
        const unsigned off = 4u * (unsigned)(row * it.ldb + 4 * (pos ^ rck_swz(r)));
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const float*>(bbr + off), st + LY::A + ins * 256, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const float*>(bbi + off), st + LY::A + LY::B + ins * 256, 16,
                                         0, 0);
    }
}

// Left-factor entries L[rowb + r][col] of one 16-row block's NJ 16-column blocks (the
// identity when dl_re is null).  Lane (lr, lk) holds C[4 lk + r][lr] of each 16 x 16 block;
// rowb = the block's first row + 4 lk.
template <int NJ>
struct RckEpi {
    double lre[NJ][4], lim[NJ][4];
};

template <int NJ>
__device__ __forceinline__ void rck_epi_load(const GemmItem<float>& it, int rowb, int colb, int lr, RckEpi<NJ>& e) {
    const bool rows_in = rowb < it.m;  // m % 4 == 0: all four rows or none
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = colb + j * 16 + lr;
            const bool in = rows_in && col < it.n;
            if (it.dl_re == nullptr) {  // uniform over the workgroup
                e.lre[j][r] = in && rowb + r == col ? 1.0 : 0.0;
                e.lim[j][r] = 0.0;
            } else {
                const long o = in ? (long)(rowb + r) * it.n + col : 0;
                const double x = ld_global(it.dl_re + o), y = ld_global(it.dl_im + o);
                e.lre[j][r] = in ? x : 0.0;
                e.lim[j][r] = in ? y : 0.0;
            }
        }
}

// The row sums of L .* C per 16-column block (from the unscaled product) into dpart, and the
// product scaled by the row factors (pr, pi: the tile's, from LDS) stored transposed:
// Ct[col][rowb .. rowb + 3], one 16-byte store per lane.
template <int NJ>
__device__ __forceinline__ void rck_epi_store(const GemmItem<float>& it, int rowb, int rloc, int colb, int lr,
                                              const float* pr, const float* pi, const RckEpi<NJ>& e,
                                              const flt4 (&are)[NJ], const flt4 (&aim)[NJ]) {
    const bool rows_in = rowb < it.m;
    float fr[4], fi[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        fr[r] = pr[rloc + r];
        fi[r] = pi[rloc + r];
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int col = colb + j * 16 + lr;
        flt4 outr, outi;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float vr = are[j][r], vi = aim[j][r];
            const double sr = row16_sum(e.lre[j][r] * (double)vr - e.lim[j][r] * (double)vi);
            const double si = row16_sum(e.lre[j][r] * (double)vi + e.lim[j][r] * (double)vr);
            const int slot = (colb + j * 16) / 16;
            if (it.dpart != nullptr && lr == 15 && rows_in && slot < it.n / 16) {
                st_global(it.dpart + ((long)slot * it.m + rowb + r) * 2, sr);
                st_global(it.dpart + ((long)slot * it.m + rowb + r) * 2 + 1, si);
            }
            outr[r] = vr * fr[r] - vi * fi[r];
            outi[r] = vr * fi[r] + vi * fr[r];
        }
        if (rows_in && col < it.n) {
            const long off = (long)col * it.ldc + rowb;
            st_global(reinterpret_cast<flt4*>(it.c_re + off), outr);
            st_global(reinterpret_cast<flt4*>(it.c_im + off), outi);
        }
    }
}

template <int BM, int BN, int BK, int STAGES, int WM = 2, int WN = 2>
__global__ __launch_bounds__(64 * WM * WN, (WM * WN >= 12 ? 6 : 1))
void gemm_rck_kernel(const GemmItem<float>* __restrict__ items, int tiles_n_max) {
    constexpr int NW = WM * WN;
    typedef RckLayout<BM, BN, BK, STAGES> LY;
    static_assert(rck_shape_ok<BM, BN, BK, WM, WN>(), "tile shape");
    static_assert(STAGES >= 2 && STAGES <= 4, "one to three slabs in flight");
    constexpr int MI = BM / (16 * WM), NJ = BN / (16 * WN);
    constexpr int RPI = 64 / (BK / 4);
    constexpr int PER_SLAB = BM / RPI / NW + 2 * (BN / RPI / NW);  // direct loads per wave and slab
    extern __shared__ __align__(16) unsigned char rck_raw[];
    float* lds = reinterpret_cast<float*>(rck_raw);

    int tile, item;
    xcd_map(gridDim.x, gridDim.y, tile, item);
    const GemmItem<float> it = items[item];
    const int m0 = (tile / tiles_n_max) * BM, n0 = (tile % tiles_n_max) * BN;
    if (m0 >= it.m || n0 >= it.n) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave / WN) * (BM / WM), wn = (wave % WN) * (BN / WN);
    const int lr = lane & 15, lk = lane >> 4;
    const int nslab = it.k / BK;

    flt4 acc_re[MI][NJ], acc_im[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc_re[i][j] = acc_im[i][j] = flt4{0, 0, 0, 0};

    for (int j = 0; j < STAGES - 1 && j < nslab; ++j)
        rck_issue<BM, BN, BK, STAGES, NW>(it, lds, j % STAGES, j, m0, n0, wave, lane);
    // the tile's row factors into LDS, read by the stores after the k loop's barriers (its
    // load in flight with the first slabs)
    if (tid < 2 * BM) {
        const int r = tid % BM, row = m0 + r;
        const float* src = tid < BM ? it.post_re : it.post_im;
        float v = tid < BM ? 1.0f : 0.0f;
        if (it.post_re != nullptr && row < it.m) v = ld_global(src + row);
        lds[LY::POST + tid] = v;
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    for (int s = 0; s < nslab; ++s) {
        // this wave's copies of slab s done (later slabs stay in flight), then all waves'
        const int ahead = min(STAGES - 2, nslab - 1 - s);
        if (ahead >= 2)
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(2 * PER_SLAB) : "memory");
        else if (ahead == 1)
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(PER_SLAB) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        if (s + STAGES - 1 < nslab) {  // into the buffer every wave finished reading
            // the lane index read again (v_mbcnt, no inputs), so the copies' per-lane offsets
            // are formed here instead of held in registers over the loop
            int lanev;
            asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lanev));
            rck_issue<BM, BN, BK, STAGES, NW>(it, lds, (s + STAGES - 1) % STAGES, s + STAGES - 1, m0, n0, wave, lanev);
        }
        const float* st = lds + (s % STAGES) * LY::STAGE;
        const float* As = st;
        const float* Br = st + LY::A;
        const float* Bi = st + LY::A + LY::B;
#pragma unroll
        for (int kg = 0; kg < BK; kg += 16) {
            const int c = kg / 4 + lk;
            flt4 a[MI], br[NJ], bi[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const int r = wm + i * 16 + lr;
                a[i] = *reinterpret_cast<const flt4*>(As + r * BK + 4 * (c ^ rck_swz(r)));
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int r = wn + j * 16 + lr;
                const int o = r * BK + 4 * (c ^ rck_swz(r));
                br[j] = *reinterpret_cast<const flt4*>(Br + o);
                bi[j] = *reinterpret_cast<const flt4*>(Bi + o);
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        acc_re[i][j] = mfma16(a[i][ks], br[j][ks], acc_re[i][j]);
                        acc_im[i][j] = mfma16(a[i][ks], bi[j][ks], acc_im[i][j]);
                    }
            // keep the next group's operand reads below this group's MFMAs (six waves per SIMD
            // hide their latency; hoisted, they would double the operand registers)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (nslab == 0) __syncthreads();  // the row factors
    // Epilogue per 16-row block: left-factor entries, then the row sums and the stores
    const float* pr = lds + LY::POST;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        RckEpi<NJ> ep;
        const int rloc = wm + i * 16 + 4 * lk;
        rck_epi_load<NJ>(it, m0 + rloc, n0 + wn, lr, ep);
        rck_epi_store<NJ>(it, m0 + rloc, rloc, n0 + wn, lr, pr, pr + BM, ep, acc_re[i], acc_im[i]);
    }
}

template <int BM, int BN, int BK, int ST, int WM, int WN>
inline bool launch_gemm_rck(const GemmItem<float>* d_items, int nitems, int max_m, int max_n, hipStream_t s) {
    if constexpr (!rck_shape_ok<BM, BN, BK, WM, WN>()) {
        return false;
    } else {
        const size_t lds = RckLayout<BM, BN, BK, ST>::BYTES;
        if (lds > 160 * 1024) return false;
        auto kern = gemm_rck_kernel<BM, BN, BK, ST, WM, WN>;
        ISLE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const int tm = (max_m + BM - 1) / BM, tn = (max_n + BN - 1) / BN;
        hipLaunchKernelGGL(kern, dim3(tm * tn, nitems), dim3(64 * WM * WN), lds, s, d_items, tn);
        ISLE_HIP(hipGetLastError());
        return true;
    }
}

// Tile shapes of the k-contiguous FP32 chain product (the autotuner times them per step):
//   0: 96x96, 12 waves of 48x16, 2 stages   1: the same, 3 stages   2: 96x96, 6 waves of 48x32,
//   2 stages   3: 64x64, 4 waves of 32x32, 3 stages; other values report "does not fit".
inline bool launch_gemm_rck_tile(const GemmItem<float>* d_items, int nitems, int max_m, int max_n, hipStream_t s,
                                 int tile) {
    switch (tile) {
    case 0: return launch_gemm_rck<96, 96, 32, 2, 2, 6>(d_items, nitems, max_m, max_n, s);
    case 1: return launch_gemm_rck<96, 96, 32, 3, 2, 6>(d_items, nitems, max_m, max_n, s);
    case 2: return launch_gemm_rck<96, 96, 32, 2, 2, 3>(d_items, nitems, max_m, max_n, s);
    case 3: return launch_gemm_rck<64, 64, 32, 3, 2, 2>(d_items, nitems, max_m, max_n, s);
    default: return false;
    }
}

// Tile shapes the launcher can pick (the plan's autotuner times them per step):
//   0: 64x64, K slab 16   1: 32x32, K slab 32   2: 32x32, K slab 64   3: 64x64, K slab 32
//   4: 96x96, K slab 32   5: 96x96, K slab 16   (96 = 288 / 3: no partial tiles at Nx = 288)
//   direct-to-LDS (gemm_rcd_kernel; real symmetric A times a complex B only, other operand
//   kinds report "does not fit"), K slab S = 32 in FP32, 16 in FP64:
//   6: 64x64 slab S   7: 96x96 slab S   8: 32x32 slab S   9: 64x64 slab 2S   10: 64x64 slab S, 4 stages
//   11: 32x32 slab S, 2 stages   12: 32x32 slab 2S, 2 stages   13: 32x32 slab S, 4 stages
//   14: 32x32 slab 2S  (fewer stages: less LDS per workgroup, more workgroups per CU)
//   15: 96x32 slab S   16: 96x32 slab S, 2 stages   17: 96x32 slab 2S, 2 stages
//   (48x16 per wave: 5 LDS fragment reads per 6 MFMAs instead of 3 per 2 at 32x32)
//   register-staged again (any operand kinds): 18: 96x32, K slab 32   19: 96x32, K slab 16
//   direct, more waves sharing one LDS ring (48x16 per wave): 20: 96x96 slab S, 12 waves, 2 stages
//   21: 96x96 slab S, 12 waves, 3 stages   22: 96x48 slab S, 6 waves, 2 stages (FP64 only)
//   23: 96x96 slab S, 6 waves of 48x32, 2 stages   24: the same, 3 stages
//   register-staged with one LDS buffer (half the LDS: twice the workgroups per CU, a second
//   barrier per slab): 25: 96x32, K slab 16   26: 96x32, K slab 32   27: 64x64, K slab 16
enum : int { N_TILES = 28 };

template <typename T, int POST, bool PRE, int BM, int BN, int BK, int ST = 3, int WM = 2, int WN = 2>
inline bool launch_gemm_rcd(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s) {
    if constexpr (!rcd_shape_ok<T, BM, BN, BK, WM, WN>()) {
        return false;
    } else {
        const size_t lds = RcdLayout<T, BM, BN, BK, ST, WM * WN>::BYTES;
        if (lds > 160 * 1024) return false;
        auto kern = gemm_rcd_kernel<T, POST, PRE, BM, BN, BK, ST, WM, WN>;
        ISLE_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const int tm = (max_m + BM - 1) / BM, tn = (max_n + BN - 1) / BN;
        hipLaunchKernelGGL(kern, dim3(tm * tn, nitems), dim3(64 * WM * WN), lds, s, d_items, tn);
        ISLE_HIP(hipGetLastError());
        return true;
    }
}

template <typename T, bool AC, bool BC, int BM, int BN, int BK, bool SB = false>
constexpr size_t gemm_lds_bytes() {
    return sizeof(T) * (SB ? 1 : 2) * BK * ((BM + 1) * (AC ? 2 : 1) + BN * (BC ? 2 : 1));
}

// Launches the GEMM with tile shape `tile`; returns false if that shape does not fit
// this operand type (LDS budget), in which case nothing is launched.
template <typename T, bool AC, bool BC, bool PRE, int POST, int BM, int BN, int BK, bool G4 = false, bool SB = false>
inline bool launch_gemm_tile(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s) {
    if constexpr (gemm_lds_bytes<T, AC, BC, BM, BN, BK, SB>() > 96 * 1024) {
        return false;
    } else {
        const int tm = (max_m + BM - 1) / BM, tn = (max_n + BN - 1) / BN;
        hipLaunchKernelGGL((gemm_kernel<T, AC, BC, PRE, POST, BM, BN, BK, G4, SB>), dim3(tm * tn, nitems),
                           dim3(GEMM_THREADS), 0, s, d_items, tn);
        ISLE_HIP(hipGetLastError());
        return true;
    }
}

// Direct tiles (6..10) need: A real, square and symmetric, B complex, no A pre-scaling,
// every dimension and leading dimension a multiple of 4 and 16-byte aligned operand /
// phase pointers (the caller vouches for all of it by passing direct_k, the K dimension
// shared by every item, or 0 when not eligible); K must be a multiple of the tile's K slab,
// which is checked here.
template <typename T, bool AC, bool BC, bool PRE, int POST, bool G4 = false>
inline bool launch_gemm(const GemmItem<T>* d_items, int nitems, int max_m, int max_n, hipStream_t s,
                        int tile = 0, int direct_k = 0) {
    if ((tile >= 6 && tile < 18) || (tile >= 20 && tile < 25)) {
        if constexpr (!AC && BC && POST != SUBTRACT) {
            constexpr int BK = sizeof(T) == 4 ? 32 : 16;
            const int slab = (tile == 9 || tile == 12 || tile == 14 || tile == 17) ? 2 * BK : BK;
            if (direct_k <= 0 || direct_k % slab != 0) return false;
            switch (tile) {
            case 6: return launch_gemm_rcd<T, POST, PRE, 64, 64, BK>(d_items, nitems, max_m, max_n, s);
            case 7: return launch_gemm_rcd<T, POST, PRE, 96, 96, BK>(d_items, nitems, max_m, max_n, s);
            case 8: return launch_gemm_rcd<T, POST, PRE, 32, 32, BK>(d_items, nitems, max_m, max_n, s);
            case 9: return launch_gemm_rcd<T, POST, PRE, 64, 64, 2 * BK>(d_items, nitems, max_m, max_n, s);
            case 11: return launch_gemm_rcd<T, POST, PRE, 32, 32, BK, 2>(d_items, nitems, max_m, max_n, s);
            case 12: return launch_gemm_rcd<T, POST, PRE, 32, 32, 2 * BK, 2>(d_items, nitems, max_m, max_n, s);
            case 13: return launch_gemm_rcd<T, POST, PRE, 32, 32, BK, 4>(d_items, nitems, max_m, max_n, s);
            case 14: return launch_gemm_rcd<T, POST, PRE, 32, 32, 2 * BK>(d_items, nitems, max_m, max_n, s);
            case 15: return launch_gemm_rcd<T, POST, PRE, 96, 32, BK>(d_items, nitems, max_m, max_n, s);
            case 16: return launch_gemm_rcd<T, POST, PRE, 96, 32, BK, 2>(d_items, nitems, max_m, max_n, s);
            case 17: return launch_gemm_rcd<T, POST, PRE, 96, 32, 2 * BK, 2>(d_items, nitems, max_m, max_n, s);
            case 20: return launch_gemm_rcd<T, POST, PRE, 96, 96, BK, 2, 2, 6>(d_items, nitems, max_m, max_n, s);
            case 21: return launch_gemm_rcd<T, POST, PRE, 96, 96, BK, 3, 2, 6>(d_items, nitems, max_m, max_n, s);
            case 22: return launch_gemm_rcd<T, POST, PRE, 96, 48, BK, 2, 2, 3>(d_items, nitems, max_m, max_n, s);
            case 23: return launch_gemm_rcd<T, POST, PRE, 96, 96, BK, 2, 2, 3>(d_items, nitems, max_m, max_n, s);
            case 24: return launch_gemm_rcd<T, POST, PRE, 96, 96, BK, 3, 2, 3>(d_items, nitems, max_m, max_n, s);
            default: return launch_gemm_rcd<T, POST, PRE, 64, 64, BK, 4>(d_items, nitems, max_m, max_n, s);
            }
        } else {
            return false;
        }
    }
    if constexpr (POST == FUSED_DIAG) {
        return false;  // only the direct-to-LDS tiles fuse the row sums
    } else {
        switch (tile) {
        case 1: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 32, G4>(d_items, nitems, max_m, max_n, s);
        case 2: return launch_gemm_tile<T, AC, BC, PRE, POST, 32, 32, 64, G4>(d_items, nitems, max_m, max_n, s);
        case 3: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 32, G4>(d_items, nitems, max_m, max_n, s);
        case 4: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 96, 32, G4>(d_items, nitems, max_m, max_n, s);
        case 5: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 96, 16, G4>(d_items, nitems, max_m, max_n, s);
        case 18: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 32, 32, G4>(d_items, nitems, max_m, max_n, s);
        case 19: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 32, 16, G4>(d_items, nitems, max_m, max_n, s);
        case 25: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 32, 16, G4, true>(d_items, nitems, max_m, max_n, s);
        case 26: return launch_gemm_tile<T, AC, BC, PRE, POST, 96, 32, 32, G4, true>(d_items, nitems, max_m, max_n, s);
        case 27: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 16, G4, true>(d_items, nitems, max_m, max_n, s);
        default: return launch_gemm_tile<T, AC, BC, PRE, POST, 64, 64, 16, G4>(d_items, nitems, max_m, max_n, s);
        }
    }
}

}  // namespace isle
