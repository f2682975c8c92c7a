// Batched complex FP64 LU with partial pivoting, its log-determinant, and a
// right-hand-side-parallel triangular solve.
//
// Pivoting follows LAPACK zgetrf: the pivot of column k is the first row of maximal
// |re| + |im| (izamax), full rows are swapped, and piv[k] records the swapped row, so
// the factors are in LAPACK's layout (L unit lower with later interchanges applied).
// The log-determinant is sum_i log(U_ii) + i*pi*(number of k with piv[k] != k),
// with the imaginary part reduced by the IEEE remainder to [-pi, pi]
// (src/isle/cpp/math.hpp:265-293).
//
// The factorisation is right-looking in panels of LU_NB columns, two launches per panel:
//   lu_panel_kernel  (one workgroup per matrix, one matrix row per thread, held in
//                     registers): pivoted factorisation of the panel, L11^{-1} by
//                     recursive doubling, the panel's interchanges on the columns right
//                     of the panel (the new A12 rows are staged in the work buffer, the
//                     displaced rows written in place), and on the item's last panel
//                     the log-determinant;
//   lu_update_kernel (32x32 tiles, every CU): U12 = L11^{-1} A12 and
//                     A22 -= L21 U12 on FP64 MFMA, plus the panel's interchanges on the
//                     columns left of it.
//
// Besides square matrices the same kernels factorise the first kf columns of an m x ncols
// matrix whose rows [k0+NB, nt) are, in the columns not yet factorised, known to be rows of
// the identity (a block row [I  X  Y] stacked on a general block row, as in one step of the
// structured elimination of a block-cyclic matrix): such rows can neither be pivots nor
// receive an update (their multipliers are zero), so the candidates of the panel at k0 are
// its own rows k0..k0+NB-1 and the rows nt..m-1.  Panel positions p index those candidates,
// lu_row() maps them to matrix rows.
#pragma once

#include "common.hpp"

namespace isle {

constexpr int LU_NB = 32;          // panel width
constexpr int LU_MAX_N = 512;      // one panel row per thread
constexpr int LU_UPD_THREADS = 256;
constexpr int LU_TILE = 32;        // trailing-update tile
constexpr int LU_SWAP_COLS = 64;   // columns per left-interchange workgroup

struct LuItem {
    double* re;      // n x n (or m x ncols), row-major, row stride ld (factorised in place)
    double* im;
    int* piv;        // length kf (0-based row index swapped with row k)
    double* logdet;  // 2 doubles (re, im) or null: sum over the kf factorised columns
    double* work;    // lu_work_doubles(kf, ncols) doubles of scratch
    int n;           // order of a square matrix (m = ncols = kf = n when the fields below are 0)
    int ld;
    int m = 0, ncols = 0, kf = 0;  // rows, columns, columns to factorise (kf % LU_NB == 0 if nt > 0)
    int nt = 0;      // > 0: rows [k0 + LU_NB, nt) are identity rows, never candidates (see above)
    int nupd = 0;    // > 0: columns [nupd, ncols) take the row interchanges only (their
                     // elimination is left to the caller: U12 = L11^{-1} A12, A22 -= L21 U12)
};
__host__ __device__ inline int lu_m(const LuItem& it) { return it.m ? it.m : it.n; }
__host__ __device__ inline int lu_ncols(const LuItem& it) { return it.ncols ? it.ncols : it.n; }
__host__ __device__ inline int lu_kf(const LuItem& it) { return it.kf ? it.kf : it.n; }
__host__ __device__ inline int lu_nupd(const LuItem& it) { return it.nupd ? it.nupd : lu_ncols(it); }
// Candidate rows of the panel at k0 and the matrix row of panel position p.
__host__ __device__ inline bool lu_structured(int k0, int nt) { return nt > 0 && k0 + LU_NB < nt; }
__host__ __device__ inline int lu_rows(int k0, int m, int nt) {
    return lu_structured(k0, nt) ? LU_NB + m - nt : m - k0;
}
__host__ __device__ inline int lu_row(int k0, int nt, int p) {
    return (lu_structured(k0, nt) && p >= LU_NB) ? nt + p - LU_NB : k0 + p;
}

// Work layout: for every panel p, L11^{-1} and U11^{-1} of that panel ([NB][NB] re and im
// planes each; a partial last panel is padded with the identity), then the staged A12
// rows [NB][n] re and im planes, then the panel's composed interchange as ints
// {count, dst[2NB], src[2NB]}.  The diagonal-block inverses stay valid after the
// factorisation and are used by the triangular solve.
constexpr int LU_DINV = 4 * LU_NB * LU_NB;  // doubles per panel: L11^{-1}, U11^{-1} (re, im)
__host__ __device__ inline int lu_panels(int n) { return (n + LU_NB - 1) / LU_NB; }
__host__ __device__ inline size_t lu_work_doubles(int kf, int ncols) {
    return (size_t)LU_DINV * lu_panels(kf) + 2 * (size_t)LU_NB * ncols + 2 * LU_NB + 2;
}
__host__ __device__ inline size_t lu_work_doubles(int n) { return lu_work_doubles(n, n); }
__device__ inline double* lu_linv(const LuItem& it, int panel) { return it.work + (size_t)LU_DINV * panel; }
__device__ inline double* lu_uinv(const LuItem& it, int panel) {
    return it.work + (size_t)LU_DINV * panel + 2 * LU_NB * LU_NB;
}
__device__ inline double* lu_stage(const LuItem& it) { return it.work + (size_t)LU_DINV * lu_panels(lu_kf(it)); }
__device__ inline int* lu_perm(const LuItem& it) {
    return reinterpret_cast<int*>(lu_stage(it) + 2 * (size_t)LU_NB * lu_ncols(it));
}

__device__ __forceinline__ double cabs1(double r, double i) { return fabs(r) + fabs(i); }

// Block-wide sum over all threads (any multiple of 64 up to 1024).  s: >= 16 doubles.
__device__ inline double block_sum(double v, double* s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) s[wave] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s[w];
    __syncthreads();
    return t;
}

// Pivot key: the bits of |re|+|im| (a non-negative double, so its bit pattern orders like
// its value) with the low 9 mantissa bits replaced by 511 - position.  The maximal key is
// the largest candidate, ties (and values equal to 2^-43 relative) going to the first
// position, as izamax.  Positions must be < 512.
__device__ __forceinline__ unsigned long long pivot_key(double v, int pos) {
    return (static_cast<unsigned long long>(__double_as_longlong(v)) & ~0x1FFull) |
           static_cast<unsigned long long>(511 - pos);
}

// One DPP step on a 64-bit value: both halves move with the same lane pattern, so the
// result is always some lane's whole value (or 0 where the pattern has no source lane).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, ROW_MASK, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, ROW_MASK, 0xf, true);
    return ((unsigned long long)hi << 32) | lo;
}

// Maximum of a non-negative 64-bit key over the 64 lanes of a wave, in registers: prefix
// maxima within each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's maximum into row 1
// and row 2's into row 3 (row_bcast:15), then lane 31's into rows 2 and 3 (row_bcast:31);
// lane 63 ends with the wave's maximum, which every lane reads.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
    unsigned long long o;
    o = dpp_u64<0x111, 0xf>(k); k = o > k ? o : k;
    o = dpp_u64<0x112, 0xf>(k); k = o > k ? o : k;
    o = dpp_u64<0x114, 0xf>(k); k = o > k ? o : k;
    o = dpp_u64<0x118, 0xf>(k); k = o > k ? o : k;
    o = dpp_u64<0x142, 0xa>(k); k = o > k ? o : k;
    o = dpp_u64<0x143, 0xc>(k); k = o > k ? o : k;
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)k, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(k >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// L11^{-1} and U11^{-1} of a factorised panel by recursive doubling (lt: L11 unit lower,
// ut: U11 with its diagonal already inverted, both [NB][NB + 1] in LDS, tl / tu temporaries),
// then written to the panel's slots of the work buffer.  Ends with a barrier-free copy; the
// caller's threads must all take part (block-wide loops and barriers inside).
//   [A 0; C D]^{-1} = [A^-1 0; -D^-1 C A^-1  D^-1],  [A B; 0 D]^{-1} = [A^-1 -A^-1 B D^-1; 0 D^-1]
template <int NB>
__device__ void lu_invert_panel_blocks(const LuItem& it, int panel, double (*lt_re)[NB + 1], double (*lt_im)[NB + 1],
                                       double (*ut_re)[NB + 1], double (*ut_im)[NB + 1], double (*tl_re)[NB + 1],
                                       double (*tl_im)[NB + 1], double (*tu_re)[NB + 1], double (*tu_im)[NB + 1]) {
    const int tid = threadIdx.x, nth = blockDim.x;
    // Each level's dot products run over the whole block length, the terms outside the
    // triangle replaced by zero (fixed trip counts: the loads of a product are issued together
    // instead of one loop iteration at a time; every index stays inside the NB x NB blocks).
#pragma unroll
    for (int b = 1; b < NB; b *= 2) {
        const int work = (NB / (2 * b)) * b * b;
        for (int e = tid; e < work; e += nth) {
            const int blk = e / (b * b), i = (e / b) % b, j = e % b, o = blk * 2 * b;
            double tr = 0, ti = 0, sr = 0, si = 0;
#pragma unroll 8
            for (int q = 0; q < b; ++q) {
                {  // C A^-1: terms q >= j
                    const double cr = lt_re[o + b + i][o + q], ci = lt_im[o + b + i][o + q];
                    const double xr = lt_re[o + q][o + j], xi = lt_im[o + q][o + j];
                    tr += q >= j ? cr * xr - ci * xi : 0.0;
                    ti += q >= j ? cr * xi + ci * xr : 0.0;
                }
                {  // B D^-1: terms q <= j
                    const double br = ut_re[o + i][o + b + q], bi = ut_im[o + i][o + b + q];
                    const double xr = ut_re[o + b + q][o + b + j], xi = ut_im[o + b + q][o + b + j];
                    sr += q <= j ? br * xr - bi * xi : 0.0;
                    si += q <= j ? br * xi + bi * xr : 0.0;
                }
            }
            tl_re[o + b + i][o + j] = tr;
            tl_im[o + b + i][o + j] = ti;
            tu_re[o + i][o + b + j] = sr;
            tu_im[o + i][o + b + j] = si;
        }
        __syncthreads();
        for (int e = tid; e < work; e += nth) {
            const int blk = e / (b * b), i = (e / b) % b, j = e % b, o = blk * 2 * b;
            double xr_ = 0, xi_ = 0, yr = 0, yi = 0;
#pragma unroll 8
            for (int q = 0; q < b; ++q) {
                {  // D^-1 (C A^-1): terms q <= i
                    const double dr = lt_re[o + b + i][o + b + q], di = lt_im[o + b + i][o + b + q];
                    const double tr = tl_re[o + b + q][o + j], ti = tl_im[o + b + q][o + j];
                    xr_ += q <= i ? dr * tr - di * ti : 0.0;
                    xi_ += q <= i ? dr * ti + di * tr : 0.0;
                }
                {  // A^-1 (B D^-1): terms q >= i
                    const double ar_ = ut_re[o + i][o + q], ai_ = ut_im[o + i][o + q];
                    const double tr = tu_re[o + q][o + b + j], ti = tu_im[o + q][o + b + j];
                    yr += q >= i ? ar_ * tr - ai_ * ti : 0.0;
                    yi += q >= i ? ar_ * ti + ai_ * tr : 0.0;
                }
            }
            lt_re[o + b + i][o + j] = -xr_;
            lt_im[o + b + i][o + j] = -xi_;
            ut_re[o + i][o + b + j] = -yr;
            ut_im[o + i][o + b + j] = -yi;
        }
        __syncthreads();
    }
    double* li_re = lu_linv(it, panel);
    double* li_im = li_re + NB * NB;
    double* ui_re = lu_uinv(it, panel);
    double* ui_im = ui_re + NB * NB;
    for (int e = tid; e < NB * NB; e += nth) {
        li_re[e] = lt_re[e / NB][e % NB];
        li_im[e] = lt_im[e / NB][e % NB];
        ui_re[e] = ut_re[e / NB][e % NB];
        ui_im[e] = ut_im[e / NB][e % NB];
    }
}

// On the item's last panel (k0 + NB >= kf), where every U_ii and pivot of the kf columns is
// final: log det = sum log U_ii + i pi (number of interchanges), into it.logdet.  rsum: >= 16
// doubles of LDS.
__device__ inline void lu_panel_logdet(const LuItem& it, int k0, int nb, int kf, double* rsum) {
    if (!it.logdet || k0 + nb < kf) return;
    __syncthreads();
    const int tid = threadIdx.x, nth = blockDim.x, ld = it.ld;
    double lr_ = 0, la = 0, pc = 0;
    for (int i = tid; i < kf; i += nth) {
        const double dr = it.re[(long)i * ld + i], di = it.im[(long)i * ld + i];
        lr_ += log(hypot(dr, di));
        la += atan2(di, dr);
        pc += (it.piv[i] != i) ? 1.0 : 0.0;
    }
    lr_ = block_sum(lr_, rsum);
    la = block_sum(la, rsum);
    pc = block_sum(pc, rsum);
    if (tid == 0) {
        if (((long)pc) & 1) la += M_PI;
        it.logdet[0] = lr_;
        it.logdet[1] = remainder(la, 2.0 * M_PI);
    }
}

template <int NB>
__global__ __launch_bounds__(LU_MAX_N) void lu_panel_kernel(const LuItem* __restrict__ items, int k0,
                                                            int add_identity) {
    __shared__ unsigned long long pkey[NB];
    __shared__ int pvt[NB];  // pivot position of each column
    // the pivot row (re, im pairs); a column's writes and reads are separated by that
    // column's barriers
    __shared__ __align__(16) double xrow[2 * NB];
    __shared__ double lt_re[NB][NB + 1], lt_im[NB][NB + 1];  // L11, inverted in place
    __shared__ double ut_re[NB][NB + 1], ut_im[NB][NB + 1];  // U11, inverted in place
    __shared__ double tl_re[NB][NB + 1], tl_im[NB][NB + 1];  // doubling temporaries
    __shared__ double tu_re[NB][NB + 1], tu_im[NB][NB + 1];
    __shared__ int slot_cnt, slot_dst[2 * NB], slot_src[2 * NB];
    __shared__ double rsum[LU_MAX_N / 64];

    const LuItem it = items[blockIdx.x];
    const int n = it.n, m = lu_m(it), ncols = lu_ncols(it), kf = lu_kf(it), ntop = it.nt;
    if (k0 >= kf) return;
    const int ld = it.ld, tid = threadIdx.x, nth = blockDim.x;
    const int rows = lu_rows(k0, m, ntop), kb = min(NB, kf - k0);
    const int panel = k0 / NB;
    double* re = it.re;
    double* im = it.im;
    const bool own = tid < rows;
    if (tid == 0) slot_cnt = 0;
    if (tid < NB) pkey[tid] = 0ull;
    if (add_identity) {
        for (int i = NB + tid; i < n; i += nth) re[(long)i * ld + i] += 1.0;
    }

    double ar[NB], ai[NB];
    int pos = tid;  // current position of this thread's original panel row
    // full panels of 16-byte aligned rows move in 16-byte pieces (uniform over the workgroup)
    const bool vec = kb == NB && ld % 2 == 0 && ((reinterpret_cast<uintptr_t>(re) | reinterpret_cast<uintptr_t>(im)) & 15) == 0;
    if (vec) {
        const long o = own ? (long)lu_row(k0, ntop, tid) * ld + k0 : 0;
#pragma unroll
        for (int c = 0; c < NB; c += 2) {
            const double2 vr = own ? *reinterpret_cast<const double2*>(re + o + c) : make_double2(0.0, 0.0);
            const double2 vi = own ? *reinterpret_cast<const double2*>(im + o + c) : make_double2(0.0, 0.0);
            ar[c] = vr.x;
            ar[c + 1] = vr.y;
            ai[c] = vi.x;
            ai[c + 1] = vi.y;
        }
        if (add_identity && own && tid < NB) {
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c == tid) ar[c] += 1.0;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            double vr = 0, vi = 0;
            if (own && c < kb) {
                vr = re[(long)lu_row(k0, ntop, tid) * ld + k0 + c];
                vi = im[(long)lu_row(k0, ntop, tid) * ld + k0 + c];
                if (add_identity && c == tid) vr += 1.0;
            }
            ar[c] = vr;
            ai[c] = vi;
        }
    }

    __syncthreads();
    // Column loop.  Rows never move between threads: thread t keeps original panel row t in
    // registers and `pos` is its current position (LAPACK's interchange of rows KK and p is
    // the exchange of two positions).  Columns go in groups of LU_CG: in the group starting at
    // column K2 slot Q (0 .. LU_CG-1) holds column K2 + Q, and after the group the register
    // row is rotated by LU_CG slots, the freed slots taking their columns' multipliers (rows
    // still to be pivoted) or departing values (rows already pivoted).  After NB / LU_CG
    // rotations every row is back in column order, L multipliers then U entries.  This keeps
    // all register indices static inside a real loop (a fully unrolled column loop does not
    // fit the instruction cache) at 1 / LU_CG of the register moves of a rotation per column.
    constexpr int LU_CG = 8;
    static_assert(NB % LU_CG == 0, "columns go in groups");
    for (int k2 = 0; k2 < NB; k2 += LU_CG) {
        const int live = NB - k2;  // slots Q+1 .. live-1 hold the columns right of column K2 + Q
#pragma unroll
        for (int Q = 0; Q < LU_CG; ++Q) {
            const int kk = k2 + Q;
            double nr = ar[Q], ni = ai[Q];  // stays in slot Q: the multiplier or the departing value
            if (kk < kb) {
                {  // max over the wave first (lanes to one LDS address serialise), then one atomic per wave
                    const unsigned long long key =
                        wave_max_u64((own && pos >= kk) ? pivot_key(cabs1(ar[Q], ai[Q]), pos) : 0ull);
                    if ((tid & 63) == 0 && key != 0ull) atomicMax(&pkey[kk], key);
                }
                __syncthreads();
                // pivot position; an exactly zero column (no nonzero key) keeps row kk in place
                const unsigned long long kmax = pkey[kk];
                const int p = kmax ? 511 - static_cast<int>(kmax & 0x1FFull) : kk;
                if (own && pos == p) {  // the raw register row (see the update for the other slots)
                    double2* dst = reinterpret_cast<double2*>(xrow);
#pragma unroll
                    for (int j = 0; j < NB; ++j) dst[j] = make_double2(ar[j], ai[j]);
                }
                if (own) {  // the interchange of positions kk and p
                    if (pos == p)
                        pos = kk;
                    else if (pos == kk)
                        pos = p;
                }
                if (tid == 0) pvt[kk] = p;  // global stores here would stall every LDS wait (FLAT)
                __syncthreads();
                if (own && pos > kk) {  // rows not yet pivoted: multiplier and rank-1 update
                    const double2* xv = reinterpret_cast<const double2*>(xrow);
                    const double2 u0 = xv[Q];
                    const double ur = u0.x, ui = u0.y;
                    if (ur != 0.0 || ui != 0.0) {
                        const double rden = 1.0 / (ur * ur + ui * ui);
                        const double ir = ur * rden, ii = -ui * rden;
                        const double t = nr * ir - ni * ii;
                        ni = nr * ii + ni * ir;
                        nr = t;
                    }
                    // the pivot row in batches of 8 entries (a batch's reads in flight together).
                    // Slots 0..Q and live.. of the pivot row hold its own multipliers (or, for
                    // slots below Q, entries of columns already done): batches
                    // from live on are skipped and the one batch that straddles live is masked
                    // (the owner writes its row unmasked, a single lane's stores on the critical
                    // path).  Two fused multiply-adds per real part:
                    // a_j - m b_j = (a_j - m_r b_r) + m_i b_i, likewise for the imaginary part.
                    constexpr int RB = 8;
#pragma unroll
                    for (int j0 = 0; j0 < NB; j0 += RB) {
                        if (j0 < live) {
                            double2 bv[RB];
#pragma unroll
                            for (int q = 0; q < RB; ++q) bv[q] = xv[j0 + q];
                            if (j0 + RB <= live) {
#pragma unroll
                                for (int q = 0; q < RB; ++q) {
                                    const int j = j0 + q;
                                    if (j > Q) {
                                        ar[j] = fma(ni, bv[q].y, fma(-nr, bv[q].x, ar[j]));
                                        ai[j] = fma(-ni, bv[q].x, fma(-nr, bv[q].y, ai[j]));
                                    }
                                }
                            } else {
#pragma unroll
                                for (int q = 0; q < RB; ++q) {
                                    const int j = j0 + q;
                                    if (j > Q && j < live) {
                                        ar[j] = fma(ni, bv[q].y, fma(-nr, bv[q].x, ar[j]));
                                        ai[j] = fma(-ni, bv[q].x, fma(-nr, bv[q].y, ai[j]));
                                    }
                                }
                            }
                        }
                    }
                }
            }
            ar[Q] = nr;
            ai[Q] = ni;
        }
        double hr[LU_CG], hi[LU_CG];
#pragma unroll
        for (int q = 0; q < LU_CG; ++q) {
            hr[q] = ar[q];
            hi[q] = ai[q];
        }
#pragma unroll
        for (int j = 0; j < NB - LU_CG; ++j) {
            ar[j] = ar[j + LU_CG];
            ai[j] = ai[j + LU_CG];
        }
#pragma unroll
        for (int q = 0; q < LU_CG; ++q) {
            ar[NB - LU_CG + q] = hr[q];
            ai[NB - LU_CG + q] = hi[q];
        }
    }
    if (tid < kb) it.piv[k0 + tid] = lu_row(k0, ntop, pvt[tid]);
    if (own) {
        const long orow = lu_row(k0, ntop, pos);
        if (vec) {
#pragma unroll
            for (int c = 0; c < NB; c += 2) {
                *reinterpret_cast<double2*>(re + orow * ld + k0 + c) = make_double2(ar[c], ar[c + 1]);
                *reinterpret_cast<double2*>(im + orow * ld + k0 + c) = make_double2(ai[c], ai[c + 1]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < kb) {
                    re[orow * ld + k0 + c] = ar[c];
                    im[orow * ld + k0 + c] = ai[c];
                }
            }
        }
        if (pos != tid) {  // at most 2*NB positions change: the panel rows and their partners
            const int sl = atomicAdd(&slot_cnt, 1);
            if (sl < 2 * NB) {
                slot_dst[sl] = pos;
                slot_src[sl] = tid;
            }
        }
    }
    // L11 and U11 (diagonal inverted) into LDS from the rows now at positions 0..kb-1;
    // positions kb..NB-1 of a partial last panel are padded with the identity
    if ((own && pos < NB) || (!own && tid < NB)) {
        const int r = own ? pos : tid;
        const bool in = own;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            lt_re[r][c] = (in && c < r) ? ar[c] : (c == r ? 1.0 : 0.0);
            lt_im[r][c] = (in && c < r) ? ai[c] : 0.0;
            const bool u = in && c >= r && c < kb;
            ut_re[r][c] = u ? ar[c] : (c == r ? 1.0 : 0.0);
            ut_im[r][c] = u ? ai[c] : 0.0;
        }
    }
    __syncthreads();
    if (tid < NB) {  // the diagonal of U11^{-1} is 1/u_ii (one division per row, not per column)
        const double vr = ut_re[tid][tid], vi = ut_im[tid][tid];
        const double den = vr * vr + vi * vi;
        ut_re[tid][tid] = vr / den;
        ut_im[tid][tid] = -vi / den;
    }
    const int cnt = min(slot_cnt, 2 * NB);
    int* perm = lu_perm(it);
    for (int e = tid; e < cnt; e += nth) {
        perm[1 + e] = slot_dst[e];
        perm[1 + 2 * NB + e] = slot_src[e];
    }
    if (tid == 0) perm[0] = cnt;
    __syncthreads();

    lu_invert_panel_blocks<NB>(it, panel, lt_re, lt_im, ut_re, ut_im, tl_re, tl_im, tu_re, tu_im);
    lu_panel_logdet(it, k0, NB, kf, rsum);
}

// Panels with more candidate rows than one workgroup holds in registers (lu_rows > LU_MAX_N,
// up to LU_TALL_MAX): the same factorisation with the panel in global memory (in place; an
// n x NB panel is a few MB, resident in L2), each thread taking every blockDim-th candidate
// row.  Per column: pivot search (keys as in pivot_key, 14 position bits), the pivot row
// swapped with row kk across the panel's columns and copied to LDS, then the multipliers
// and the rank-1 update of the rows below.  The interchange record, L11^{-1} / U11^{-1} and
// the log-determinant are those of lu_panel_kernel.
constexpr int LU_TALL_MAX = 16383;
constexpr int LU_TALL_THREADS = 1024;

__device__ __forceinline__ unsigned long long pivot_key_tall(double v, int pos) {
    return (static_cast<unsigned long long>(__double_as_longlong(v)) & ~0x3FFFull) |
           static_cast<unsigned long long>(0x3FFF - pos);
}

template <int NB>
__global__ __launch_bounds__(LU_TALL_THREADS) void lu_panel_tall_kernel(const LuItem* __restrict__ items, int k0,
                                                                        int add_identity) {
    __shared__ unsigned long long wkey[LU_TALL_THREADS / 64];
    __shared__ int pvt[NB];
    __shared__ double xrow[2 * NB];  // the pivot row (re, im interleaved)
    __shared__ double lt_re[NB][NB + 1], lt_im[NB][NB + 1];
    __shared__ double ut_re[NB][NB + 1], ut_im[NB][NB + 1];
    __shared__ double tl_re[NB][NB + 1], tl_im[NB][NB + 1];
    __shared__ double tu_re[NB][NB + 1], tu_im[NB][NB + 1];
    __shared__ double rsum[LU_TALL_THREADS / 64];

    const LuItem it = items[blockIdx.x];
    const int n = it.n, m = lu_m(it), kf = lu_kf(it), ntop = it.nt;
    if (k0 >= kf) return;
    const int ld = it.ld, tid = threadIdx.x, nth = blockDim.x;
    const int rows = lu_rows(k0, m, ntop), kb = min(NB, kf - k0);
    const int panel = k0 / NB;
    double* re = it.re;
    double* im = it.im;
    if (add_identity)
        for (int i = tid; i < n; i += nth) re[(long)i * ld + i] += 1.0;
    __syncthreads();

    for (int kk = 0; kk < kb; ++kk) {
        unsigned long long key = 0ull;
        for (int p = kk + tid; p < rows; p += nth) {
            const long o = (long)lu_row(k0, ntop, p) * ld + k0 + kk;
            const unsigned long long k2 = pivot_key_tall(cabs1(re[o], im[o]), p);
            key = k2 > key ? k2 : key;
        }
        key = wave_max_u64(key);
        if ((tid & 63) == 0) wkey[tid >> 6] = key;
        __syncthreads();
        unsigned long long kmax = 0ull;
        for (int w = 0; w < (nth >> 6); ++w) kmax = wkey[w] > kmax ? wkey[w] : kmax;
        // an exactly zero column keeps row kk in place
        const int p = kmax ? 0x3FFF - static_cast<int>(kmax & 0x3FFFull) : kk;
        if (tid < kb) {  // exchange positions kk and p across the panel's columns
            const long okk = (long)lu_row(k0, ntop, kk) * ld + k0 + tid;
            const long op = (long)lu_row(k0, ntop, p) * ld + k0 + tid;
            const double ar = re[okk], ai = im[okk];
            const double br = re[op], bi = im[op];
            if (p != kk) {
                re[okk] = br;
                im[okk] = bi;
                re[op] = ar;
                im[op] = ai;
            }
            xrow[2 * tid] = br;
            xrow[2 * tid + 1] = bi;
        }
        if (tid == 0) pvt[kk] = p;
        __syncthreads();
        const double ur = xrow[2 * kk], ui = xrow[2 * kk + 1];
        const bool nz = ur != 0.0 || ui != 0.0;
        double ir = 0.0, ii = 0.0;
        if (nz) {
            const double rden = 1.0 / (ur * ur + ui * ui);
            ir = ur * rden;
            ii = -ui * rden;
        }
        for (int q = kk + 1 + tid; q < rows; q += nth) {
            const long ob = (long)lu_row(k0, ntop, q) * ld + k0;
            double lr = re[ob + kk], li = im[ob + kk];
            if (nz) {
                const double t = lr * ir - li * ii;
                li = lr * ii + li * ir;
                lr = t;
                re[ob + kk] = lr;
                im[ob + kk] = li;
            }
            for (int j = kk + 1; j < kb; ++j) {
                const double br = xrow[2 * j], bi = xrow[2 * j + 1];
                re[ob + j] -= lr * br - li * bi;
                im[ob + j] -= lr * bi + li * br;
            }
        }
        __syncthreads();
    }
    if (tid < kb) it.piv[k0 + tid] = lu_row(k0, ntop, pvt[tid]);
    if (tid == 0) {
        // composed interchange of the panel: (final position, original position) of every
        // row that moved, replayed from the pivot sequence on at most 2 NB positions
        int pos[2 * NB], org[2 * NB], cnt = 0;
        for (int kk = 0; kk < kb; ++kk) {
            const int p = pvt[kk];
            if (p == kk) continue;
            int a = -1, b = -1;
            for (int e = 0; e < cnt; ++e) {
                if (pos[e] == kk) a = e;
                if (pos[e] == p) b = e;
            }
            if (a < 0) {
                pos[cnt] = kk;
                org[cnt] = kk;
                a = cnt++;
            }
            if (b < 0) {
                pos[cnt] = p;
                org[cnt] = p;
                b = cnt++;
            }
            const int t = org[a];
            org[a] = org[b];
            org[b] = t;
        }
        int* perm = lu_perm(it);
        int w = 0;
        for (int e = 0; e < cnt; ++e)
            if (org[e] != pos[e]) {
                perm[1 + w] = pos[e];
                perm[1 + 2 * NB + w] = org[e];
                ++w;
            }
        perm[0] = w;
    }
    // L11 and U11 (diagonal inverted) from the rows now at positions 0..NB-1; positions
    // kb..NB-1 of a partial last panel are padded with the identity
    if (tid < NB) {
        const int r = tid;
        const bool in = r < kb;
        const long ob = (long)lu_row(k0, ntop, r) * ld + k0;
#pragma unroll 4
        for (int c = 0; c < NB; ++c) {
            const bool cin = in && c < kb;
            const double ar = cin ? re[ob + c] : 0.0, ai = cin ? im[ob + c] : 0.0;
            lt_re[r][c] = (cin && c < r) ? ar : (c == r ? 1.0 : 0.0);
            lt_im[r][c] = (cin && c < r) ? ai : 0.0;
            double vr = (cin && c >= r) ? ar : (c == r ? 1.0 : 0.0), vi = (cin && c >= r) ? ai : 0.0;
            if (c == r) {  // the diagonal of U11^{-1} is 1/u_ii
                const double den = vr * vr + vi * vi;
                vr = vr / den;
                vi = -vi / den;
            }
            ut_re[r][c] = vr;
            ut_im[r][c] = vi;
        }
    }
    __syncthreads();
    lu_invert_panel_blocks<NB>(it, panel, lt_re, lt_im, ut_re, ut_im, tl_re, tl_im, tu_re, tu_im);
    lu_panel_logdet(it, k0, NB, kf, rsum);
}

// The panel's interchanges on the columns right of it, on every CU: the new rows
// k0..k0+NB-1 (A12) go to the stage (position q reads the original row of the row now at q);
// then the rows displaced below the panel take their new content, which is always an
// original panel row (rows k0..k0+NB-1 are not written here).  A staged read of a row
// below the panel precedes the displaced write to that row: both are in the same column,
// and every column belongs to one workgroup (64 columns, 4 threads per column, a barrier
// between the two phases).  grid.x: column blocks; grid.y: item.
__global__ __launch_bounds__(256) void lu_rswap_kernel(const LuItem* __restrict__ items, int k0) {
    constexpr int NB = LU_NB, CW = 64, PARTS = 256 / CW, QP = NB / PARTS;
    const LuItem it = items[blockIdx.y];
    const int kf = lu_kf(it), ncols = lu_ncols(it), ntop = it.nt, ld = it.ld;
    if (k0 >= kf) return;
    const int nrt = ncols - k0 - NB;
    const int c0 = blockIdx.x * CW;
    if (c0 >= nrt) return;
    __shared__ int orig[NB], dst[2 * NB], src[2 * NB];
    __shared__ int cnt_s;
    const int* perm = lu_perm(it);
    const int tid = threadIdx.x;
    if (tid == 0) cnt_s = min(perm[0], 2 * NB);
    if (tid < NB) orig[tid] = tid;
    __syncthreads();
    const int cnt = cnt_s;
    if (tid < cnt) {
        dst[tid] = perm[1 + tid];
        src[tid] = perm[1 + 2 * NB + tid];
    }
    __syncthreads();
    if (tid < cnt && dst[tid] < NB) orig[dst[tid]] = src[tid];
    __syncthreads();
    double* re = it.re;
    double* im = it.im;
    const int c = c0 + (tid % CW), part = tid / CW;
    const bool ok = c < nrt;
    const long col = k0 + NB + c;
    const int prow = min(NB, lu_rows(k0, lu_m(it), ntop));  // positions that hold rows of the matrix
    double* st_re = lu_stage(it);
    double* st_im = st_re + (size_t)NB * ncols;
    {
        double vr[QP], vi[QP];
#pragma unroll
        for (int j = 0; j < QP; ++j) {
            const bool in = ok && part + PARTS * j < prow;
            const long o = in ? (long)lu_row(k0, ntop, orig[part + PARTS * j]) * ld + col : 0;
            vr[j] = in ? re[o] : 0.0;
            vi[j] = in ? im[o] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < QP; ++j)
            if (ok) {
                st_re[(long)(part + PARTS * j) * ncols + c] = vr[j];
                st_im[(long)(part + PARTS * j) * ncols + c] = vi[j];
            }
    }
    __syncthreads();
    constexpr int EP = 2 * NB / PARTS;  // displaced entries per thread, at most
    double wr[EP], wi[EP];
#pragma unroll
    for (int j = 0; j < EP; ++j) {
        const int e = part + PARTS * j;
        if (ok && e < cnt && dst[e] >= NB) {
            const long o = (long)lu_row(k0, ntop, src[e]) * ld + col;
            wr[j] = re[o];
            wi[j] = im[o];
        }
    }
#pragma unroll
    for (int j = 0; j < EP; ++j) {
        const int e = part + PARTS * j;
        if (ok && e < cnt && dst[e] >= NB) {
            const long o = (long)lu_row(k0, ntop, dst[e]) * ld + col;
            re[o] = wr[j];
            im[o] = wi[j];
        }
    }
    // columns left to the caller's elimination (col >= nupd): the panel rows take their new
    // content as it is (the reads of phase two, which may read panel rows, are done first)
    if (nrt > 0 && k0 + NB + c0 + CW > lu_nupd(it)) {  // uniform over the workgroup
        __syncthreads();
        if (ok && col >= lu_nupd(it)) {
#pragma unroll
            for (int j = 0; j < QP; ++j) {
                const int q = part + PARTS * j;
                if (q < prow) {
                    re[(long)(k0 + q) * ld + col] = st_re[(long)q * ncols + c];
                    im[(long)(k0 + q) * ld + col] = st_im[(long)q * ncols + c];
                }
            }
        }
    }
}

// grid.x: tm*tm trailing tiles (tm = tiles per dimension of the largest item's A22), then
// swap_wgs workgroups of LU_SWAP_COLS left columns each; grid.y: item.
__global__ __launch_bounds__(LU_UPD_THREADS) void lu_update_kernel(const LuItem* __restrict__ items, int k0,
                                                                    int tmr, int tmc) {
    constexpr int T = LU_TILE, NB = LU_NB;
    static_assert(T == 32 && NB == 32, "tile code assumes 32");
    const LuItem it = items[blockIdx.y];
    const int m = lu_m(it), ncols = lu_ncols(it), kf = lu_kf(it), ntop = it.nt;
    const int ld = it.ld, tid = threadIdx.x;
    if (k0 >= kf) return;
    double* re = it.re;
    double* im = it.im;
    if ((int)blockIdx.x >= tmr * tmc) {
        // ---- interchanges of this panel on LU_SWAP_COLS columns left of it
        const int c0 = ((int)blockIdx.x - tmr * tmc) * LU_SWAP_COLS;
        if (c0 >= k0) return;
        const int* perm = lu_perm(it);
        const int cnt = perm[0];
        constexpr int PER = 2 * NB * LU_SWAP_COLS / LU_UPD_THREADS;
        double vr[PER], vi[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + j * LU_UPD_THREADS, q = e / LU_SWAP_COLS, c = c0 + e % LU_SWAP_COLS;
            if (q < cnt && c < k0) {
                const long o = (long)lu_row(k0, ntop, perm[1 + 2 * NB + q]) * ld + c;
                vr[j] = re[o];
                vi[j] = im[o];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + j * LU_UPD_THREADS, q = e / LU_SWAP_COLS, c = c0 + e % LU_SWAP_COLS;
            if (q < cnt && c < k0) {
                const long o = (long)lu_row(k0, ntop, perm[1 + q]) * ld + c;
                re[o] = vr[j];
                im[o] = vi[j];
            }
        }
        return;
    }
    // ---- trailing tile (I, J) of A22 = panel positions NB.. x columns c0.. of the matrix
    const int c0 = k0 + NB, ntr = lu_rows(k0, m, ntop) - NB, ntc = lu_nupd(it) - c0;
    const int I = blockIdx.x / tmc, J = blockIdx.x % tmc;
    // tile row 0 always runs when there are columns to update: it writes U12, also for a
    // last panel with no rows below it (ntr <= 0: a matrix with columns beyond its kf)
    if ((I > 0 && I * T >= ntr) || J * T >= ntc) return;
    __shared__ double sa_re[T][T + 1], sa_im[T][T + 1];  // L11^{-1}
    __shared__ double sl_re[T][T + 1], sl_im[T][T + 1];  // L21 tile
    __shared__ double sb_re[T][T + 1], sb_im[T][T + 1];  // staged A12 tile, then U12 tile
    const double* li_re = lu_linv(it, k0 / NB);
    const double* li_im = li_re + NB * NB;
    const double* st_re = lu_stage(it);
    const double* st_im = st_re + (size_t)NB * ncols;
    for (int e = tid; e < T * T; e += LU_UPD_THREADS) {
        const int r = e / T, c = e % T;
        sa_re[r][c] = li_re[e];
        sa_im[r][c] = li_im[e];
        const int col = J * T + c;
        sb_re[r][c] = col < ntc ? st_re[(long)r * ncols + col] : 0.0;
        sb_im[r][c] = col < ntc ? st_im[(long)r * ncols + col] : 0.0;
        const int row = I * T + r;
        const long mrow = lu_row(k0, ntop, NB + row);
        sl_re[r][c] = row < ntr ? re[mrow * ld + k0 + c] : 0.0;
        sl_im[r][c] = row < ntr ? im[mrow * ld + k0 + c] : 0.0;
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
    const int li = lane & 15, lk = lane >> 4;
    // prefetch this lane's C entries: rows wm + lk + 4r, column wn + li
    double cr[4], ci[4];
    const int ccol = J * T + wn + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int crow = I * T + wm + lk + 4 * r;
        cr[r] = ci[r] = 0.0;
        if (crow < ntr && ccol < ntc) {
            const long o = (long)lu_row(k0, ntop, NB + crow) * ld + c0 + ccol;
            cr[r] = re[o];
            ci[r] = im[o];
        }
    }
    __syncthreads();
    // U12 tile = L11^{-1} A12 tile
    dbl4 rr = {0, 0, 0, 0}, ii = rr, ri = rr, ir = rr;
#pragma unroll
    for (int ks = 0; ks < NB; ks += 4) {
        const double a_r = sa_re[wm + li][ks + lk], a_i = sa_im[wm + li][ks + lk];
        const double b_r = sb_re[ks + lk][wn + li], b_i = sb_im[ks + lk][wn + li];
        rr = mfma16(a_r, b_r, rr);
        ii = mfma16(a_i, b_i, ii);
        ri = mfma16(a_r, b_i, ri);
        ir = mfma16(a_i, b_r, ir);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        sb_re[wm + lk + 4 * r][wn + li] = rr[r] - ii[r];
        sb_im[wm + lk + 4 * r][wn + li] = ri[r] + ir[r];
    }
    __syncthreads();
    if (I == 0) {  // U12 into rows k0.. of the matrix (read by no other workgroup of this launch)
        const int prow = NB + min(ntr, 0);  // the panel's rows (fewer than NB in a short last panel)
        for (int e = tid; e < T * T; e += LU_UPD_THREADS) {
            const int r = e / T, c = e % T, col = J * T + c;
            if (col < ntc && r < prow) {
                re[(long)(k0 + r) * ld + c0 + col] = sb_re[r][c];
                im[(long)(k0 + r) * ld + c0 + col] = sb_im[r][c];
            }
        }
    }
    // A22 tile -= L21 tile * U12 tile
    rr = dbl4{0, 0, 0, 0};
    ii = rr;
    ri = rr;
    ir = rr;
#pragma unroll
    for (int ks = 0; ks < NB; ks += 4) {
        const double a_r = sl_re[wm + li][ks + lk], a_i = sl_im[wm + li][ks + lk];
        const double b_r = sb_re[ks + lk][wn + li], b_i = sb_im[ks + lk][wn + li];
        rr = mfma16(a_r, b_r, rr);
        ii = mfma16(a_i, b_i, ii);
        ri = mfma16(a_r, b_i, ri);
        ir = mfma16(a_i, b_r, ir);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int crow = I * T + wm + lk + 4 * r;
        if (crow < ntr && ccol < ntc) {
            const long o = (long)lu_row(k0, ntop, NB + crow) * ld + c0 + ccol;
            re[o] = cr[r] - (rr[r] - ii[r]);
            im[o] = ci[r] - (ri[r] + ir[r]);
        }
    }
}

// Factorise the first kf columns of nitems m x ncols matrices of one shape (identity rows
// above nt as described at the top; nt = 0 for a general matrix).  Panels with at most
// LU_MAX_N candidate rows run in registers (lu_panel_kernel), larger ones up to LU_TALL_MAX
// in global memory (lu_panel_tall_kernel).
inline void launch_lu_shape(const LuItem* d_items, int nitems, int m, int ncols, int kf, int nt, int add_identity,
                            hipStream_t s, int nupd = 0) {
    if (kf > ncols || kf > m || (nt > 0 && (kf % LU_NB != 0 || nt > m)))
        throw std::invalid_argument("inconsistent LU shape");
    for (int k0 = 0; k0 < kf; k0 += LU_NB)
        if (lu_rows(k0, m, nt) > LU_TALL_MAX) throw std::invalid_argument("matrix too large for the LU kernel");
    for (int k0 = 0; k0 < kf; k0 += LU_NB) {
        const int rows = lu_rows(k0, m, nt);
        if (rows > LU_MAX_N) {  // the panel in global memory
            hipLaunchKernelGGL(lu_panel_tall_kernel<LU_NB>, dim3(nitems), dim3(LU_TALL_THREADS), 0, s, d_items, k0,
                               k0 == 0 ? add_identity : 0);
        } else {
            // at least 4 waves: the L11 / U11 inversion after the column loop uses every thread
            const int threads = std::max(256, ((rows + 63) / 64) * 64);
            hipLaunchKernelGGL(lu_panel_kernel<LU_NB>, dim3(nitems), dim3(threads), 0, s, d_items, k0,
                               k0 == 0 ? add_identity : 0);
        }
        ISLE_HIP(hipGetLastError());
        const int ntr = rows - LU_NB, nsw = ncols - k0 - LU_NB;  // columns right of the panel
        const int ntc = (nupd > 0 ? std::min(nupd, ncols) : ncols) - k0 - LU_NB;  // of them updated
        if (nsw > 0) {
            hipLaunchKernelGGL(lu_rswap_kernel, dim3((nsw + 63) / 64, nitems), dim3(256), 0, s, d_items, k0);
            ISLE_HIP(hipGetLastError());
        }
        const int tmr = ntc > 0 ? std::max(1, (ntr + LU_TILE - 1) / LU_TILE) : 0;
        const int tmc = ntc > 0 ? (ntc + LU_TILE - 1) / LU_TILE : 0;
        const int swap_wgs = (k0 + LU_SWAP_COLS - 1) / LU_SWAP_COLS;
        if (tmr * tmc + swap_wgs > 0) {
            hipLaunchKernelGGL(lu_update_kernel, dim3(tmr * tmc + swap_wgs, nitems), dim3(LU_UPD_THREADS), 0, s,
                               d_items, k0, tmr, tmc);
            ISLE_HIP(hipGetLastError());
        }
    }
}

// Factorise nitems square matrices (n <= max_n <= LU_MAX_N each); add_identity: factorise 1 + A.
inline void launch_lu(const LuItem* d_items, int nitems, int max_n, int add_identity, hipStream_t s) {
    launch_lu_shape(d_items, nitems, max_n, max_n, max_n, 0, add_identity, s);
}

// X = A^{-1} B for the LU factors of A.  Element (row r, rhs c) of B is read from
// b_re[c*bsc + r*bsr] / b_im[...] and X is written to x_re/x_im with its own strides
// (so interleaved complex and transposed layouts need no copy).  If b_re is null,
// B is the identity.  16 right-hand sides per workgroup, X staged in LDS.  `mode` selects
// the parts: SOLVE_PERM (row interchanges piv), SOLVE_L (unit lower L), SOLVE_U (upper U);
// B and X may be the same array (each workgroup reads its columns before writing them).
struct SolveItem {
    const double* lu_re;
    const double* lu_im;
    const int* piv;
    const double* dinv;  // the LU's work buffer (its per-panel L11^{-1}, U11^{-1})
    const double* b_re;
    const double* b_im;
    long bsr, bsc;
    double* x_re;
    double* x_im;
    long xsr, xsc;
    int n, ld, nrhs;
    int mode = 7;
};
enum : int { SOLVE_PERM = 1, SOLVE_L = 2, SOLVE_U = 4 };

constexpr int SOLVE_COLS = 16;
constexpr int SOLVE_THREADS = 256;

// Dynamic LDS of lu_solve_kernel for matrices of order n.
inline size_t lu_solve_lds(int n) {
    const size_t np = (size_t)((n + 15) / 16) * 16;
    return 2 * np * SOLVE_COLS * sizeof(double) + 2 * (size_t)n * sizeof(int);
}

// Blocked forward/back substitution on 16 right-hand sides per workgroup.
// Per block of 16 rows: the product with the already solved rows runs on FP64 MFMA (the
// K range split over the 4 waves, partial sums reduced through LDS), then the block is
// multiplied by the inverse of its 16x16 diagonal block of L (or U), which is a diagonal
// sub-block of the panel inverses the LU stored.  GX = false: X staged in LDS (orders up to
// LU_MAX_N), B and X may be the same array.  GX = true: X kept in the output array itself
// (larger orders; only the row permutation in LDS), so B must be null or another array.
template <bool GX>
__global__ __launch_bounds__(SOLVE_THREADS) void lu_solve_kernel(const SolveItem* __restrict__ items) {
    extern __shared__ double lds[];
    __shared__ double part_r[4][16][17], part_i[4][16][17];
    __shared__ double dg_r[16][17], dg_i[16][17];  // inverse of the diagonal block
    __shared__ double y_r[16][17], y_i[16][17];    // right-hand side of the diagonal block
    const SolveItem it = items[blockIdx.y];
    const int n = it.n, ld = it.ld;
    const int c0 = blockIdx.x * SOLVE_COLS;
    if (c0 >= it.nrhs) return;
    const int nb = (n + 15) / 16, np = nb * 16;
    double* xr = lds;  // LDS form: [np][16]
    double* xi = lds + (GX ? 0 : (size_t)np * SOLVE_COLS);
    int* sigma = reinterpret_cast<int*>(lds + (GX ? 0 : 2 * (size_t)np * SOLVE_COLS));  // [n]
    int* pv = sigma + n;                                                                 // [n]
    // element (row, column c0 + c) of X
    auto xget = [&](int row, int c, double& vr, double& vi) {
        if constexpr (GX) {
            const int gc = c0 + c;
            if (row < n && gc < it.nrhs) {
                vr = it.x_re[gc * it.xsc + row * it.xsr];
                vi = it.x_im[gc * it.xsc + row * it.xsr];
            } else {
                vr = vi = 0.0;
            }
        } else {
            vr = xr[row * SOLVE_COLS + c];
            vi = xi[row * SOLVE_COLS + c];
        }
    };
    auto xset = [&](int row, int c, double vr, double vi) {
        if constexpr (GX) {
            const int gc = c0 + c;
            if (row < n && gc < it.nrhs) {
                it.x_re[gc * it.xsc + row * it.xsr] = vr;
                it.x_im[gc * it.xsc + row * it.xsr] = vi;
            }
        } else {
            xr[row * SOLVE_COLS + c] = vr;
            xi[row * SOLVE_COLS + c] = vi;
        }
    };
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const double* lre = it.lu_re;
    const double* lim = it.lu_im;

    // row permutation: original row r ends at position pos(r) after the swaps piv[0..n-1];
    // sigma[pos(r)] = r, so (P B)[i] = B[sigma[i]]
    const int mode = it.mode;
    if (mode & SOLVE_PERM) {
        for (int i = tid; i < n; i += SOLVE_THREADS) pv[i] = it.piv[i];
        __syncthreads();
        for (int r = tid; r < n; r += SOLVE_THREADS) {
            int pos = r;
            for (int k = 0; k < n; ++k) {
                const int p = pv[k];
                pos = pos == k ? p : (pos == p ? k : pos);
            }
            sigma[pos] = r;
        }
    } else {
        for (int r = tid; r < n; r += SOLVE_THREADS) sigma[r] = r;
    }
    __syncthreads();
    for (int e = tid; e < np * SOLVE_COLS; e += SOLVE_THREADS) {
        const int row = e / SOLVE_COLS, cc = e % SOLVE_COLS, gc = c0 + cc;
        double vr = 0, vi = 0;
        if (row < n && gc < it.nrhs) {
            const int src = sigma[row];
            if (it.b_re) {
                vr = it.b_re[gc * it.bsc + src * it.bsr];
                vi = it.b_im[gc * it.bsc + src * it.bsr];
            } else {
                vr = (src == gc) ? 1.0 : 0.0;
            }
        }
        xset(row, cc, vr, vi);
    }
    __syncthreads();

    const int ei = tid >> 4, ec = tid & 15;  // this thread's element of a 16x16 block
    for (int sweep = 0; sweep < 2; ++sweep) {  // 0: L (unit lower), 1: U (upper)
        if (!(mode & (sweep == 0 ? SOLVE_L : SOLVE_U))) continue;
        for (int bi = 0; bi < nb; ++bi) {
            const int b = sweep == 0 ? bi : nb - 1 - bi;
            const int r0 = b * 16;
            const int arow = r0 + li;
            // ---- partial = A[r0..r0+15, K] X[K], K = [0, r0) (L) or [r0+16, n) (U)
            dbl4 rr = {0, 0, 0, 0}, ii = rr, ri = rr, ir = rr;
            const int kbeg = sweep == 0 ? 0 : r0 + 16;
            const int kend = sweep == 0 ? r0 : np;
#pragma unroll 4
            for (int ks = kbeg + wave * 4; ks < kend; ks += 16) {
                const int k = ks + lk;
                const bool ok = arow < n && k < n;
                const double a_r = ok ? lre[(long)arow * ld + k] : 0.0;
                const double a_i = ok ? lim[(long)arow * ld + k] : 0.0;
                double b_r, b_i;
                xget(k, li, b_r, b_i);
                rr = mfma16(a_r, b_r, rr);
                ii = mfma16(a_i, b_i, ii);
                ri = mfma16(a_r, b_i, ri);
                ir = mfma16(a_i, b_r, ir);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                part_r[wave][lk + 4 * r][li] = rr[r] - ii[r];
                part_i[wave][lk + 4 * r][li] = ri[r] + ir[r];
            }
            {  // inverse of the 16x16 diagonal block: a diagonal sub-block of the panel inverse
                const int off = r0 % LU_NB;
                const double* d = it.dinv + (size_t)LU_DINV * (r0 / LU_NB) + (sweep == 0 ? 0 : 2 * LU_NB * LU_NB);
                dg_r[ei][ec] = d[(off + ei) * LU_NB + off + ec];
                dg_i[ei][ec] = d[LU_NB * LU_NB + (off + ei) * LU_NB + off + ec];
            }
            __syncthreads();
            double xv_r, xv_i;
            xget(r0 + ei, ec, xv_r, xv_i);
            y_r[ei][ec] = xv_r - (part_r[0][ei][ec] + part_r[1][ei][ec] + part_r[2][ei][ec] + part_r[3][ei][ec]);
            y_i[ei][ec] = xv_i - (part_i[0][ei][ec] + part_i[1][ei][ec] + part_i[2][ei][ec] + part_i[3][ei][ec]);
            __syncthreads();
            if (wave == 0) {  // X block = Dinv * Y
                dbl4 pr = {0, 0, 0, 0}, pi = pr, qr = pr, qi = pr;
#pragma unroll
                for (int ks = 0; ks < 16; ks += 4) {
                    const double a_r = dg_r[li][ks + lk], a_i = dg_i[li][ks + lk];
                    const double b_r = y_r[ks + lk][li], b_i = y_i[ks + lk][li];
                    pr = mfma16(a_r, b_r, pr);
                    pi = mfma16(a_i, b_i, pi);
                    qr = mfma16(a_r, b_i, qr);
                    qi = mfma16(a_i, b_r, qi);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) xset(r0 + lk + 4 * r, li, pr[r] - pi[r], qr[r] + qi[r]);
            }
            __syncthreads();
        }
    }
    if constexpr (!GX) {
        for (int e = tid; e < n * SOLVE_COLS; e += SOLVE_THREADS) {
            const int row = e / SOLVE_COLS, cc = e % SOLVE_COLS;
            const int gc = c0 + cc;
            if (gc < it.nrhs) {
                it.x_re[gc * it.xsc + row * it.xsr] = xr[row * SOLVE_COLS + cc];
                it.x_im[gc * it.xsc + row * it.xsr] = xi[row * SOLVE_COLS + cc];
            }
        }
    }
}

inline void launch_lu_solve(const SolveItem* d_items, int nitems, int max_n, int max_rhs, hipStream_t s) {
    dim3 grid((max_rhs + SOLVE_COLS - 1) / SOLVE_COLS, nitems);
    const size_t lds = lu_solve_lds(max_n);
    if (max_n <= LU_MAX_N && lds <= 134 * 1024) {
        ISLE_HIP(hipFuncSetAttribute((const void*)lu_solve_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
        hipLaunchKernelGGL(lu_solve_kernel<false>, grid, dim3(SOLVE_THREADS), lds, s, d_items);
    } else {
        if (max_n > LU_TALL_MAX) throw std::invalid_argument("matrix too large for the LU solve kernel");
        const size_t lds_gx = 2 * (size_t)max_n * sizeof(int);
        ISLE_HIP(hipFuncSetAttribute((const void*)lu_solve_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_gx));
        hipLaunchKernelGGL(lu_solve_kernel<true>, grid, dim3(SOLVE_THREADS), lds_gx, s, d_items);
    }
    ISLE_HIP(hipGetLastError());
}

}  // namespace isle
