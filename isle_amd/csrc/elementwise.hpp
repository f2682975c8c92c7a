// Streaming kernels: phase vectors, chain initialisation, diagonal extraction,
// precision conversion, force combination, reductions and leapfrog updates.
// All are HBM/L2-bound; they read and write each element once, coalesced.
#pragma once

#include "common.hpp"

namespace isle {

// ---------------------------------------------------------------------------
// phase vectors d[t][x] = exp(sgn * i * phi[((t - 1) mod nt) * nx + x]),
// phi interleaved complex128.  Written as split planes in FP64 and (optionally) FP32.
struct PhaseItem {
    int chain;     // index into the batch of fields (pointer = phi + chain * phi_stride)
    int spin;      // 0: exp(sgn*i*phi); 1: exp(sgn*phi)  (the spin basis uses aux = -i*phi)
    double sgn;
    double* re64;
    double* im64;
    float* re32;
    float* im32;
    int nx, nt;
};

__global__ void phase_kernel(const PhaseItem* __restrict__ items, const double* __restrict__ phi,
                             long phi_stride) {
    const PhaseItem it = items[blockIdx.y];
    const double* p = phi + it.chain * phi_stride;
    const long total = (long)it.nx * it.nt;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int t = (int)(e / it.nx), x = (int)(e % it.nx);
        const int tm1 = t == 0 ? it.nt - 1 : t - 1;
        const long src = 2 * ((long)tm1 * it.nx + x);
        const double a = p[src], b = p[src + 1];
        // z = sgn*i*(a+ib) = -sgn*b + i*sgn*a ;  spin: z = sgn*(a+ib)
        const double zr = it.spin ? it.sgn * a : -it.sgn * b;
        const double zi = it.spin ? it.sgn * b : it.sgn * a;
        const double mag = exp(zr);
        double s, c;
        sincos(zi, &s, &c);
        const double vr = mag * c, vi = mag * s;
        if (it.re64) {
            it.re64[e] = vr;
            it.im64[e] = vi;
        }
        if (it.re32) {
            it.re32[e] = (float)vr;
            it.im32[e] = (float)vi;
        }
    }
}

// ---------------------------------------------------------------------------
// Chain start: out = diag(d) * A (ROW) or A * diag(d) (COL) for real A (n x n).
template <typename T>
struct ScaleItem {
    const T* a;
    const T* d_re;
    const T* d_im;
    T* out_re;
    T* out_im;
    int n;
    int mode;  // SCALE_ROW / SCALE_COL
};

template <typename T>
__global__ void scale_real_kernel(const ScaleItem<T>* __restrict__ items) {
    const ScaleItem<T> it = items[blockIdx.y];
    const long total = (long)it.n * it.n;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / it.n), j = (int)(e % it.n);
        const int di = it.mode == 1 ? i : j;
        const T v = it.a[e];
        it.out_re[e] = v * it.d_re[di];
        it.out_im[e] = v * it.d_im[di];
    }
}

// Complex matrix scaled by a diagonal: out = X diag(d) (COL) or diag(d) X (ROW).
template <typename T>
struct CScaleItem {
    const T* x_re;
    const T* x_im;
    const T* d_re;
    const T* d_im;
    T* out_re;
    T* out_im;
    int n;
    int mode;
};

template <typename T>
__global__ void scale_complex_kernel(const CScaleItem<T>* __restrict__ items) {
    const CScaleItem<T> it = items[blockIdx.y];
    const long total = (long)it.n * it.n;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / it.n), j = (int)(e % it.n);
        const int di = it.mode == 1 ? i : j;
        const T xr = it.x_re[e], xi = it.x_im[e];
        const T dr = it.d_re[di], dm = it.d_im[di];
        it.out_re[e] = xr * dr - xi * dm;
        it.out_im[e] = xr * dm + xi * dr;
    }
}

// ---------------------------------------------------------------------------
// out[i] = scale * sum_k L[i][k] R[k][i]  (L == null: identity), FP64 accumulation.
// Result written interleaved complex128 at out[2*i], optional per-i divisor q.
template <typename TL, typename TR = TL>
struct DiagItem {
    const TL* l_re;
    const TL* l_im;
    const TR* r_re;  // R, or R^T when rowdot != 0
    const TR* r_im;
    double* out;
    int n;
    int rowdot;      // 1: out[i] = sum_k L[i][k] RT[i][k]  (R stored transposed, both rows contiguous)
};

template <typename TL, typename TR = TL>
__global__ __launch_bounds__(256) void diag_prod_kernel(const DiagItem<TL, TR>* __restrict__ items) {
    const DiagItem<TL, TR> it = items[blockIdx.y];
    const int n = it.n;
    const int ii = threadIdx.x >> 4;   // 0..15
    const int kk = threadIdx.x & 15;   // k lane
    const int i = blockIdx.x * 16 + ii;
    if (blockIdx.x * 16 >= n) return;
    double sr = 0, si = 0;
    if (i < n) {
        if (it.l_re && it.rowdot) {
            for (int k = kk; k < n; k += 16) {
                const double ar = it.l_re[(long)i * n + k], ai = it.l_im[(long)i * n + k];
                const double br = it.r_re[(long)i * n + k], bi = it.r_im[(long)i * n + k];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
        } else if (it.l_re) {
            for (int k = kk; k < n; k += 16) {
                const double ar = it.l_re[(long)i * n + k], ai = it.l_im[(long)i * n + k];
                const double br = it.r_re[(long)k * n + i], bi = it.r_im[(long)k * n + i];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
        } else if (kk == 0) {
            sr = it.r_re[(long)i * n + i];
            si = it.r_im[(long)i * n + i];
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 16);
        si += __shfl_xor(si, off, 16);
    }
    if (kk == 0 && i < n) {
        it.out[2 * i] = sr;
        it.out[2 * i + 1] = si;
    }
}

// out[i] (interleaved complex) = sum over s < slots of part[(s * n + i) * 2 + {0, 1}]: the
// per-16-column partial row sums written by a FUSED_DIAG chain product.
struct DReduceItem {
    const double* part;
    double* out;
    int n, slots;
};

__global__ __launch_bounds__(256) void diag_reduce_kernel(const DReduceItem* __restrict__ items) {
    const DReduceItem it = items[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= it.n) return;
    double sr = 0, si = 0;
    for (int s = 0; s < it.slots; ++s) {
        sr += it.part[((long)s * it.n + i) * 2];
        si += it.part[((long)s * it.n + i) * 2 + 1];
    }
    it.out[2 * i] = sr;
    it.out[2 * i + 1] = si;
}

// ---------------------------------------------------------------------------
// precision conversion between split planes (n elements)
template <typename S, typename D>
__global__ void convert_kernel(const S* __restrict__ src, D* __restrict__ dst, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        dst[e] = (D)src[e];
}

// split planes -> interleaved complex128
__global__ void split_to_interleaved_kernel(const double* __restrict__ re, const double* __restrict__ im,
                                            double* __restrict__ out, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        out[2 * e] = re[e];
        out[2 * e + 1] = im[e];
    }
}

// interleaved complex128 -> split planes
// Interleaved complex [nblk][n][n] to split planes (the inverse of blocks_to_interleaved_kernel):
// block k of the planes starts at element k * boff, row stride ld.  Grid (ceil(n*n / 256), nblk).
__global__ __launch_bounds__(256) void interleaved_to_blocks_kernel(const double* __restrict__ in, long boff, long ld,
                                                                     int n, double* __restrict__ re,
                                                                     double* __restrict__ im) {
    const long k = blockIdx.y, nn = (long)n * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        const long r = e / n, c = e % n, dst = k * boff + r * ld + c;
        re[dst] = in[2 * (k * nn + e)];
        im[dst] = in[2 * (k * nn + e) + 1];
    }
}

// Rows of right-hand sides (interleaved complex [nrhs][len]) to and from split planes holding
// them as columns ([len][nrhs]).  to_planes = 1: rows -> planes; 0: planes -> rows.
__global__ __launch_bounds__(256) void rhs_columns_kernel(double* __restrict__ rows, double* __restrict__ re,
                                                          double* __restrict__ im, long len, int nrhs, int to_planes) {
    const long total = len * nrhs;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long i = e / nrhs, c = e % nrhs, o = 2 * (c * len + i);
        if (to_planes) {
            re[e] = rows[o];
            im[e] = rows[o + 1];
        } else {
            rows[o] = re[e];
            rows[o + 1] = im[e];
        }
    }
}

// Stacks of n x n blocks in split planes to interleaved complex [nblk][n][n]: block k starts at
// element k * boff of the planes, row stride ld.  Grid (ceil(n*n / 256), nblk).
__global__ __launch_bounds__(256) void blocks_to_interleaved_kernel(const double* __restrict__ re,
                                                                     const double* __restrict__ im, long boff,
                                                                     long ld, int n, double* __restrict__ out) {
    const long k = blockIdx.y, nn = (long)n * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        const long r = e / n, c = e % n, src = k * boff + r * ld + c;
        out[2 * (k * nn + e)] = re[src];
        out[2 * (k * nn + e) + 1] = im[src];
    }
}

__global__ void interleaved_to_split_kernel(const double* __restrict__ in, double* __restrict__ re,
                                            double* __restrict__ im, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        re[e] = in[2 * e];
        im[e] = in[2 * e + 1];
    }
}

// ---------------------------------------------------------------------------
// Combine the two species' diagonal terms into the fermion force (interleaved).
//  mode 0 (particle/hole):          out = -i (fp - fh)
//  mode 1 (particle/hole shortcut): out = -i (fp - conj(fp))
//  mode 2 (spin):                   out = fh - fp
// ``accumulate`` adds to out instead of overwriting.  blockIdx.y selects one of a batch of
// fields: its parts start in_stride and its output out_stride complex elements after the
// previous field's.
__global__ void combine_force_kernel(const double* __restrict__ fp, const double* __restrict__ fh,
                                     double* __restrict__ out, long n, long in_stride, long out_stride, int mode,
                                     int accumulate) {
    fp += 2 * (long)blockIdx.y * in_stride;
    fh += 2 * (long)blockIdx.y * in_stride;
    out += 2 * (long)blockIdx.y * out_stride;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const double pr = fp[2 * e], pi = fp[2 * e + 1];
        double hr, hi;
        if (mode == 1) {
            hr = pr;
            hi = -pi;
        } else {
            hr = fh[2 * e];
            hi = fh[2 * e + 1];
        }
        double vr, vi;
        if (mode == 2) {
            vr = hr - pr;
            vi = hi - pi;
        } else {  // -i (a + i b) = b - i a
            const double ar = pr - hr, ai = pi - hi;
            vr = ai;
            vi = -ar;
        }
        if (accumulate) {
            out[2 * e] += vr;
            out[2 * e + 1] += vi;
        } else {
            out[2 * e] = vr;
            out[2 * e + 1] = vi;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-vector sums of phi (complex) and of phi*phi (bilinear), one workgroup per vector.
// out[4*b + 0..3] = (Re sum phi, Im sum phi, Re sum phi^2, Im sum phi^2)
// Field b starts `stride` complex elements (2*stride doubles) after field b-1.
__global__ __launch_bounds__(256) void sums_kernel(const double* __restrict__ phi, long n, long stride,
                                                   double* __restrict__ out) {
    __shared__ double s[4][256];
    const double* p = phi + 2 * blockIdx.x * stride;
    double a = 0, b = 0, c = 0, d = 0;
    for (long e = threadIdx.x; e < n; e += blockDim.x) {
        const double r = p[2 * e], i = p[2 * e + 1];
        a += r;
        b += i;
        c += r * r - i * i;
        d += 2 * r * i;
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    s[3][threadIdx.x] = d;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Leapfrog pieces on interleaved complex128 vectors (n complex elements).
//  kick:  pi += eps * Re(force) [ - eps * Re(phi) / utilde  when utilde > 0 ]
//  drift: phi += eps * pi
__global__ void kick_kernel(double* __restrict__ pi, const double* __restrict__ force,
                            const double* __restrict__ phi, long n, double eps, double inv_utilde) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        double f = force ? force[2 * e] : 0.0;
        if (phi) f -= phi[2 * e] * inv_utilde;
        pi[2 * e] += eps * f;
    }
}

__global__ void drift_kernel(double* __restrict__ phi, const double* __restrict__ pi, long n, double eps) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        phi[2 * e] += eps * pi[2 * e];
        phi[2 * e + 1] += eps * pi[2 * e + 1];
    }
}

// gauge force -phi/utilde, accumulated or written
__global__ void gauge_force_kernel(const double* __restrict__ phi, double* __restrict__ out, long n,
                                   double inv_utilde, int accumulate) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const double vr = -phi[2 * e] * inv_utilde, vi = -phi[2 * e + 1] * inv_utilde;
        if (accumulate) {
            out[2 * e] += vr;
            out[2 * e + 1] += vi;
        } else {
            out[2 * e] = vr;
            out[2 * e + 1] = vi;
        }
    }
}

// real symmetric helper for exp(X) by Taylor series: out = I + X/1!(I + X/2(I + ...)), Horner steps
// done with GEMMs; this kernel forms out = I + c * Y (n x n).
__global__ void identity_plus_kernel(const double* __restrict__ y, double c, double* __restrict__ out, int n) {
    const long total = (long)n * n;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / n), j = (int)(e % n);
        out[e] = c * y[e] + (i == j ? 1.0 : 0.0);
    }
}


// ---------------------------------------------------------------------------
// itemised copies / conversions (count elements each)
template <typename S, typename D>
struct CopyItem {
    const S* src;
    D* dst;
    long count;
};

template <typename S, typename D>
__global__ void copy_items_kernel(const CopyItem<S, D>* __restrict__ items) {
    const CopyItem<S, D> it = items[blockIdx.y];
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < it.count; e += (long)gridDim.x * blockDim.x)
        it.dst[e] = (D)it.src[e];
}

// n x n complex matrices (re plane then im plane) copied transposed, with conversion
template <typename S, typename D>
struct TCopyItem {
    const S* src;
    D* dst;
    int n;
};

template <typename S, typename D>
__global__ void transpose_items_kernel(const TCopyItem<S, D>* __restrict__ items) {
    __shared__ S tile[32][33];
    const TCopyItem<S, D> it = items[blockIdx.z];
    const int n = it.n;
    const int tiles = (n + 31) / 32;
    const int plane = blockIdx.y;
    const long nn = (long)n * n;
    for (int tb = blockIdx.x; tb < tiles * tiles; tb += gridDim.x) {
        const int bi = tb / tiles, bj = tb % tiles;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
        for (int r = ty; r < 32; r += 8) {
            const int i = bi * 32 + r, j = bj * 32 + tx;
            tile[r][tx] = (i < n && j < n) ? it.src[plane * nn + (long)i * n + j] : S(0);
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int i = bj * 32 + r, j = bi * 32 + tx;  // dst[i][j] = src[j][i]
            if (i < n && j < n) it.dst[plane * nn + (long)i * n + j] = (D)tile[tx][r];
        }
        __syncthreads();
    }
}

// dst = sign diag(d) src^T for complex n x n matrices: FP64 src as re / im planes with row
// stride lds (a block of a larger matrix is fine), FP32 dst as split planes (im n*n after
// re), d complex (d_re, d_im; FP64).  The product is formed in FP64 and rounded once.
// 32 x 32 tiles of both planes through LDS; grid (64, 1, items).
struct TScaleItem {
    const double* s_re;
    const double* s_im;
    long lds;
    double sign;
    float* dst;
    const double* d_re;
    const double* d_im;
    int n;
};

__global__ __launch_bounds__(256) void transpose_scale_kernel(const TScaleItem* __restrict__ items) {
    __shared__ double tr[32][33], ti[32][33];
    const TScaleItem it = items[blockIdx.z];
    const int n = it.n;
    const int tiles = (n + 31) / 32;
    const long nn = (long)n * n;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int tb = blockIdx.x; tb < tiles * tiles; tb += gridDim.x) {
        const int bi = tb / tiles, bj = tb % tiles;
        for (int r = ty; r < 32; r += 8) {
            const int i = bi * 32 + r, j = bj * 32 + tx;
            const bool in = i < n && j < n;
            tr[r][tx] = in ? it.s_re[(long)i * it.lds + j] : 0.0;
            ti[r][tx] = in ? it.s_im[(long)i * it.lds + j] : 0.0;
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int i = bj * 32 + r, j = bi * 32 + tx;  // dst[i][j] = sign d[i] src[j][i]
            if (i < n && j < n) {
                const double vr = it.sign * tr[tx][r], vi = it.sign * ti[tx][r];
                const double dr = it.d_re[i], di = it.d_im[i];
                it.dst[(long)i * n + j] = (float)(vr * dr - vi * di);
                it.dst[nn + (long)i * n + j] = (float)(vr * di + vi * dr);
            }
        }
        __syncthreads();
    }
}

// dst = sign src diag(d) for complex n x n matrices: FP64 src as re / im planes with row
// stride lds (a block of a larger matrix is fine), FP32 dst as split planes (im n*n after
// re), d complex (FP64); the product is formed in FP64 and rounded once.  Grid-stride over
// the n*n elements of each item (grid.y = item).
struct ColScaleItem {
    const double* s_re;
    const double* s_im;
    long lds;
    double sign;
    float* dst;
    const double* d_re;
    const double* d_im;
    int n;
};

__global__ __launch_bounds__(256) void colscale_convert_kernel(const ColScaleItem* __restrict__ items) {
    const ColScaleItem it = items[blockIdx.y];
    const long nn = (long)it.n * it.n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        const long i = e / it.n, j = e % it.n;
        const double vr = it.sign * it.s_re[i * it.lds + j], vi = it.sign * it.s_im[i * it.lds + j];
        const double dr = it.d_re[j], di = it.d_im[j];
        it.dst[e] = (float)(vr * dr - vi * di);
        it.dst[nn + e] = (float)(vr * di + vi * dr);
    }
}

// M(phi) applied to rows (row form of hubbardFermiMatrixExp.cpp:150-166 / Dia.cpp:140-160):
// rows v = x[r] (interleaved complex, time-major nt x n), y_t = sgn_t D_t v_{t-1} with
// D_t = exp(s i phi_{t-1}) (s = +1 particle, -1 hole), sgn_0 = -1 (the antiperiodic wrap),
// sgn_t = +1 otherwise.  EXP: (M v)_t = v_t - S y_t; DIA: (M v)_t = K v_t - y_t (S, K the real
// symmetric slice matrix; the product is a GEMM over the rows of each slice).
// applym_prep_kernel writes the GEMM operands as split planes [re/im][t][r][x]:
//   EXP: ya = y, c = v (the GEMM then subtracts ya S from c);  DIA: ya = v.
__device__ __forceinline__ void applym_y(const double* __restrict__ x, const double* __restrict__ phi, int n, int nt,
                                         double sgn, long r, int t, int i, double& yr, double& yi) {
    const int tp = t == 0 ? nt - 1 : t - 1;
    const long src = 2 * ((r * nt + tp) * (long)n + i);
    const double vr = x[src], vi = x[src + 1];
    const double a = phi[2 * ((long)tp * n + i)], b = phi[2 * ((long)tp * n + i) + 1];
    // exp(s i (a + i b)) = exp(-s b) (cos(s a) + i sin(s a))
    const double mag = exp(-sgn * b) * (t == 0 ? -1.0 : 1.0);
    double sn, cs;
    sincos(sgn * a, &sn, &cs);
    const double pr = mag * cs, pi = mag * sn;
    yr = pr * vr - pi * vi;
    yi = pr * vi + pi * vr;
}

__global__ __launch_bounds__(256) void applym_prep_kernel(const double* __restrict__ x, const double* __restrict__ phi,
                                                          int n, int nt, long nrhs, double sgn, int exp_hop,
                                                          double* __restrict__ ya, double* __restrict__ c) {
    const long total = nrhs * nt * (long)n, plane = total;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = (int)(e % n);
        const long rt = e / n;
        const int t = (int)(rt % nt);
        const long r = rt / nt;
        const long dst = ((long)t * nrhs + r) * n + i;  // [t][r][x]
        if (exp_hop) {
            double yr, yi;
            applym_y(x, phi, n, nt, sgn, r, t, i, yr, yi);
            ya[dst] = yr;
            ya[plane + dst] = yi;
            c[dst] = x[2 * e];
            c[plane + dst] = x[2 * e + 1];
        } else {
            ya[dst] = x[2 * e];
            ya[plane + dst] = x[2 * e + 1];
        }
    }
}

// out[r][t][x] = c (EXP) or c - y (DIA), interleaved.
__global__ __launch_bounds__(256) void applym_finish_kernel(const double* __restrict__ x, const double* __restrict__ phi,
                                                            int n, int nt, long nrhs, double sgn, int exp_hop,
                                                            const double* __restrict__ c, double* __restrict__ out) {
    const long total = nrhs * nt * (long)n, plane = total;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = (int)(e % n);
        const long rt = e / n;
        const int t = (int)(rt % nt);
        const long r = rt / nt;
        const long src = ((long)t * nrhs + r) * n + i;
        double vr = c[src], vi = c[plane + src];
        if (!exp_hop) {
            double yr, yi;
            applym_y(x, phi, n, nt, sgn, r, t, i, yr, yi);
            vr -= yr;
            vi -= yi;
        }
        out[2 * e] = vr;
        out[2 * e + 1] = vi;
    }
}

// Right-hand sides between the caller's layout (nrhs interleaved rows of nt*n complex) and
// blocks [t][re/im][n][nrhs]: 32 x 32 tiles through LDS so that both the row reads and the
// block writes are contiguous.  Grid (ceil(n/32), ceil(nrhs/32), nt), 256 threads.
__global__ __launch_bounds__(256) void rhs_transpose_kernel(const double* __restrict__ rhss, double* __restrict__ out,
                                                            int n, int nt, int nrhs) {
    __shared__ double tr[32][33], ti[32][33];
    const int kb = blockIdx.x * 32, cb = blockIdx.y * 32, t = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {  // rows c = cb + r, sites k = kb + tx
        const int c = cb + r, k = kb + tx;
        double vr = 0, vi = 0;
        if (c < nrhs && k < n) {
            const long src = 2 * ((long)c * nt * n + (long)t * n + k);
            vr = rhss[src];
            vi = rhss[src + 1];
        }
        tr[r][tx] = vr;
        ti[r][tx] = vi;
    }
    __syncthreads();
    double* blk = out + (long)t * 2 * n * nrhs;
    for (int r = ty; r < 32; r += 8) {  // sites k = kb + r, rows c = cb + tx
        const int k = kb + r, c = cb + tx;
        if (k < n && c < nrhs) {
            blk[(long)k * nrhs + c] = tr[tx][r];
            blk[(long)n * nrhs + (long)k * nrhs + c] = ti[tx][r];
        }
    }
}

// inverse of rhs_transpose_kernel (same grid)
__global__ __launch_bounds__(256) void rhs_untranspose_kernel(const double* __restrict__ in, double* __restrict__ rhss,
                                                              int n, int nt, int nrhs) {
    __shared__ double tr[32][33], ti[32][33];
    const int kb = blockIdx.x * 32, cb = blockIdx.y * 32, t = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const double* blk = in + (long)t * 2 * n * nrhs;
    for (int r = ty; r < 32; r += 8) {  // sites k = kb + r, rows c = cb + tx
        const int k = kb + r, c = cb + tx;
        double vr = 0, vi = 0;
        if (k < n && c < nrhs) {
            vr = blk[(long)k * nrhs + c];
            vi = blk[(long)n * nrhs + (long)k * nrhs + c];
        }
        tr[r][tx] = vr;
        ti[r][tx] = vi;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {  // rows c = cb + r, sites k = kb + tx
        const int c = cb + r, k = kb + tx;
        if (c < nrhs && k < n) {
            const long dst = 2 * ((long)c * nt * n + (long)t * n + k);
            rhss[dst] = tr[tx][r];
            rhss[dst + 1] = ti[tx][r];
        }
    }
}

// z_t += z_{t-1} along t (blocks of 2*n*nrhs doubles)
__global__ void prefix_sum_t_kernel(double* __restrict__ z, long blk, int nt) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < blk; e += (long)gridDim.x * blockDim.x) {
        double acc = z[e];
        for (int t = 1; t < nt; ++t) {
            acc += z[(long)t * blk + e];
            z[(long)t * blk + e] = acc;
        }
    }
}

// z_t -= x for t < nt-1 (blocks of blk doubles)
__global__ void sub_last_kernel(double* __restrict__ z, const double* __restrict__ x, long blk, int count) {
    const long total = blk * count;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
        z[e] -= x[e % blk];
}

// out = c * in (complex scalar c) on interleaved complex128 vectors
__global__ void cscale_vec_kernel(const double* __restrict__ in, double* __restrict__ out, long n, double cr,
                                  double ci) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const double a = in[2 * e], b = in[2 * e + 1];
        out[2 * e] = a * cr - b * ci;
        out[2 * e + 1] = a * ci + b * cr;
    }
}

inline unsigned ew_blocks(long n) {
    long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace isle
