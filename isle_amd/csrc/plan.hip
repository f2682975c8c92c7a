// Implementation of the fermion-matrix plan (see plan.hpp) and the C ABI.
#include "plan.hpp"

#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>

#include "../../include/isle_hfa.h"

namespace isle {

namespace {

constexpr int EW_THREADS = 256;

// Host-side staging of one cached operation.
struct Builder {
    OpCache& op;
    std::vector<char> blob;
    struct Pending {
        K kind;
        size_t off;
        int count, mm, mn, flag, aux;
        double flops;
        int tile;
        size_t ibytes;
    };
    std::vector<Pending> pend;

    explicit Builder(OpCache& o) : op(o) {}

    template <typename It>
    void add(K kind, const std::vector<It>& items, int mm = 0, int mn = 0, int flag = 0, int aux = 0, int tile = 0) {
        if (items.empty()) return;
        size_t off = (blob.size() + 63) & ~size_t(63);
        blob.resize(off + items.size() * sizeof(It));
        std::memcpy(blob.data() + off, items.data(), items.size() * sizeof(It));
        pend.push_back({kind, off, (int)items.size(), mm, mn, flag, aux, flops_of(kind, items), tile, sizeof(It)});
    }
    template <typename It>
    static double flops_of(K, const std::vector<It>&) { return 0.0; }
    template <typename T>
    static double flops_of(K kind, const std::vector<GemmItem<T>>& items) {
        const double reals = (kind == K::GEMM64_CC || kind == K::GEMM64_CC_SUB || kind == K::TRSM ||
                              kind == K::GEMM64_CC4 || kind == K::GEMM64_CC4_SUB) ? 4.0 : 2.0;
        double f = 0;
        for (const auto& g : items) f += reals * 2.0 * g.m * (double)g.n * g.k;
        return f;
    }
    double* d64(size_t count) {
        op.bufs.emplace_back(new DevBuf(count * sizeof(double)));
        return op.bufs.back()->as<double>();
    }
    float* d32(size_t count) {
        op.bufs.emplace_back(new DevBuf(count * sizeof(float)));
        return op.bufs.back()->as<float>();
    }
    int* di(size_t count) {
        op.bufs.emplace_back(new DevBuf(count * sizeof(int)));
        return op.bufs.back()->as<int>();
    }
    void finish() {
        op.bufs.emplace_back(new DevBuf(blob.size() ? blob.size() : 64));
        char* dev = op.bufs.back()->as<char>();
        if (!blob.empty()) ISLE_HIP(hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice));
        for (auto& p : pend)
            op.steps.push_back({p.kind, dev + p.off, p.count, p.mm, p.mn, p.flag, p.flops, p.tile, p.aux, p.ibytes, p.off});
        op.blob = std::move(blob);
    }
};

// complex n x n matrix (split planes) inside a block array
template <typename T>
struct CMat {
    T* re;
    T* im;
};
template <typename T>
inline CMat<T> cmat(T* base, long idx, long nn) {
    T* r = base + idx * 2 * nn;
    return {r, r + nn};
}

inline double sigma_of(int species) { return species == PARTICLE ? -1.0 : 1.0; }

template <typename T>
GemmItem<T> gitem(const T* a_re, const T* a_im, const T* b_re, const T* b_im, CMat<T> c, const T* pre_re,
                  const T* pre_im, const T* post_re, const T* post_im, int n) {
    GemmItem<T> g;
    g.a_re = a_re;
    g.a_im = a_im;
    g.b_re = b_re;
    g.b_im = b_im;
    g.c_re = c.re;
    g.c_im = c.im;
    g.pre_re = pre_re;
    g.pre_im = pre_im;
    g.post_re = post_re;
    g.post_im = post_im;
    g.m = g.n = g.k = n;
    g.lda = g.ldb = g.ldc = n;
    return g;
}

inline std::complex<double> to_first_log_branch(std::complex<double> z) {
    return {z.real(), std::remainder(z.imag(), 2.0 * M_PI)};
}

}  // namespace

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Blocked triangular solve X = T^{-1} R on GEMMs, T the unit lower (mode 0) or upper (mode 1)
// triangle of LU factors with the 32x32 diagonal-block inverses the LU stored per panel:
// per panel p (forward for L, backward for U) X_p = D_p^{-1} R_p, then the rows still to solve
// R_q -= T_{q,p} X_p.  R is overwritten; X must not overlap R.  n % 32 == 0.
namespace {
struct TrsmMat {
    const double* f_re;
    const double* f_im;
    long ldf;
    const double* dinv;  // the LU's work buffer (per-panel L11^{-1}, U11^{-1})
    double* r_re;
    double* r_im;
    long ldr;
    double* x_re;
    double* x_im;
    long ldx;
};
// Blocked triangular solve X = L^-1 R (mode 0: unit lower, panels first to last) or
// X = U^-1 R (mode 1: upper, last to first) of `per` matrices at once with w right-hand-side
// columns, from the LU's stored inverses D_p of its 32 x 32 diagonal blocks.  Panels go in
// pairs (p, q): x_p = D_p r_p; r_q -= F_qp x_p; x_q = D_q r_q; then the rows beyond the pair
// take one update with k = 64, half the right-hand-side traffic of two k = 32 updates.
// Every product is a step of its own (flag 1: C -= A B); R is overwritten.
void add_trsm(Builder& b, int mode, int n, int w, const std::vector<TrsmMat>& mats) {
    if (n % LU_NB != 0) throw std::invalid_argument("blocked solve needs n % 32 == 0");
    const int np = n / LU_NB, per = (int)mats.size();
    auto diag = [&](int p) {  // x_p = D_p r_p
        std::vector<GemmItem<double>> gi(per);
        for (int i = 0; i < per; ++i) {
            const TrsmMat& t = mats[i];
            GemmItem<double> d{};
            const double* dv = t.dinv + (size_t)LU_DINV * p + (mode == 0 ? 0 : 2 * LU_NB * LU_NB);
            d.a_re = dv;
            d.a_im = dv + LU_NB * LU_NB;
            d.b_re = t.r_re + (long)p * LU_NB * t.ldr;
            d.b_im = t.r_im + (long)p * LU_NB * t.ldr;
            d.c_re = t.x_re + (long)p * LU_NB * t.ldx;
            d.c_im = t.x_im + (long)p * LU_NB * t.ldx;
            d.m = LU_NB;
            d.n = w;
            d.k = LU_NB;
            d.lda = LU_NB;
            d.ldb = (int)t.ldr;
            d.ldc = (int)t.ldx;
            gi[i] = d;
        }
        b.add(K::TRSM, gi, LU_NB, w, 0, 0, 1);
    };
    auto update = [&](int r0, int rows, int c0, int k) {  // r[r0, rows) -= F[r0.., c0..c0+k) x[c0..c0+k)
        if (rows <= 0) return;
        std::vector<GemmItem<double>> gi(per);
        for (int i = 0; i < per; ++i) {
            const TrsmMat& t = mats[i];
            GemmItem<double> u{};
            u.a_re = t.f_re + (long)r0 * t.ldf + c0;
            u.a_im = t.f_im + (long)r0 * t.ldf + c0;
            u.b_re = t.x_re + (long)c0 * t.ldx;
            u.b_im = t.x_im + (long)c0 * t.ldx;
            u.c_re = t.r_re + (long)r0 * t.ldr;
            u.c_im = t.r_im + (long)r0 * t.ldr;
            u.m = rows;
            u.n = w;
            u.k = k;
            u.lda = (int)t.ldf;
            u.ldb = (int)t.ldx;
            u.ldc = (int)t.ldr;
            gi[i] = u;
        }
        b.add(K::TRSM, gi, rows, w, 1, 0, rows <= LU_NB ? 1 : 0);
    };
    if (mode == 0) {
        for (int p = 0; p < np; p += 2) {
            diag(p);
            if (p + 1 < np) {
                update((p + 1) * LU_NB, LU_NB, p * LU_NB, LU_NB);
                diag(p + 1);
                update((p + 2) * LU_NB, n - (p + 2) * LU_NB, p * LU_NB, 2 * LU_NB);
            } else {
                update((p + 1) * LU_NB, n - (p + 1) * LU_NB, p * LU_NB, LU_NB);
            }
        }
    } else {
        for (int p = np - 1; p >= 0; p -= 2) {
            diag(p);
            if (p >= 1) {
                update((p - 1) * LU_NB, LU_NB, p * LU_NB, LU_NB);
                diag(p - 1);
                update(0, (p - 1) * LU_NB, (p - 1) * LU_NB, 2 * LU_NB);
            } else {
                update(0, p * LU_NB, p * LU_NB, LU_NB);
            }
        }
    }
}
}  // namespace

Plan::Plan(int hopping, int nx, const double* kappa, double mu, int sk)
    : hop_(hopping), nx_(nx), mu_(mu), sk_(sk), kappa_(kappa, kappa + (size_t)nx * nx) {
    if (nx <= 0) throw std::invalid_argument("nx must be positive");
    if (hopping != DIA && hopping != EXP) throw std::invalid_argument("unknown hopping");
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nx; ++j)
            if (kappa_[(size_t)i * nx + j] != kappa_[(size_t)j * nx + i])
                throw std::invalid_argument("Hopping matrix is not symmetric.");
    const int n = nx;
    const size_t nn = (size_t)n * n;
    double trk = 0;
    for (int i = 0; i < n; ++i) trk += kappa_[(size_t)i * n + i];
    std::vector<double> id(nn, 0.0);
    for (int i = 0; i < n; ++i) id[(size_t)i * n + i] = 1.0;
    ident64_.reset(new DevBuf(nn * sizeof(double)));
    ISLE_HIP(hipMemcpy(ident64_->p, id.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    for (int s = 0; s < 2; ++s) {
        a64_[s].reset(new DevBuf(nn * sizeof(double)));
        a32_[s].reset(new DevBuf(nn * sizeof(float)));
        e64_[s].reset(new DevBuf(nn * sizeof(double)));
    }
    if (hop_ == EXP) {
        // computeExponential (hubbardFermiMatrixExp.cpp:24-45)
        std::vector<double> xp(nn), xh(nn), mxp(nn), mxh(nn);
        for (size_t e = 0; e < nn; ++e) {
            const double idv = id[e];
            xp[e] = -kappa_[e] + mu * idv;        // particle, inv
            xh[e] = -sk * kappa_[e] - mu * idv;   // hole, inv
            mxp[e] = -xp[e];
            mxh[e] = -xh[e];
        }
        build_exp(xp, *a64_[PARTICLE]);
        build_exp(xh, *a64_[HOLE]);
        build_exp(mxp, *e64_[PARTICLE]);
        build_exp(mxh, *e64_[HOLE]);
        logdet_a_[PARTICLE] = -trk + n * mu;     // log det exp(X) = tr X
        logdet_a_[HOLE] = -sk * trk - n * mu;
    } else {
        // K (hubbardFermiMatrixDia.cpp:96-103)
        std::vector<double> kp(nn), kh(nn);
        for (size_t e = 0; e < nn; ++e) {
            kp[e] = (1 + mu) * id[e] - kappa_[e];
            kh[e] = (1 - mu) * id[e] - sk * kappa_[e];
        }
        ISLE_HIP(hipMemcpy(a64_[PARTICLE]->p, kp.data(), nn * 8, hipMemcpyHostToDevice));
        ISLE_HIP(hipMemcpy(a64_[HOLE]->p, kh.data(), nn * 8, hipMemcpyHostToDevice));
        ISLE_HIP(hipMemcpy(e64_[PARTICLE]->p, id.data(), nn * 8, hipMemcpyHostToDevice));
        ISLE_HIP(hipMemcpy(e64_[HOLE]->p, id.data(), nn * 8, hipMemcpyHostToDevice));
    }
    for (int s = 0; s < 2; ++s) {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(ew_blocks(nn)), dim3(EW_THREADS), 0, nullptr,
                           a64_[s]->as<double>(), a32_[s]->as<float>(), (long)nn);
        ISLE_HIP(hipGetLastError());
    }
    ISLE_HIP(hipDeviceSynchronize());
}

// exp of a real symmetric matrix: scaling and squaring with a degree-18 Taylor
// polynomial evaluated by Horner's rule, all products on the device.
void Plan::build_exp(const std::vector<double>& x, DevBuf& out) {
    const int n = nx_;
    const size_t nn = (size_t)n * n;
    double norm1 = 0;
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += std::fabs(x[(size_t)i * n + j]);
        norm1 = std::max(norm1, s);
    }
    int sq = 0;
    while (norm1 / std::ldexp(1.0, sq) > 0.125) ++sq;
    std::vector<double> y(nn);
    for (size_t e = 0; e < nn; ++e) y[e] = std::ldexp(x[e], -sq);
    DevBuf dy(nn * 8), dt(nn * 8), dtmp(nn * 8), ditem(sizeof(GemmItem<double>));
    ISLE_HIP(hipMemcpy(dy.p, y.data(), nn * 8, hipMemcpyHostToDevice));
    const int deg = 18;
    auto gemm_real = [&](const double* a, const double* b, double* c) {
        GemmItem<double> g = gitem<double>(a, nullptr, b, nullptr, CMat<double>{c, nullptr}, nullptr, nullptr,
                                           nullptr, nullptr, n);
        ISLE_HIP(hipMemcpy(ditem.p, &g, sizeof(g), hipMemcpyHostToDevice));
        launch_gemm<double, false, false, false, SCALE_NONE>(ditem.as<GemmItem<double>>(), 1, n, n, nullptr);
        ISLE_HIP(hipDeviceSynchronize());
    };
    // T = I + Y/deg
    hipLaunchKernelGGL(identity_plus_kernel, dim3(ew_blocks(nn)), dim3(EW_THREADS), 0, nullptr, dy.as<double>(),
                       1.0 / deg, dt.as<double>(), n);
    ISLE_HIP(hipGetLastError());
    for (int k = deg - 1; k >= 1; --k) {
        gemm_real(dy.as<double>(), dt.as<double>(), dtmp.as<double>());
        hipLaunchKernelGGL(identity_plus_kernel, dim3(ew_blocks(nn)), dim3(EW_THREADS), 0, nullptr,
                           dtmp.as<double>(), 1.0 / k, dt.as<double>(), n);
        ISLE_HIP(hipGetLastError());
    }
    for (int s = 0; s < sq; ++s) {
        gemm_real(dt.as<double>(), dt.as<double>(), dtmp.as<double>());
        ISLE_HIP(hipMemcpy(dt.p, dtmp.p, nn * 8, hipMemcpyDeviceToDevice));
    }
    // symmetrise (exact symmetry of exp of a symmetric matrix)
    std::vector<double> h(nn);
    ISLE_HIP(hipMemcpy(h.data(), dt.p, nn * 8, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double v = 0.5 * (h[(size_t)i * n + j] + h[(size_t)j * n + i]);
            h[(size_t)i * n + j] = h[(size_t)j * n + i] = v;
        }
    ISLE_HIP(hipMemcpy(out.p, h.data(), nn * 8, hipMemcpyHostToDevice));
}

void Plan::slice_matrix(int species, int inv, double* out_host) {
    const size_t nn = (size_t)nx_ * nx_;
    const DevBuf& src = (hop_ == EXP && !inv) ? *e64_[species] : *a64_[species];
    ISLE_HIP(hipMemcpy(out_host, src.p, nn * 8, hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
void Plan::run(OpCache& op, const double* phi) {
    for (const Step& st : op.steps)
        if (!launch_step(op, st, phi)) throw std::runtime_error("a step's launch shape does not fit it");
}

void Plan::set_tile(int kind, int tile) {
    if (kind < 0 || kind >= (int)K::KIND_END) throw std::invalid_argument("unknown step kind");
    if (tile < -1 || tile >= N_TILES) throw std::invalid_argument("tile out of range");
    if (tile >= 0)
        forced_[kind] = tile;
    else
        forced_.erase(kind);
    auto apply = [&](OpCache& op) {
        for (Step& st : op.steps) {
            if ((int)st.kind != kind) continue;
            if (tile >= 0) {
                st.tile = tile;
            } else {
                auto f = tuned_.find(std::make_tuple((int)st.kind, st.max_m, st.max_n, st.count));
                if (f != tuned_.end()) st.tile = f->second;
            }
        }
    };
    for (auto& kv : cache_) apply(*kv.second);
    for (auto& kv : multi_) apply(kv.second->merged);
}

static bool is_gemm(K k) {
    return k == K::GEMM64_RC_PRE || k == K::GEMM64_RC_ROW || k == K::GEMM64_CR_PRE || k == K::GEMM64_CC ||
           k == K::GEMM32_RC_ROW || k == K::GEMM32_RC_PRE || k == K::GEMM64_CC_SUB || k == K::GEMM32_RC_PRE_DIAG ||
           k == K::TRSM || k == K::GEMM32_RC_DIAG || k == K::GEMM32_RCK_DIAG || k == K::GEMM64_CC4 ||
           k == K::GEMM64_CC4_SUB;
}

void Plan::tune(OpCache& op) {
    hipEvent_t e0, e1;
    ISLE_HIP(hipEventCreate(&e0));
    ISLE_HIP(hipEventCreate(&e1));
    for (Step& st : op.steps) {
        if (!is_gemm(st.kind)) continue;
        auto pinned = forced_.find((int)st.kind);
        if (pinned != forced_.end()) {
            st.tile = pinned->second;
            continue;
        }
        auto key = std::make_tuple((int)st.kind, st.max_m, st.max_n, st.count);
        auto f = tuned_.find(key);
        if (f != tuned_.end()) {
            st.tile = f->second;
            continue;
        }
        float best = 1e30f;
        int best_tile = 0;
        for (int tile = 0; tile < N_TILES; ++tile) {
            st.tile = tile;
            if (!launch_step(op, st, nullptr)) continue;  // shape does not fit; warm up otherwise
            ISLE_HIP(hipEventRecord(e0, stream_));
            for (int r = 0; r < 3; ++r) launch_step(op, st, nullptr);
            ISLE_HIP(hipEventRecord(e1, stream_));
            ISLE_HIP(hipEventSynchronize(e1));
            float ms = 0;
            ISLE_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_tile = tile;
            }
        }
        st.tile = best_tile;
        tuned_[key] = best_tile;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void Plan::profile(const double* phi, int opk, int nt, int batch, int group, int nspecies, int spin, int reps,
                   double* out, int nkinds) {
    OpCache& op = opk == 0   ? get_force_plain(nt, batch, nspecies, spin)
                  : opk == 1 ? get_force_mixed(nt, batch, group, nspecies, spin)
                  : opk == 2 ? get_logdet_plain(nt, batch, nspecies, spin)
                  : opk == 3 ? get_logdet_grouped(nt, batch, group, nspecies, spin)
                  : opk == 4 ? get_force_stable(nt, batch, group, nspecies, spin)
                  : opk == 6 ? get_force_stable(nt, batch, group, nspecies, spin, true)
                             : get_logdet_stable(nt, batch, group, nspecies, spin);
    run(op, phi);  // warm up
    const size_t ns = op.steps.size();
    std::vector<hipEvent_t> ev(ns + 1);
    for (auto& e : ev) ISLE_HIP(hipEventCreate(&e));
    std::vector<double> ms(ns, 0.0);
    for (int r = 0; r < reps; ++r) {
        ISLE_HIP(hipEventRecord(ev[0], stream_));
        for (size_t i = 0; i < ns; ++i) {
            launch_step(op, op.steps[i], phi);
            ISLE_HIP(hipEventRecord(ev[i + 1], stream_));
        }
        ISLE_HIP(hipEventSynchronize(ev[ns]));
        for (size_t i = 0; i < ns; ++i) {
            float t = 0;
            ISLE_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            ms[i] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 4 * nkinds; ++k) out[k] = 0;
    for (size_t i = 0; i < ns; ++i) {
        const int k = (int)op.steps[i].kind;
        if (k >= nkinds) continue;
        out[4 * k] += 1;
        out[4 * k + 1] += ms[i] / reps;
        out[4 * k + 2] += op.steps[i].flops;
    }
}

bool Plan::launch_step(OpCache& op, const Step& st, const double* phi) {
    const int n = nx_;
    const long nn = (long)n * n;
    hipStream_t s = stream_;
    {
        switch (st.kind) {
        case K::PHASE:
            hipLaunchKernelGGL(phase_kernel, dim3(ew_blocks((long)n * op.nt), st.count), dim3(EW_THREADS), 0, s,
                               (const PhaseItem*)st.items, phi, op.phi_stride);
            break;
        // every real x complex chain step multiplies a symmetric slice matrix (A_s, E_p) into
        // square n x n planes: the direct-to-LDS tiles apply when n % 32 == 0
        case K::GEMM64_RC_PRE:
            if (!launch_gemm<double, false, true, true, SCALE_NONE>((const GemmItem<double>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile, n % 4 == 0 ? n : 0))
                return false;
            break;
        case K::GEMM64_RC_ROW:
            if (!launch_gemm<double, false, true, false, SCALE_ROW>((const GemmItem<double>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile, n % 4 == 0 ? n : 0))
                return false;
            break;
        case K::GEMM64_CR_PRE:
            if (!launch_gemm<double, true, false, true, SCALE_NONE>((const GemmItem<double>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile)) return false;
            break;
        case K::GEMM64_CC:
            if (!launch_gemm<double, true, true, false, SCALE_NONE>((const GemmItem<double>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile)) return false;
            break;
        case K::GEMM32_RC_ROW:
            if (!launch_gemm<float, false, true, false, SCALE_ROW>((const GemmItem<float>*)st.items, st.count, st.max_m,
                                                              st.max_n, s, st.tile, n % 4 == 0 ? n : 0)) return false;
            break;
        case K::GEMM32_RC_PRE:
            if (!launch_gemm<float, false, true, true, SCALE_NONE>((const GemmItem<float>*)st.items, st.count, st.max_m,
                                                              st.max_n, s, st.tile, n % 4 == 0 ? n : 0)) return false;
            break;
        case K::TCVT64TO32:
            hipLaunchKernelGGL((transpose_items_kernel<double, float>), dim3(64, 2, st.count), dim3(256), 0, s,
                               (const TCopyItem<double, float>*)st.items);
            break;
        case K::SCALE64:
            hipLaunchKernelGGL(scale_real_kernel<double>, dim3(ew_blocks(nn), st.count), dim3(EW_THREADS), 0, s,
                               (const ScaleItem<double>*)st.items);
            break;
        case K::SCALE32:
            hipLaunchKernelGGL(scale_real_kernel<float>, dim3(ew_blocks(nn), st.count), dim3(EW_THREADS), 0, s,
                               (const ScaleItem<float>*)st.items);
            break;
        case K::LU:
            launch_lu((const LuItem*)st.items, st.count, st.max_m, st.flag, s);
            break;
        case K::SOLVE:
            launch_lu_solve((const SolveItem*)st.items, st.count, st.max_m, st.max_n, s);
            break;
        case K::DIAG64:
            hipLaunchKernelGGL(diag_prod_kernel<double>, dim3(cdiv(n, 16), st.count), dim3(256), 0, s,
                               (const DiagItem<double>*)st.items);
            break;
        case K::DIAG32:
            hipLaunchKernelGGL(diag_prod_kernel<float>, dim3(cdiv(n, 16), st.count), dim3(256), 0, s,
                               (const DiagItem<float>*)st.items);
            break;
        case K::DIAG64X32:
            hipLaunchKernelGGL((diag_prod_kernel<double, float>), dim3(cdiv(n, 16), st.count), dim3(256), 0, s,
                               (const DiagItem<double, float>*)st.items);
            break;
        case K::GEMM64_CC_SUB:
            if (!launch_gemm<double, true, true, false, SUBTRACT>((const GemmItem<double>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile)) return false;
            break;
        case K::GEMM64_CC4:
            if (!launch_gemm<double, true, true, false, SCALE_NONE, true>((const GemmItem<double>*)st.items, st.count,
                                                                       st.max_m, st.max_n, s, st.tile)) return false;
            break;
        case K::GEMM64_CC4_SUB:
            if (!launch_gemm<double, true, true, false, SUBTRACT, true>((const GemmItem<double>*)st.items, st.count,
                                                                     st.max_m, st.max_n, s, st.tile)) return false;
            break;
        case K::LU_SHAPE:
            launch_lu_shape((const LuItem*)st.items, st.count, st.max_m, st.max_n, st.flag, st.aux, 0, s);
            break;
        case K::LU_SHAPE_DEFER:  // the columns beyond kf (= flag) are eliminated by later steps
            launch_lu_shape((const LuItem*)st.items, st.count, st.max_m, st.max_n, st.flag, st.aux, 0, s, st.flag);
            break;
        case K::BLOCK:
            hipLaunchKernelGGL(block_fill_kernel, dim3(ew_blocks(st.aux), st.count), dim3(256), 0, s,
                               (const BlockItem*)st.items);
            break;
        case K::SIGMA:
            hipLaunchKernelGGL(sigma_kernel, dim3(st.count), dim3(256), 0, s, (const SigmaItem*)st.items);
            break;
        case K::GATHER:
            hipLaunchKernelGGL(stack_gather_kernel, dim3(cdiv(st.aux, 256), st.max_m, st.count), dim3(256), 0, s,
                               (const GatherItem*)st.items);
            break;
        case K::TBLOCK:
            hipLaunchKernelGGL(tblock_kernel, dim3(64, 2, st.count), dim3(256), 0, s, (const TBlockItem*)st.items);
            break;
        case K::GEMM32_RC_PRE_DIAG:
            if (!launch_gemm<float, false, true, true, FUSED_DIAG>((const GemmItem<float>*)st.items, st.count, st.max_m,
                                                              st.max_n, s, st.tile, n % 4 == 0 ? n : 0)) return false;
            break;
        case K::GEMM32_RC_DIAG:
            if (!launch_gemm<float, false, true, false, FUSED_DIAG>((const GemmItem<float>*)st.items, st.count,
                                                               st.max_m, st.max_n, s, st.tile, n % 4 == 0 ? n : 0))
                return false;
            break;
        case K::GEMM32_RCK_DIAG:
            if (n % 32 != 0 || n % 4 != 0) return false;
            if (!launch_gemm_rck_tile((const GemmItem<float>*)st.items, st.count, st.max_m, st.max_n, s, st.tile))
                return false;
            break;
        case K::COLSCALE64TO32:
            hipLaunchKernelGGL(colscale_convert_kernel, dim3(std::min<unsigned>(cdiv(nn, 256), 1024), st.count),
                               dim3(256), 0, s, (const ColScaleItem*)st.items);
            break;
        case K::TSCALE64TO32:
            hipLaunchKernelGGL(transpose_scale_kernel, dim3(64, 1, st.count), dim3(256), 0, s,
                               (const TScaleItem*)st.items);
            break;
        case K::DREDUCE:
            hipLaunchKernelGGL(diag_reduce_kernel, dim3(cdiv(n, 256), st.count), dim3(256), 0, s,
                               (const DReduceItem*)st.items);
            break;
        case K::RGATHER:
            hipLaunchKernelGGL(row_gather_kernel, dim3(std::min<int>(cdiv(st.aux, 256), 64), st.max_m, st.count), dim3(256),
                               0, s, (const RowGatherItem*)st.items);
            break;
        case K::TRSM:  // one product of a blocked triangular solve (flag 1: C -= A B)
            if (st.flag) {
                if (!launch_gemm<double, true, true, false, SUBTRACT>((const GemmItem<double>*)st.items, st.count,
                                                                      st.max_m, st.max_n, s, st.tile))
                    return false;
            } else if (!launch_gemm<double, true, true, false, SCALE_NONE>((const GemmItem<double>*)st.items, st.count,
                                                                           st.max_m, st.max_n, s, st.tile)) {
                return false;
            }
            break;
        case K::COPY64:
            hipLaunchKernelGGL((copy_items_kernel<double, double>), dim3(ew_blocks(st.max_m), st.count),
                               dim3(EW_THREADS), 0, s, (const CopyItem<double, double>*)st.items);
            break;
        case K::CVT32TO64:
            hipLaunchKernelGGL((copy_items_kernel<float, double>), dim3(ew_blocks(st.max_m), st.count),
                               dim3(EW_THREADS), 0, s, (const CopyItem<float, double>*)st.items);
            break;
        case K::CVT64TO32:
            hipLaunchKernelGGL((copy_items_kernel<double, float>), dim3(ew_blocks(st.max_m), st.count),
                               dim3(EW_THREADS), 0, s, (const CopyItem<double, float>*)st.items);
            break;
        }
        ISLE_HIP(hipGetLastError());
    }
    return true;
}

// ---------------------------------------------------------------------------
// Plain FP64 DIRECT_SINGLE force (hubbardFermiAction.cpp:16-60), all (field, species)
// pairs advanced in lock step:
//   lefts   Y_{nt-1} = X_{nt-1},  Y_t = X_t Y_{t+1}  with X_t = D_t A_s (D_t = exp(sigma i phi_{t-1}))
//   Mid     = (1 + Y_0)^{-1}
//   right   R_{-1} = Mid, R_tau = R_{tau-1} X_tau, carried transposed (A_s is symmetric):
//           RT_tau = A_s (D_tau RT_{tau-1}), so every product is a real x complex GEMM
//   f[nt-1] = diag(Y_0 Mid),  f[tau] = diag(Y_{tau+1} R_tau) = sum_k Y_{tau+1}[i][k] RT_tau[i][k]
OpCache& Plan::get_force_plain(int nt, int batch, int nspecies, int spin) {
    auto key = std::make_tuple(0, nt, batch, 0, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    double* ph = b.d64((size_t)nb * 2 * nt * n);
    double* Y = b.d64((size_t)nb * nt * 2 * nn);
    double* R = b.d64((size_t)nb * nt * 2 * nn);
    double* LUb = b.d64((size_t)nb * 2 * nn);
    int* piv = b.di((size_t)nb * n);
    op->parts = b.d64((size_t)nb * nt * n * 2);
    auto D = [&](int item, int t) { return CMat<double>{ph + (long)item * 2 * nt * n + (long)t * n,
                                                        ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n}; };
    auto mY = [&](int item, int t) { return cmat(Y, (long)item * nt + t, nn); };
    auto mR = [&](int item, int t) { return cmat(R, (long)item * nt + t, nn); };
    auto spec = [&](int item) { return item % nspecies; };

    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item) {
        PhaseItem p{item / nspecies, spin, sigma_of(spec(item)), D(item, 0).re, D(item, 0).im, nullptr, nullptr, n, nt};
        phs.push_back(p);
    }
    b.add(K::PHASE, phs);
    std::vector<ScaleItem<double>> sc;
    for (int item = 0; item < nb; ++item)
        sc.push_back({a64_[spec(item)]->as<double>(), D(item, nt - 1).re, D(item, nt - 1).im, mY(item, nt - 1).re,
                      mY(item, nt - 1).im, n, SCALE_ROW});
    b.add(K::SCALE64, sc);
    for (int t = nt - 2; t >= 0; --t) {
        std::vector<GemmItem<double>> g;
        for (int item = 0; item < nb; ++item) {
            const CMat<double> src = mY(item, t + 1);
            g.push_back(gitem<double>(a64_[spec(item)]->as<double>(), nullptr, src.re, src.im, mY(item, t), nullptr,
                                      nullptr, D(item, t).re, D(item, t).im, n));
        }
        b.add(K::GEMM64_RC_ROW, g, n, n);
    }
    std::vector<CopyItem<double, double>> cp;
    for (int item = 0; item < nb; ++item) cp.push_back({mY(item, 0).re, LUb + (long)item * 2 * nn, 2 * nn});
    b.add(K::COPY64, cp, (int)(2 * nn));
    std::vector<LuItem> lu;
    double* luw = b.d64(nb * lu_work_doubles(n));
    for (int item = 0; item < nb; ++item)
        lu.push_back({LUb + (long)item * 2 * nn, LUb + (long)item * 2 * nn + nn, piv + (long)item * n, nullptr,
                      luw + item * lu_work_doubles(n), n, n});
    b.add(K::LU, lu, n, 0, 1);
    std::vector<SolveItem> sv;
    for (int item = 0; item < nb; ++item) {
        SolveItem si{};
        si.lu_re = LUb + (long)item * 2 * nn;
        si.lu_im = si.lu_re + nn;
        si.piv = piv + (long)item * n;
        si.dinv = luw + item * lu_work_doubles(n);
        si.b_re = si.b_im = nullptr;
        si.x_re = mR(item, nt - 1).re;  // Mid^T: element (r, c) stored at [c][r]
        si.x_im = mR(item, nt - 1).im;
        si.xsr = 1;
        si.xsc = n;
        si.n = n;
        si.ld = n;
        si.nrhs = n;
        sv.push_back(si);
    }
    b.add(K::SOLVE, sv, n, n);
    for (int tau = 0; tau < nt - 1; ++tau) {
        std::vector<GemmItem<double>> g;
        for (int item = 0; item < nb; ++item) {
            const CMat<double> prev = tau == 0 ? mR(item, nt - 1) : mR(item, tau - 1);
            g.push_back(gitem<double>(a64_[spec(item)]->as<double>(), nullptr, prev.re, prev.im, mR(item, tau),
                                      D(item, tau).re, D(item, tau).im, nullptr, nullptr, n));
        }
        b.add(K::GEMM64_RC_PRE, g, n, n);
    }
    std::vector<DiagItem<double>> dg;
    for (int item = 0; item < nb; ++item)
        for (int tau = 0; tau < nt; ++tau) {
            const CMat<double> l = tau == nt - 1 ? mY(item, 0) : mY(item, tau + 1);
            const CMat<double> r = tau == nt - 1 ? mR(item, nt - 1) : mR(item, tau);
            dg.push_back({l.re, l.im, r.re, r.im, op->parts + ((long)item * nt + tau) * 2 * n, n, 1});
        }
    b.add(K::DIAG64, dg);
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// ---------------------------------------------------------------------------
// Mixed-precision DIRECT_SINGLE force.  Slices are split into ng = nt/g groups
// [s_j, e_j].  Per group (FP64, MFMA, all groups/species/fields in one launch):
//   Suf_{e_j} = X_{e_j},  Suf_t = X_t Suf_{t+1}  (Suf_{s_j} = P_j, the group product)
// (FP32 here would put its rounding into P_j, which the top level amplifies by the
// conditioning of 1 + A^{-1}: 1e-4 relative force error near zeros of det M at Nx=288.)
// Top level (FP64):
//   GL_{ng-1} = I, GL_j = P_{j+1} GL_{j+1};  Mid = (1 + P_0 GL_0)^{-1}
//   RR_0 = Mid, RR_j = RR_{j-1} P_{j-1};     K_j = GL_j RR_j
// Per group again (FP32):  Rt_{s_j} = K_j X_{s_j}, Rt_tau = Rt_{tau-1} X_tau, carried
// transposed (A_s symmetric): RT_{s_j} = A_s (D_{s_j} K_j^T), RT_tau = A_s (D_tau RT_{tau-1})
// (with n % 32 == 0 the products read and store the untransposed Rt, see below);
//   f[tau] = diag(Suf_{tau+1} Rt_tau) = sum_k Suf_{tau+1}[i][k] RT_tau[i][k]  (tau < e_j),
//   f[e_j] = diag(RT_{e_j}).
// This is the reference's lefts/right factorisation applied inside each group with the
// global factors folded into K_j; FP32 chains never span more than g slices.
//
// Memory: with n % 32 == 0 (the force diagonal fused into the right-chain products) a right
// chain keeps only its last two products.  With a recompute interval k (set_recompute;
// 0 < k < g, k | g) the suffix chains keep only Suf at s_j + m k (m = 0 .. g/k - 1; m = 0 is
// P_j) and recompute the k - 1 matrices between two of them just before the right chain
// needs them, segment by segment: about g more FP64 products per group for
// (g/k + k + 1) instead of g stored suffix matrices.
OpCache& Plan::get_force_mixed(int nt, int batch, int g, int nspecies, int spin) {
    if (g <= 0 || nt % g != 0) throw std::invalid_argument("group size must divide nt");
    const int ck = (recompute_ > 0 && recompute_ < g && g % recompute_ == 0 && nx_ % 32 == 0) ? recompute_ : 0;
    auto key = std::make_tuple(1 + 16 * ck, nt, batch, g, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    const int ng = nt / g;
    float* ph = b.d32((size_t)nb * 2 * nt * n);
    double* ph64 = b.d64((size_t)nb * 2 * nt * n);
    const bool fuse = n % 32 == 0;
    const int ncp = ck ? g / ck : g;  // stored suffix matrices per group (checkpoints)
    // suffix storage: all of them, or the checkpoints + 2 ping-pong matrices of the first
    // pass + k - 1 recomputed matrices of the current segment, per (item, group)
    const int suf_per_group = ck ? ncp + 2 + (ck - 1) : g;
    double* Suf = b.d64((size_t)nb * ng * suf_per_group * 2 * nn);
    // right chains: every product, or the last two per group when the diagonal is fused
    float* Rt = b.d32((size_t)nb * (fuse ? 2 * ng : nt) * 2 * nn);
    double* GL = b.d64((size_t)nb * ng * 2 * nn);
    double* RR = b.d64((size_t)nb * ng * 2 * nn);
    double* KK = b.d64((size_t)nb * ng * 2 * nn);
    float* K32 = b.d32((size_t)nb * ng * 2 * nn);
    double* LUb = b.d64((size_t)nb * 2 * nn);
    int* piv = b.di((size_t)nb * n);
    op->parts = b.d64((size_t)nb * nt * n * 2);
    auto D = [&](int item, int t) { return CMat<float>{ph + (long)item * 2 * nt * n + (long)t * n,
                                                       ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n}; };
    auto D64 = [&](int item, int t) {
        return CMat<double>{ph64 + (long)item * 2 * nt * n + (long)t * n,
                            ph64 + (long)item * 2 * nt * n + (long)nt * n + (long)t * n};
    };
    // slot s of group j's suffix storage
    auto sufSlot = [&](int item, int j, int slot) {
        return cmat(Suf, ((long)item * ng + j) * suf_per_group + slot, nn);
    };
    // where the first (full) pass keeps Suf_t: every t, or checkpoints / ping-pong
    auto mS = [&](int item, int t) {
        const int j = t / g, r = t % g;
        if (!ck) return sufSlot(item, j, r);
        return r % ck == 0 ? sufSlot(item, j, r / ck) : sufSlot(item, j, ncp + (t & 1));
    };
    // Suf_t as the right chain reads it (t = s_j + 1 .. e_j): stored, or recomputed into the
    // segment slots (the top of a segment below the last is the next checkpoint)
    auto mSL = [&](int item, int t) {
        const int j = t / g, r = t % g;
        if (!ck) return sufSlot(item, j, r);
        const int m = (r - 1) / ck;
        if (r == (m + 1) * ck) return sufSlot(item, j, m + 1);
        return sufSlot(item, j, ncp + 2 + (r - m * ck - 1));
    };
    auto mRt = [&](int item, int t) {
        return fuse ? cmat(Rt, ((long)item * ng + t / g) * 2 + (t & 1), nn) : cmat(Rt, (long)item * nt + t, nn);
    };
    auto m64 = [&](double* base, int item, int j) { return cmat(base, (long)item * ng + j, nn); };
    auto mP = [&](int item, int j) { return mS(item, j * g); };  // group product P_j = Suf_{s_j}
    auto mK32 = [&](int item, int j) { return cmat(K32, (long)item * ng + j, nn); };
    auto spec = [&](int item) { return item % nspecies; };
    const float* A32[2] = {a32_[0]->as<float>(), a32_[1]->as<float>()};
    const double* A64[2] = {a64_[0]->as<double>(), a64_[1]->as<double>()};

    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item)
        phs.push_back({item / nspecies, spin, sigma_of(spec(item)), D64(item, 0).re, D64(item, 0).im, D(item, 0).re,
                       D(item, 0).im, n, nt});
    b.add(K::PHASE, phs);
    // ---- local suffix chains in FP64: their first elements are the group products P_j,
    //      whose rounding the top level would amplify by the conditioning of 1 + A^{-1}
    {
        std::vector<ScaleItem<double>> sc;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int e = j * g + g - 1;
                sc.push_back({A64[spec(item)], D64(item, e).re, D64(item, e).im, mS(item, e).re, mS(item, e).im, n,
                              SCALE_ROW});
            }
        b.add(K::SCALE64, sc);
    }
    for (int q = 1; q < g; ++q) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int t = j * g + g - 1 - q;
                const CMat<double> src = mS(item, t + 1);
                gi.push_back(gitem<double>(A64[spec(item)], nullptr, src.re, src.im, mS(item, t), nullptr, nullptr,
                                           D64(item, t).re, D64(item, t).im, n));
            }
        b.add(K::GEMM64_RC_ROW, gi, n, n);
    }
    // ---- top level (FP64)
    if (ng == 1) {
        std::vector<CopyItem<double, double>> cp;
        for (int item = 0; item < nb; ++item) cp.push_back({mP(item, 0).re, LUb + (long)item * 2 * nn, 2 * nn});
        b.add(K::COPY64, cp, (int)(2 * nn));
    } else {
        std::vector<CopyItem<double, double>> cp;
        for (int item = 0; item < nb; ++item) cp.push_back({mP(item, ng - 1).re, m64(GL, item, ng - 2).re, 2 * nn});
        b.add(K::COPY64, cp, (int)(2 * nn));
        for (int j = ng - 3; j >= 0; --j) {
            std::vector<GemmItem<double>> gi;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> a = mP(item, j + 1), bb = m64(GL, item, j + 1);
                gi.push_back(gitem<double>(a.re, a.im, bb.re, bb.im, m64(GL, item, j), nullptr, nullptr, nullptr,
                                           nullptr, n));
            }
            b.add(K::GEMM64_CC, gi, n, n);
        }
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item) {
            const CMat<double> a = mP(item, 0), bb = m64(GL, item, 0);
            gi.push_back(gitem<double>(a.re, a.im, bb.re, bb.im, cmat(LUb, item, nn), nullptr, nullptr, nullptr,
                                       nullptr, n));
        }
        b.add(K::GEMM64_CC, gi, n, n);
    }
    {
        std::vector<LuItem> lu;
        double* luw = b.d64(nb * lu_work_doubles(n));
        for (int item = 0; item < nb; ++item)
            lu.push_back({LUb + (long)item * 2 * nn, LUb + (long)item * 2 * nn + nn, piv + (long)item * n, nullptr,
                          luw + item * lu_work_doubles(n), n, n});
        b.add(K::LU, lu, n, 0, 1);
        if (n % LU_NB == 0) {
            // Mid = (1 + P_0 GL_0)^{-1} = U^{-1} L^{-1} P: the row-permuted identity, then two
            // blocked triangular solves on GEMMs
            int* sig = b.di((size_t)nb * n);
            double* S1 = b.d64((size_t)nb * 2 * nn);
            double* S2 = b.d64((size_t)nb * 2 * nn);
            std::vector<SigmaItem> sv;
            std::vector<GatherItem> ga;
            std::vector<TrsmMat> tl, tu;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> s1 = cmat(S1, item, nn), s2 = cmat(S2, item, nn), rr = m64(RR, item, 0);
                const double* fre = LUb + (long)item * 2 * nn;
                sv.push_back({piv + (long)item * n, sig + (long)item * n, n, n});
                ga.push_back({sig + (long)item * n, nullptr, nullptr, s1.re, s1.im, nullptr, nullptr, n, n, n, n, 0});
                tl.push_back({fre, fre + nn, n, luw + item * lu_work_doubles(n), s1.re, s1.im, n, s2.re, s2.im, n});
                tu.push_back({fre, fre + nn, n, luw + item * lu_work_doubles(n), s2.re, s2.im, n, rr.re, rr.im, n});
            }
            b.add(K::SIGMA, sv);
            b.add(K::GATHER, ga, n, 0, 0, n);
            add_trsm(b, 0, n, n, tl);
            add_trsm(b, 1, n, n, tu);
        } else {
            std::vector<SolveItem> sv;
            for (int item = 0; item < nb; ++item) {
                SolveItem si{};
                si.lu_re = LUb + (long)item * 2 * nn;
                si.lu_im = si.lu_re + nn;
                si.piv = piv + (long)item * n;
                si.dinv = luw + item * lu_work_doubles(n);
                si.x_re = m64(RR, item, 0).re;
                si.x_im = m64(RR, item, 0).im;
                si.xsr = n;
                si.xsc = 1;
                si.n = n;
                si.ld = n;
                si.nrhs = n;
                sv.push_back(si);
            }
            b.add(K::SOLVE, sv, n, n);
        }
    }
    for (int j = 1; j < ng; ++j) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item) {
            const CMat<double> a = m64(RR, item, j - 1), bb = mP(item, j - 1);
            gi.push_back(gitem<double>(a.re, a.im, bb.re, bb.im, m64(RR, item, j), nullptr, nullptr, nullptr, nullptr, n));
        }
        b.add(K::GEMM64_CC, gi, n, n);
    }
    if (ng > 1) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng - 1; ++j) {
                const CMat<double> a = m64(GL, item, j), bb = m64(RR, item, j);
                gi.push_back(gitem<double>(a.re, a.im, bb.re, bb.im, m64(KK, item, j), nullptr, nullptr, nullptr,
                                           nullptr, n));
            }
        b.add(K::GEMM64_CC, gi, n, n);
    }
    // ---- local right chains (FP32), transposed.  With n % 32 == 0 the chain products also
    //      produce the force diagonal f[tau] = rowdot(Suf_{tau+1}, Rt_tau) (diag Rt at a group's
    //      last slice) as partial row sums per 16 columns, summed by one reduction at the end,
    //      and every phase is applied outside the MFMA loop: the group's first phase D_{s_j}
    //      in the conversion of K_j (K32 = K_j D_{s_j}), the next slice's D_{tau+1} to the
    //      product as it is stored (its row sums use the product before that scaling).
    //      Otherwise the phase scales the complex operand along k as it is read.
    const int slots = n / 16;
    double* part = fuse ? b.d64((size_t)nb * nt * slots * n * 2) : nullptr;
    if (fuse) {
        // the right chains are carried untransposed in the fused path: Rt_{s_j} starts from
        // K_j D_{s_j} (FP32), Rt_tau = (Rt_{tau-1} D_tau) A_s, the next phase applied as the
        // product is stored
        std::vector<ColScaleItem> cv;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const double* src = (j < ng - 1) ? m64(KK, item, j).re : m64(RR, item, ng - 1).re;
                cv.push_back({src, src + nn, n, 1.0, mK32(item, j).re, D64(item, j * g).re, D64(item, j * g).im, n});
            }
        b.add(K::COLSCALE64TO32, cv);
    } else {
        std::vector<TCopyItem<double, float>> cv;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const double* src = (j < ng - 1) ? m64(KK, item, j).re : m64(RR, item, ng - 1).re;
                cv.push_back({src, mK32(item, j).re, n});
            }
        b.add(K::TCVT64TO32, cv);
    }
    for (int q = 0; q < g; ++q) {
        if (ck && q % ck == 0) {
            // recompute the suffix matrices the next ck right-chain steps read:
            // Suf_t, t = s_j + q + 1 .. s_j + q + ck (the top one is a checkpoint or, in the
            // last segment, X_{e_j} itself)
            const int top = q + ck;  // offset of the segment's top in its group
            if (top == g) {
                std::vector<ScaleItem<double>> sc;
                for (int item = 0; item < nb; ++item)
                    for (int j = 0; j < ng; ++j) {
                        const int e = j * g + g - 1;
                        sc.push_back({A64[spec(item)], D64(item, e).re, D64(item, e).im, mSL(item, e).re,
                                      mSL(item, e).im, n, SCALE_ROW});
                    }
                b.add(K::SCALE64, sc);
            }
            const int first = top == g ? top - 2 : top - 1;  // offset of the first product to form
            for (int r = first; r > q; --r) {
                std::vector<GemmItem<double>> gi;
                for (int item = 0; item < nb; ++item)
                    for (int j = 0; j < ng; ++j) {
                        const int t = j * g + r;
                        const CMat<double> src = mSL(item, t + 1);
                        gi.push_back(gitem<double>(A64[spec(item)], nullptr, src.re, src.im, mSL(item, t), nullptr,
                                                   nullptr, D64(item, t).re, D64(item, t).im, n));
                    }
                b.add(K::GEMM64_RC_ROW, gi, n, n);
            }
        }
        std::vector<GemmItem<float>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int tau = j * g + q;
                const CMat<float> prev = q == 0 ? mK32(item, j) : mRt(item, tau - 1);
                GemmItem<float> it;
                if (fuse) {
                    const bool more = q < g - 1;
                    it = gitem<float>(A32[spec(item)], nullptr, prev.re, prev.im, mRt(item, tau), nullptr, nullptr,
                                      more ? D(item, tau + 1).re : nullptr, more ? D(item, tau + 1).im : nullptr, n);
                    if (more) {
                        it.dl_re = mSL(item, tau + 1).re;
                        it.dl_im = mSL(item, tau + 1).im;
                    }
                    it.dpart = part + ((long)item * nt + tau) * slots * n * 2;
                } else {
                    it = gitem<float>(A32[spec(item)], nullptr, prev.re, prev.im, mRt(item, tau), D(item, tau).re,
                                      D(item, tau).im, nullptr, nullptr, n);
                }
                gi.push_back(it);
            }
        b.add(fuse ? K::GEMM32_RCK_DIAG : K::GEMM32_RC_PRE, gi, n, n);
    }
    if (fuse) {
        std::vector<DReduceItem> dr;
        for (int item = 0; item < nb; ++item)
            for (int tau = 0; tau < nt; ++tau)
                dr.push_back({part + ((long)item * nt + tau) * slots * n * 2, op->parts + ((long)item * nt + tau) * 2 * n,
                              n, slots});
        b.add(K::DREDUCE, dr);
    } else {
        std::vector<DiagItem<double, float>> dg;
        for (int item = 0; item < nb; ++item)
            for (int tau = 0; tau < nt; ++tau) {
                const bool last = (tau % g) == g - 1;
                const CMat<double> l = last ? CMat<double>{nullptr, nullptr} : mS(item, tau + 1);
                const CMat<float> r = mRt(item, tau);
                dg.push_back({l.re, l.im, r.re, r.im, op->parts + ((long)item * nt + tau) * 2 * n, n, 1});
            }
        b.add(K::DIAG64X32, dg);
    }
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// ---------------------------------------------------------------------------
// Plain FP64 logdetM, the reference's chains (hubbardFermiMatrixExp.cpp:574-612,
// hubbardFermiMatrixDia.cpp:571-597):
//   EXP particle: B = F(nt-1)...F(0), F(t) = E_p diag(exp(i phi_{t-1}))  (B <- E_p (D'_t B))
//   EXP hole / DIA: aux = X_0 X_1 ... X_{nt-1}  carried transposed (A_s symmetric):
//                   auxT <- A_s (D_t auxT), auxT_0 = A_s D_0;  det(1 + aux) = det(1 + auxT)
//   then LU of (1 + .) with its log-determinant.
OpCache& Plan::get_logdet_plain(int nt, int batch, int nspecies, int spin) {
    auto key = std::make_tuple(2, nt, batch, 0, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    double* ph = b.d64((size_t)nb * 2 * nt * n);
    double* W = b.d64((size_t)nb * 2 * 2 * nn);
    int* piv = b.di((size_t)nb * n);
    op->logdet = b.d64((size_t)nb * 2);
    auto D = [&](int item, int t) { return CMat<double>{ph + (long)item * 2 * nt * n + (long)t * n,
                                                        ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n}; };
    auto mW = [&](int item, int k) { return cmat(W, (long)item * 2 + k, nn); };
    auto spec = [&](int item) { return item % nspecies; };
    auto bchain = [&](int item) { return hop_ == EXP && spec(item) == PARTICLE; };

    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item) {
        const double sgn = bchain(item) ? 1.0 : sigma_of(spec(item));
        phs.push_back({item / nspecies, spin, sgn, D(item, 0).re, D(item, 0).im, nullptr, nullptr, n, nt});
    }
    b.add(K::PHASE, phs);
    std::vector<ScaleItem<double>> sc;
    for (int item = 0; item < nb; ++item) {
        if (bchain(item))
            sc.push_back({e64_[PARTICLE]->as<double>(), D(item, 0).re, D(item, 0).im, mW(item, 0).re, mW(item, 0).im, n,
                          SCALE_COL});
        else
            sc.push_back({a64_[spec(item)]->as<double>(), D(item, 0).re, D(item, 0).im, mW(item, 0).re,
                          mW(item, 0).im, n, SCALE_COL});
    }
    b.add(K::SCALE64, sc);
    for (int t = 1; t < nt; ++t) {
        std::vector<GemmItem<double>> gb;
        for (int item = 0; item < nb; ++item) {
            const CMat<double> prev = mW(item, (t - 1) & 1), cur = mW(item, t & 1);
            const double* A = bchain(item) ? e64_[PARTICLE]->as<double>() : a64_[spec(item)]->as<double>();
            gb.push_back(gitem<double>(A, nullptr, prev.re, prev.im, cur, D(item, t).re, D(item, t).im, nullptr,
                                       nullptr, n));
        }
        b.add(K::GEMM64_RC_PRE, gb, n, n);
    }
    std::vector<LuItem> lu;
    double* luw = b.d64(nb * lu_work_doubles(n));
    for (int item = 0; item < nb; ++item) {
        const CMat<double> last = mW(item, (nt - 1) & 1);
        lu.push_back({last.re, last.im, piv + (long)item * n, op->logdet + 2 * item, luw + item * lu_work_doubles(n),
                      n, n});
    }
    b.add(K::LU, lu, n, 0, 1);
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// Grouped FP64 logdetM (inverse form for both species):  P_j = X_{s_j} ... X_{e_j}
// for all groups in parallel, then A^{-1} = P_0 ... P_{ng-1}, then LU of 1 + A^{-1}.
OpCache& Plan::get_logdet_grouped(int nt, int batch, int g, int nspecies, int spin) {
    if (g <= 0 || nt % g != 0) throw std::invalid_argument("group size must divide nt");
    auto key = std::make_tuple(3, nt, batch, g, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    const int ng = nt / g;
    double* ph = b.d64((size_t)nb * 2 * nt * n);
    double* Wg = b.d64((size_t)nb * ng * 2 * 2 * nn);
    double* T = b.d64((size_t)nb * 2 * 2 * nn);
    int* piv = b.di((size_t)nb * n);
    op->logdet = b.d64((size_t)nb * 2);
    auto D = [&](int item, int t) { return CMat<double>{ph + (long)item * 2 * nt * n + (long)t * n,
                                                        ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n}; };
    auto mW = [&](int item, int j, int k) { return cmat(Wg, ((long)item * ng + j) * 2 + k, nn); };
    auto mT = [&](int item, int k) { return cmat(T, (long)item * 2 + k, nn); };
    auto spec = [&](int item) { return item % nspecies; };

    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item)
        phs.push_back({item / nspecies, spin, sigma_of(spec(item)), D(item, 0).re, D(item, 0).im, nullptr, nullptr, n, nt});
    b.add(K::PHASE, phs);
    std::vector<ScaleItem<double>> sc;
    for (int item = 0; item < nb; ++item)
        for (int j = 0; j < ng; ++j) {
            const int e = j * g + g - 1;
            sc.push_back({a64_[spec(item)]->as<double>(), D(item, e).re, D(item, e).im, mW(item, j, 0).re,
                          mW(item, j, 0).im, n, SCALE_ROW});
        }
    b.add(K::SCALE64, sc);
    for (int q = 1; q < g; ++q) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int t = j * g + g - 1 - q;
                const CMat<double> src = mW(item, j, (q - 1) & 1);
                gi.push_back(gitem<double>(a64_[spec(item)]->as<double>(), nullptr, src.re, src.im, mW(item, j, q & 1),
                                           nullptr, nullptr, D(item, t).re, D(item, t).im, n));
            }
        b.add(K::GEMM64_RC_ROW, gi, n, n);
    }
    auto P = [&](int item, int j) { return mW(item, j, (g - 1) & 1); };
    std::vector<LuItem> lu;
    double* luw = b.d64(nb * lu_work_doubles(n));
    auto work = [&](int item) { return luw + item * lu_work_doubles(n); };
    if (ng == 1) {
        for (int item = 0; item < nb; ++item)
            lu.push_back({P(item, 0).re, P(item, 0).im, piv + (long)item * n, op->logdet + 2 * item, work(item), n, n});
    } else {
        for (int j = 1; j < ng; ++j) {
            std::vector<GemmItem<double>> gi;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> a = j == 1 ? P(item, 0) : mT(item, (j - 2) & 1);
                const CMat<double> bb = P(item, j);
                gi.push_back(gitem<double>(a.re, a.im, bb.re, bb.im, mT(item, (j - 1) & 1), nullptr, nullptr, nullptr,
                                           nullptr, n));
            }
            b.add(K::GEMM64_CC, gi, n, n);
        }
        for (int item = 0; item < nb; ++item) {
            const CMat<double> last = mT(item, (ng - 2) & 1);
            lu.push_back({last.re, last.im, piv + (long)item * n, op->logdet + 2 * item, work(item), n, n});
        }
    }
    b.add(K::LU, lu, n, 0, 1);
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// ---------------------------------------------------------------------------
// Structured elimination of the block-cyclic matrix R of the group products (I on the
// diagonal, -P_j right of it, P_{ng-1} bottom left; det R = det(1 + P_0 ... P_{ng-1})).
// GEPP of R only ever chooses pivots in the current block row and the last one, so step k
// factorises the 2n x 3n matrix W_k = [[I, -P_k, 0], [B_k, 0, C_k]] (columns k, k+1, ng-1 of
// R) over its first n columns, and the Schur complement of its bottom rows is [B_{k+1}, C_{k+1}];
// the last step factorises [[I, -P_{ng-2}], [B, C]] completely.  The pivots are those of GEPP
// on R, and no product of more than one group is ever formed, so the factorisation is as
// stable as the group products are accurate.  Requires ng >= 2, n % 32 == 0, n + 32 <= 512.
namespace {
struct Cyclic {
    int n, ng, nb;
    double* W;     // [nb][ng-1] step matrices, re plane then im plane of 2n x 3n each
    double* work;  // [nb][ng-1] LU scratch
    int* piv;      // [nb][ng-1][2n]
    size_t wstride, lwork;
    int ldw(int k) const { return k < ng - 2 ? 3 * n : 2 * n; }
    int kf(int k) const { return k < ng - 2 ? n : 2 * n; }
    double* wre(int item, int k) const { return W + ((size_t)item * (ng - 1) + k) * wstride; }
    double* wim(int item, int k) const { return wre(item, k) + wstride / 2; }
    double* wk(int item, int k) const { return work + ((size_t)item * (ng - 1) + k) * lwork; }
    int* pv(int item, int k) const { return piv + ((size_t)item * (ng - 1) + k) * 2 * n; }
};

// P(item, j) -> CMat<double> of group product j; logdet slots [nb][ng-1][2] (or null)
template <typename PF>
Cyclic add_cyclic_factor(Builder& b, int n, int ng, int nb, PF P, double* logdet) {
    if (ng < 2 || n % LU_NB != 0 || n + LU_NB > LU_TALL_MAX)
        throw std::invalid_argument("block-cyclic factorisation needs nx % 32 == 0 and nx <= 16351");
    Cyclic c;
    c.n = n;
    c.ng = ng;
    c.nb = nb;
    c.wstride = 2 * (size_t)(2 * n) * (3 * n);
    c.lwork = std::max(lu_work_doubles(n, 3 * n), lu_work_doubles(2 * n, 2 * n));
    c.W = b.d64((size_t)nb * (ng - 1) * c.wstride);
    c.work = b.d64((size_t)nb * (ng - 1) * c.lwork);
    c.piv = b.di((size_t)nb * (ng - 1) * 2 * n);
    const long nn = (long)n * n;
    auto blk = [&](std::vector<BlockItem>& v, int item, int k, int r0, int c0, const double* sre, const double* sim,
                   long lds, int mode) {
        const int ld = c.ldw(k);
        v.push_back({c.wre(item, k) + (long)r0 * ld + c0, c.wim(item, k) + (long)r0 * ld + c0, sre, sim, ld, lds, n, n,
                     mode});
    };
    // top block rows of every step: [I, -P_k, 0]
    {
        std::vector<BlockItem> v;
        for (int item = 0; item < nb; ++item)
            for (int k = 0; k < ng - 1; ++k) {
                blk(v, item, k, 0, 0, nullptr, nullptr, 0, 1);
                const CMat<double> p = P(item, k);
                blk(v, item, k, 0, n, p.re, p.im, n, 3);
                if (k < ng - 2) blk(v, item, k, 0, 2 * n, nullptr, nullptr, 0, 0);
            }
        b.add(K::BLOCK, v, 0, 0, 0, (int)nn);
    }
    for (int k = 0; k < ng - 1; ++k) {
        std::vector<BlockItem> v;
        for (int item = 0; item < nb; ++item) {
            const bool last = k == ng - 2;
            if (k == 0) {
                const CMat<double> p = P(item, ng - 1);
                blk(v, item, k, n, 0, p.re, p.im, n, 2);
                if (last) {
                    blk(v, item, k, n, n, nullptr, nullptr, 0, 1);
                } else {
                    blk(v, item, k, n, n, nullptr, nullptr, 0, 0);
                    blk(v, item, k, n, 2 * n, nullptr, nullptr, 0, 1);
                }
            } else {  // Schur complement of the previous step's bottom rows
                const int lp = c.ldw(k - 1);
                const double* sre = c.wre(item, k - 1) + (long)n * lp;
                const double* sim = c.wim(item, k - 1) + (long)n * lp;
                blk(v, item, k, n, 0, sre + n, sim + n, lp, 2);
                if (last) {
                    blk(v, item, k, n, n, sre + 2 * n, sim + 2 * n, lp, 2);
                } else {
                    blk(v, item, k, n, n, nullptr, nullptr, 0, 0);
                    blk(v, item, k, n, 2 * n, sre + 2 * n, sim + 2 * n, lp, 2);
                }
            }
        }
        b.add(K::BLOCK, v, 0, 0, 0, (int)nn);
        // Steps that factorise only the first kf columns leave the columns beyond to one
        // blocked triangular solve and one GEMM with k = kf (U12 = L11^{-1} A12 in place,
        // A22 -= L21 U12) instead of a rank-32 update of them per panel: a third of the
        // trailing-update traffic, and the bulk of the flops on the tuned GEMM tiles.
        const bool defer = c.kf(k) < c.ldw(k);
        std::vector<LuItem> lu;
        for (int item = 0; item < nb; ++item) {
            LuItem it{c.wre(item, k), c.wim(item, k), c.pv(item, k),
                      logdet ? logdet + ((long)item * (ng - 1) + k) * 2 : nullptr, c.wk(item, k), c.kf(k), c.ldw(k)};
            it.m = 2 * n;
            it.ncols = c.ldw(k);
            it.kf = c.kf(k);
            it.nt = n;
            if (defer) it.nupd = c.kf(k);
            lu.push_back(it);
        }
        b.add(defer ? K::LU_SHAPE_DEFER : K::LU_SHAPE, lu, 2 * n, c.ldw(k), c.kf(k), n);
        if (defer) {
            const int kf = c.kf(k), ld = c.ldw(k), w = ld - kf, mb = 2 * n - kf;
            std::vector<TrsmMat> tm;
            std::vector<GemmItem<double>> gs;
            for (int item = 0; item < nb; ++item) {
                double* wre = c.wre(item, k);
                double* wim = c.wim(item, k);
                tm.push_back({wre, wim, ld, c.wk(item, k), wre + kf, wim + kf, ld, wre + kf, wim + kf, ld});
                GemmItem<double> g{};
                g.a_re = wre + (long)kf * ld;  // L21: rows kf.., columns 0..kf-1
                g.a_im = wim + (long)kf * ld;
                g.b_re = wre + kf;  // U12: rows 0..kf-1, columns kf..
                g.b_im = wim + kf;
                g.c_re = wre + (long)kf * ld + kf;  // A22
                g.c_im = wim + (long)kf * ld + kf;
                g.m = mb;
                g.n = w;
                g.k = kf;
                g.lda = g.ldb = g.ldc = ld;
                gs.push_back(g);
            }
            add_trsm(b, 0, kf, w, tm);
            b.add(K::GEMM64_CC_SUB, gs, mb, w);
        }
    }
    return c;
}
}  // namespace

// Stable log det M: FP64 group products (as in the grouped form), then the structured
// factorisation of their block-cyclic matrix; per-step log-determinants are summed on the host.
OpCache& Plan::get_logdet_stable(int nt, int batch, int g, int nspecies, int spin) {
    if (g <= 0 || nt % g != 0) throw std::invalid_argument("group size must divide nt");
    const int ng = nt / g;
    if (ng == 1) return get_logdet_grouped(nt, batch, g, nspecies, spin);
    auto key = std::make_tuple(5, nt, batch, g, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    double* ph = b.d64((size_t)nb * 2 * nt * n);
    double* Wg = b.d64((size_t)nb * ng * 2 * 2 * nn);
    op->logdet = b.d64((size_t)nb * (ng - 1) * 2);
    auto D = [&](int item, int t) { return CMat<double>{ph + (long)item * 2 * nt * n + (long)t * n,
                                                        ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n}; };
    auto mW = [&](int item, int j, int k) { return cmat(Wg, ((long)item * ng + j) * 2 + k, nn); };
    auto spec = [&](int item) { return item % nspecies; };
    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item)
        phs.push_back({item / nspecies, spin, sigma_of(spec(item)), D(item, 0).re, D(item, 0).im, nullptr, nullptr, n, nt});
    b.add(K::PHASE, phs);
    std::vector<ScaleItem<double>> sc;
    for (int item = 0; item < nb; ++item)
        for (int j = 0; j < ng; ++j) {
            const int e = j * g + g - 1;
            sc.push_back({a64_[spec(item)]->as<double>(), D(item, e).re, D(item, e).im, mW(item, j, 0).re,
                          mW(item, j, 0).im, n, SCALE_ROW});
        }
    b.add(K::SCALE64, sc);
    for (int q = 1; q < g; ++q) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int t = j * g + g - 1 - q;
                const CMat<double> src = mW(item, j, (q - 1) & 1);
                gi.push_back(gitem<double>(a64_[spec(item)]->as<double>(), nullptr, src.re, src.im, mW(item, j, q & 1),
                                           nullptr, nullptr, D(item, t).re, D(item, t).im, n));
            }
        b.add(K::GEMM64_RC_ROW, gi, n, n);
    }
    add_cyclic_factor(b, n, ng, nb, [&](int item, int j) { return mW(item, j, (g - 1) & 1); }, op->logdet);
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// Stable DIRECT_SINGLE force.  Local FP64 suffix chains per group give P_j; the top level
// needs K_j = lefts (1 + A^{-1})^{-1} rights restricted to group j, which is a block of R^{-1}:
// K_j = -Y[j+1, j] (j < ng-1), K_{ng-1} = Y[0, ng-1].  Those blocks come from the structured
// factorisation by block forward/back substitution on the identity, columns ordered
// (ng-1, 0, 1, ..., ng-2) so that every step works on a prefix of the columns; then the FP64
// right chains Rt_tau = A_s (D_tau Rt_{tau-1}) from Rt_{s_j - 1} = K_j^T and
// f[tau] = diag(Suf_{tau+1} R_tau) as in the other forms.
OpCache& Plan::get_force_stable(int nt, int batch, int g, int nspecies, int spin, bool right32) {
    if (g <= 0 || nt % g != 0) throw std::invalid_argument("group size must divide nt");
    if (right32 && nt / g < 2) throw std::invalid_argument("FP32 right chains of the stable force need two groups or more");
    auto key = std::make_tuple(right32 ? 6 : 4, nt, batch, g, nspecies, spin);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = batch;
    op->nspecies = nspecies;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nb = batch * nspecies;
    const int ng = nt / g;
    const long wd = (long)ng * n;  // columns of the block solves
    double* ph = b.d64((size_t)nb * 2 * nt * n);
    float* ph32 = right32 ? b.d32((size_t)nb * 2 * nt * n) : nullptr;
    double* Suf = b.d64((size_t)nb * nt * 2 * nn);
    double* Rt = right32 ? nullptr : b.d64((size_t)nb * nt * 2 * nn);
    float* Rt32 = right32 ? b.d32((size_t)nb * ng * 2 * 2 * nn) : nullptr;  // last two products per group
    double* Kt = right32 ? nullptr : b.d64((size_t)nb * ng * 2 * nn);
    float* K32 = right32 ? b.d32((size_t)nb * ng * 2 * nn) : nullptr;
    double* Tb = b.d64((size_t)nb * ng * 2 * n * wd);
    op->parts = b.d64((size_t)nb * nt * n * 2);
    auto D = [&](int item, int t) {
        return CMat<double>{ph + (long)item * 2 * nt * n + (long)t * n,
                            ph + (long)item * 2 * nt * n + (long)nt * n + (long)t * n};
    };
    auto mS = [&](int item, int t) { return cmat(Suf, (long)item * nt + t, nn); };
    auto mRt = [&](int item, int t) { return cmat(Rt, (long)item * nt + t, nn); };
    auto mK = [&](int item, int j) { return cmat(Kt, (long)item * ng + j, nn); };
    auto mT = [&](int item, int k) { return cmat(Tb, (long)item * ng + k, (long)n * wd); };
    auto spec = [&](int item) { return item % nspecies; };
    const double* A64[2] = {a64_[0]->as<double>(), a64_[1]->as<double>()};

    auto D32 = [&](int item, int t) {
        return CMat<float>{ph32 + (long)item * 2 * nt * n + (long)t * n,
                           ph32 + (long)item * 2 * nt * n + (long)nt * n + (long)t * n};
    };
    std::vector<PhaseItem> phs;
    for (int item = 0; item < nb; ++item)
        phs.push_back({item / nspecies, spin, sigma_of(spec(item)), D(item, 0).re, D(item, 0).im,
                       right32 ? D32(item, 0).re : nullptr, right32 ? D32(item, 0).im : nullptr, n, nt});
    b.add(K::PHASE, phs);
    {
        std::vector<ScaleItem<double>> sc;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int e = j * g + g - 1;
                sc.push_back({A64[spec(item)], D(item, e).re, D(item, e).im, mS(item, e).re, mS(item, e).im, n,
                              SCALE_ROW});
            }
        b.add(K::SCALE64, sc);
    }
    for (int q = 1; q < g; ++q) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int t = j * g + g - 1 - q;
                const CMat<double> src = mS(item, t + 1);
                gi.push_back(gitem<double>(A64[spec(item)], nullptr, src.re, src.im, mS(item, t), nullptr, nullptr,
                                           D(item, t).re, D(item, t).im, n));
            }
        b.add(K::GEMM64_RC_ROW, gi, n, n);
    }
    auto mP = [&](int item, int j) { return mS(item, j * g); };
    auto gsub = [&](std::vector<GemmItem<double>>& v, const double* are, const double* aim, long lda, const double* bre,
                    const double* bim, double* cre, double* cim, int cols) {
        GemmItem<double> gm{};
        gm.a_re = are;
        gm.a_im = aim;
        gm.b_re = bre;
        gm.b_im = bim;
        gm.c_re = cre;
        gm.c_im = cim;
        gm.m = n;
        gm.n = cols;
        gm.k = n;
        gm.lda = (int)lda;
        gm.ldb = gm.ldc = (int)wd;
        v.push_back(gm);
    };
    if (ng == 1) {
        // K_0 = (1 + P_0)^{-1}
        double* LUb = b.d64((size_t)nb * 2 * nn);
        int* piv = b.di((size_t)nb * n);
        double* luw = b.d64(nb * lu_work_doubles(n));
        std::vector<CopyItem<double, double>> cp;
        for (int item = 0; item < nb; ++item) cp.push_back({mP(item, 0).re, LUb + (long)item * 2 * nn, 2 * nn});
        b.add(K::COPY64, cp, (int)(2 * nn));
        std::vector<LuItem> lu;
        std::vector<SolveItem> sv;
        for (int item = 0; item < nb; ++item) {
            lu.push_back({LUb + (long)item * 2 * nn, LUb + (long)item * 2 * nn + nn, piv + (long)item * n, nullptr,
                          luw + item * lu_work_doubles(n), n, n});
            SolveItem si{};
            si.lu_re = LUb + (long)item * 2 * nn;
            si.lu_im = si.lu_re + nn;
            si.piv = piv + (long)item * n;
            si.dinv = luw + item * lu_work_doubles(n);
            si.x_re = mK(item, 0).re;
            si.x_im = mK(item, 0).im;
            si.xsr = 1;  // written transposed: K^T
            si.xsc = n;
            si.n = n;
            si.ld = n;
            si.nrhs = n;
            sv.push_back(si);
        }
        b.add(K::LU, lu, n, 0, 1);
        b.add(K::SOLVE, sv, n, n);
    } else {
        const Cyclic c = add_cyclic_factor(b, n, ng, nb, mP, nullptr);
        double* Bm = b.d64((size_t)nb * 2 * 2 * n * wd);
        int* sig = b.di((size_t)nb * (ng - 1) * 2 * n);
        auto mB = [&](int item, int k) { return cmat(Bm, (long)item * 2 + k, (long)n * wd); };
        auto sg = [&](int item, int k) { return sig + ((long)item * (ng - 1) + k) * 2 * n; };
        {  // bottom right-hand side: identity at columns [0, n) (block ng-1), zero elsewhere
            std::vector<BlockItem> v;
            for (int item = 0; item < nb; ++item) {
                v.push_back({mB(item, 0).re, mB(item, 0).im, nullptr, nullptr, wd, 0, n, n, 1});
                v.push_back({mB(item, 0).re + n, mB(item, 0).im + n, nullptr, nullptr, wd, 0, n, (int)(wd - n), 0});
                v.push_back({mB(item, 1).re, mB(item, 1).im, nullptr, nullptr, wd, 0, n, (int)wd, 0});
            }
            b.add(K::BLOCK, v, 0, 0, 0, (int)(n * wd));
            std::vector<SigmaItem> sv;
            for (int item = 0; item < nb; ++item)
                for (int k = 0; k < ng - 1; ++k) sv.push_back({c.pv(item, k), sg(item, k), 2 * n, c.kf(k)});
            b.add(K::SIGMA, sv);
        }
        // per-matrix scratch for the blocked solves (X must not overlap the right-hand side)
        double* Sb = b.d64((size_t)nb * 2 * n * wd);
        auto mS2 = [&](int item) { return cmat(Sb, item, (long)n * wd); };
        for (int k = 0; k < ng - 1; ++k) {
            const int w = (k + 2) * n;
            const int ld = c.ldw(k);
            std::vector<GatherItem> ga;
            std::vector<TrsmMat> tl;
            std::vector<GemmItem<double>> gs;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> bi = mB(item, k & 1), bo = mB(item, (k + 1) & 1), t = mT(item, k), sc = mS2(item);
                ga.push_back({sg(item, k), bi.re, bi.im, sc.re, sc.im, bo.re, bo.im, wd, wd, n, w, (k + 1) * n});
                tl.push_back({c.wre(item, k), c.wim(item, k), ld, c.wk(item, k), sc.re, sc.im, wd, t.re, t.im, wd});
                gsub(gs, c.wre(item, k) + (long)n * ld, c.wim(item, k) + (long)n * ld, ld, t.re, t.im, bo.re, bo.im, w);
            }
            b.add(K::GATHER, ga, 2 * n, 0, 0, w);
            add_trsm(b, 0, n, w, tl);
            b.add(K::GEMM64_CC_SUB, gs, n, w);
        }
        {  // last step: x_{ng-1} from the bottom-right LU, then x_{ng-2}
            const int k = ng - 2, ld = 2 * n;
            std::vector<TrsmMat> t1, t2, t3;
            std::vector<GemmItem<double>> gs;
            std::vector<BlockItem> cp;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> bo = mB(item, (k + 1) & 1), xl = mT(item, ng - 1), x2 = mT(item, ng - 2),
                                   sc = mS2(item);
                const double* fre = c.wre(item, k) + (long)n * ld + n;
                const double* fim = c.wim(item, k) + (long)n * ld + n;
                const double* dv = c.wk(item, k) + (size_t)LU_DINV * (n / LU_NB);
                t1.push_back({fre, fim, ld, dv, bo.re, bo.im, wd, sc.re, sc.im, wd});
                t2.push_back({fre, fim, ld, dv, sc.re, sc.im, wd, xl.re, xl.im, wd});
                gsub(gs, c.wre(item, k) + n, c.wim(item, k) + n, ld, xl.re, xl.im, x2.re, x2.im, (int)wd);
                cp.push_back({sc.re, sc.im, x2.re, x2.im, wd, wd, n, (int)wd, 2});
                t3.push_back({c.wre(item, k), c.wim(item, k), ld, c.wk(item, k), sc.re, sc.im, wd, x2.re, x2.im, wd});
            }
            add_trsm(b, 0, n, (int)wd, t1);
            add_trsm(b, 1, n, (int)wd, t2);
            b.add(K::GEMM64_CC_SUB, gs, n, (int)wd);
            b.add(K::BLOCK, cp, 0, 0, 0, (int)(n * wd));
            add_trsm(b, 1, n, (int)wd, t3);
        }
        for (int k = ng - 3; k >= 0; --k) {  // back substitution on the columns still needed
            const int w = (k + 1) * n, ld = c.ldw(k);
            std::vector<GemmItem<double>> g1, g2;
            std::vector<BlockItem> cp;
            std::vector<TrsmMat> tu;
            for (int item = 0; item < nb; ++item) {
                const CMat<double> x = mT(item, k), x1 = mT(item, k + 1), xl = mT(item, ng - 1), sc = mS2(item);
                gsub(g1, c.wre(item, k) + n, c.wim(item, k) + n, ld, x1.re, x1.im, x.re, x.im, w);
                gsub(g2, c.wre(item, k) + 2 * n, c.wim(item, k) + 2 * n, ld, xl.re, xl.im, x.re, x.im, w);
                cp.push_back({sc.re, sc.im, x.re, x.im, wd, wd, n, w, 2});
                tu.push_back({c.wre(item, k), c.wim(item, k), ld, c.wk(item, k), sc.re, sc.im, wd, x.re, x.im, wd});
            }
            b.add(K::GEMM64_CC_SUB, g1, n, w);
            b.add(K::GEMM64_CC_SUB, g2, n, w);
            b.add(K::BLOCK, cp, 0, 0, 0, n * w);
            add_trsm(b, 1, n, w, tu);
        }
        if (right32) {
            // K_j = sign X (untransposed) to FP32 with the group's first phase: K_j D_{s_j}
            std::vector<ColScaleItem> cv;
            for (int item = 0; item < nb; ++item)
                for (int j = 0; j < ng; ++j) {
                    const CMat<double> x = j < ng - 1 ? mT(item, j + 1) : mT(item, 0);
                    const long col = j < ng - 1 ? (long)(j + 1) * n : 0;
                    float* dst = K32 + ((long)item * ng + j) * 2 * nn;
                    cv.push_back({x.re + col, x.im + col, wd, j < ng - 1 ? -1.0 : 1.0, dst, D(item, j * g).re,
                                  D(item, j * g).im, n});
                }
            b.add(K::COLSCALE64TO32, cv);
        } else {
            std::vector<TBlockItem> tb;
            for (int item = 0; item < nb; ++item)
                for (int j = 0; j < ng; ++j) {
                    const CMat<double> x = j < ng - 1 ? mT(item, j + 1) : mT(item, 0);
                    const long col = j < ng - 1 ? (long)(j + 1) * n : 0;
                    tb.push_back({x.re + col, x.im + col, mK(item, j).re, wd, n, j < ng - 1 ? -1.0 : 1.0});
                }
            b.add(K::TBLOCK, tb);
        }
    }
    if (right32) {
        // ---- local right chains in FP32 on MFMA with the force diagonal fused (as in the
        //      mixed form): RT_tau = A_s (D_tau RT_{tau-1}), the phase of the next slice
        //      applied to the stored rows, f[tau] = rowdot(Suf_{tau+1}, RT_tau)
        const float* A32[2] = {a32_[0]->as<float>(), a32_[1]->as<float>()};
        const int slots = n / 16;
        double* part = b.d64((size_t)nb * nt * slots * n * 2);
        auto mR32 = [&](int item, int t) { return cmat(Rt32, ((long)item * ng + t / g) * 2 + (t & 1), nn); };
        for (int q = 0; q < g; ++q) {
            std::vector<GemmItem<float>> gi;
            for (int item = 0; item < nb; ++item)
                for (int j = 0; j < ng; ++j) {
                    const int tau = j * g + q;
                    const bool more = q < g - 1;
                    const CMat<float> prev = q == 0 ? cmat(K32, (long)item * ng + j, nn) : mR32(item, tau - 1);
                    GemmItem<float> it = gitem<float>(A32[spec(item)], nullptr, prev.re, prev.im, mR32(item, tau),
                                                      nullptr, nullptr, more ? D32(item, tau + 1).re : nullptr,
                                                      more ? D32(item, tau + 1).im : nullptr, n);
                    if (more) {
                        it.dl_re = mS(item, tau + 1).re;
                        it.dl_im = mS(item, tau + 1).im;
                    }
                    it.dpart = part + ((long)item * nt + tau) * slots * n * 2;
                    gi.push_back(it);
                }
            b.add(K::GEMM32_RCK_DIAG, gi, n, n);
        }
        std::vector<DReduceItem> dr;
        for (int item = 0; item < nb; ++item)
            for (int tau = 0; tau < nt; ++tau)
                dr.push_back({part + ((long)item * nt + tau) * slots * n * 2, op->parts + ((long)item * nt + tau) * 2 * n,
                              n, slots});
        b.add(K::DREDUCE, dr);
        b.finish();
        tune(*op);
        auto& ref = *op;
        cache_[key] = std::move(op);
        return ref;
    }
    // ---- local right chains (FP64), transposed
    for (int q = 0; q < g; ++q) {
        std::vector<GemmItem<double>> gi;
        for (int item = 0; item < nb; ++item)
            for (int j = 0; j < ng; ++j) {
                const int tau = j * g + q;
                const CMat<double> prev = q == 0 ? mK(item, j) : mRt(item, tau - 1);
                gi.push_back(gitem<double>(A64[spec(item)], nullptr, prev.re, prev.im, mRt(item, tau), D(item, tau).re,
                                           D(item, tau).im, nullptr, nullptr, n));
            }
        b.add(K::GEMM64_RC_PRE, gi, n, n);
    }
    {
        std::vector<DiagItem<double>> dg;
        for (int item = 0; item < nb; ++item)
            for (int tau = 0; tau < nt; ++tau) {
                const bool last = (tau % g) == g - 1;
                const CMat<double> l = last ? CMat<double>{nullptr, nullptr} : mS(item, tau + 1);
                const CMat<double> r = mRt(item, tau);
                dg.push_back({l.re, l.im, r.re, r.im, op->parts + ((long)item * nt + tau) * 2 * n, n, 1});
            }
        b.add(K::DIAG64, dg);
    }
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

// ---------------------------------------------------------------------------
void Plan::logdetM(const double* phi, int nt, int batch, int mode, int group, int nspecies, int spin,
                   double* out_host) {
    if (nt < 1) throw std::invalid_argument("nt must be positive");
    OpCache& op = mode == 0   ? get_logdet_plain(nt, batch, nspecies, spin)
                  : mode == 1 ? get_logdet_grouped(nt, batch, group, nspecies, spin)
                              : get_logdet_stable(nt, batch, group, nspecies, spin);
    // log-determinant slots per (field, species): one, or one per elimination step
    const int per = (mode == 2 && nt / group > 1) ? nt / group - 1 : 1;
    run(op, phi);
    // sums of phi per field
    DevBuf sums((size_t)batch * 4 * 8);
    hipLaunchKernelGGL(sums_kernel, dim3(batch), dim3(256), 0, stream_, phi, (long)nx_ * nt, (long)nx_ * nt,
                       sums.as<double>());
    ISLE_HIP(hipGetLastError());
    std::vector<double> ld((size_t)batch * nspecies * 2 * per), sm((size_t)batch * 4);
    ISLE_HIP(hipMemcpyAsync(ld.data(), op.logdet, ld.size() * 8, hipMemcpyDeviceToHost, stream_));
    ISLE_HIP(hipMemcpyAsync(sm.data(), sums.p, sm.size() * 8, hipMemcpyDeviceToHost, stream_));
    ISLE_HIP(hipStreamSynchronize(stream_));
    const std::complex<double> I(0, 1);
    for (int bi = 0; bi < batch; ++bi) {
        std::complex<double> sphi(sm[4 * bi], sm[4 * bi + 1]);
        if (spin) sphi = -I * sphi;  // sum of aux = -i phi
        for (int s = 0; s < nspecies; ++s) {
            const int item = bi * nspecies + s;
            std::complex<double> l(0, 0);
            for (int k = 0; k < per; ++k) l += std::complex<double>(ld[2 * (item * per + k)], ld[2 * (item * per + k) + 1]);
            std::complex<double> v;
            if (mode == 0 && hop_ == EXP && s == PARTICLE) {
                v = l;
            } else {
                // log det M = log det(1 + A^{-1}) - sigma_s i sum(phi) - [EXP] nt log det A_s
                v = l - sigma_of(s) * I * sphi;
                if (hop_ == EXP) v -= (double)nt * logdet_a_[s];
            }
            v = to_first_log_branch(v);
            out_host[2 * item] = v.real();
            out_host[2 * item + 1] = v.imag();
        }
    }
}

OpCache& Plan::get_force(int mode, int nt, int batch, int group, int nspecies, int spin) {
    if (nt < 2) throw std::invalid_argument("nt < 2 in HubbardFermiAction algorithm DIRECT_SINGLE not supported");
    return mode == 0   ? get_force_plain(nt, batch, nspecies, spin)
           : mode == 1 ? get_force_mixed(nt, batch, group, nspecies, spin)
                       : get_force_stable(nt, batch, group, nspecies, spin, mode == 3);
}

void Plan::force_parts(const double* phi, int nt, int batch, int mode, int group, int nspecies, int spin,
                       double* out_dev) {
    OpCache& op = get_force(mode, nt, batch, group, nspecies, spin);
    run(op, phi);
    const size_t bytes = (size_t)batch * nspecies * nt * nx_ * 2 * 8;
    ISLE_HIP(hipMemcpyAsync(out_dev, op.parts, bytes, hipMemcpyDeviceToDevice, stream_));
}

// Merging independent operations of one plan.  Each operation is a sequence of dependent
// steps; operations share nothing but read-only slice matrices, so any interleaving that
// keeps every operation's own order is correct.  Two steps can share a launch when they have
// the same kind, descriptor size and launch parameters (GEMM kinds: any m, n, the launch
// covers the largest) and neither binds the caller's fields (PHASE).  The sequences are
// merged one after another into a shortest common supersequence under that relation (dynamic
// programming over the two sequences), so the merged operation has as few launches as that
// greedy order allows; each merged step's descriptors are the concatenation of its members'.
std::unique_ptr<Plan::MultiOp> Plan::merge_ops(const std::vector<OpCache*>& ops) {
    struct Member {
        int op, step;
    };
    struct Slot {
        std::vector<Member> members;
    };
    auto step_of = [&](const Member& mb) -> const Step& { return ops[mb.op]->steps[mb.step]; };
    auto joinable = [&](const Step& a, const Step& b) {
        if (a.kind != b.kind || a.kind == K::PHASE || a.ibytes != b.ibytes || a.flag != b.flag || a.aux != b.aux)
            return false;
        return is_gemm(a.kind) || (a.max_m == b.max_m && a.max_n == b.max_n);
    };
    std::vector<Slot> seq;
    for (int i = 0; i < (int)ops[0]->steps.size(); ++i) seq.push_back({{{0, i}}});
    for (int o = 1; o < (int)ops.size(); ++o) {
        const auto& st = ops[o]->steps;
        const int a = (int)seq.size(), bn = (int)st.size();
        // len[i][j]: length of the shortest merge of seq[i..] and st[j..]
        std::vector<std::vector<int>> len(a + 1, std::vector<int>(bn + 1, 0));
        for (int i = a; i >= 0; --i)
            for (int j = bn; j >= 0; --j) {
                if (i == a || j == bn) {
                    len[i][j] = (a - i) + (bn - j);
                    continue;
                }
                int best = 1 + std::min(len[i + 1][j], len[i][j + 1]);
                if (joinable(step_of(seq[i].members[0]), st[j])) best = std::min(best, 1 + len[i + 1][j + 1]);
                len[i][j] = best;
            }
        std::vector<Slot> out;
        int i = 0, j = 0;
        while (i < a || j < bn) {
            if (i < a && j < bn && joinable(step_of(seq[i].members[0]), st[j]) && len[i][j] == 1 + len[i + 1][j + 1]) {
                Slot sl = seq[i];
                sl.members.push_back({o, j});
                out.push_back(sl);
                ++i;
                ++j;
            } else if (i < a && (j == bn || len[i][j] == 1 + len[i + 1][j])) {
                out.push_back(seq[i++]);
            } else {
                out.push_back({{{o, j}}});
                ++j;
            }
        }
        seq.swap(out);
    }
    auto mo = std::make_unique<MultiOp>();
    mo->ops = ops;
    std::vector<char> blob;
    std::vector<size_t> offs;
    for (const Slot& sl : seq) {
        size_t off = (blob.size() + 63) & ~size_t(63);
        if (sl.members.size() > 1) {
            size_t total = 0;
            for (const Member& mb : sl.members) total += step_of(mb).count * step_of(mb).ibytes;
            blob.resize(off + total);
            size_t pos = off;
            for (const Member& mb : sl.members) {
                const Step& s2 = step_of(mb);
                const size_t bytes = s2.count * s2.ibytes;
                std::memcpy(blob.data() + pos, ops[mb.op]->blob.data() + s2.off, bytes);
                pos += bytes;
            }
        }
        offs.push_back(off);
    }
    mo->merged.bufs.emplace_back(new DevBuf(blob.size() ? blob.size() : 64));
    char* dev = mo->merged.bufs.back()->as<char>();
    if (!blob.empty()) ISLE_HIP(hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice));
    for (size_t q = 0; q < seq.size(); ++q) {
        const Slot& sl = seq[q];
        Step st = step_of(sl.members[0]);
        if (sl.members.size() == 1) {  // the member's own descriptors and launch
            mo->owner.push_back(st.kind == K::PHASE ? sl.members[0].op : -1);
            mo->merged.steps.push_back(st);
            continue;
        }
        st.items = dev + offs[q];
        st.off = offs[q];
        st.count = 0;
        st.flops = 0;
        for (const Member& mb : sl.members) {
            const Step& s2 = step_of(mb);
            st.count += s2.count;
            st.flops += s2.flops;
            st.max_m = std::max(st.max_m, s2.max_m);
            st.max_n = std::max(st.max_n, s2.max_n);
        }
        mo->owner.push_back(-1);
        mo->merged.steps.push_back(st);
    }
    mo->merged.blob = std::move(blob);
    return mo;
}

void Plan::force_parts_multi(int nops, const double* const* phis, const int* nts, const int* batches, int mode,
                             const int* groups, int nspecies, int spin, double* const* outs) {
    if (nops < 1) throw std::invalid_argument("need at least one batch of fields");
    std::vector<long> key = {mode, nspecies, spin, recompute_, nops};
    for (int i = 0; i < nops; ++i) {
        if (batches[i] < 1) throw std::invalid_argument("bad batch");
        key.push_back(nts[i]);
        key.push_back(batches[i]);
        key.push_back(groups[i]);
    }
    auto found = multi_.find(key);
    MultiOp* mo = nullptr;
    if (found == multi_.end()) {
        std::vector<OpCache*> ops;
        for (int i = 0; i < nops; ++i) {
            OpCache* op = &get_force(mode, nts[i], batches[i], groups[i], nspecies, spin);
            for (OpCache* o : ops)
                if (o == op) throw std::invalid_argument("two batches of one shape: join them into one batch");
            ops.push_back(op);
        }
        auto made = merge_ops(ops);
        tune(made->merged);
        mo = made.get();
        multi_[key] = std::move(made);
    } else {
        mo = found->second.get();
    }
    for (size_t q = 0; q < mo->merged.steps.size(); ++q) {
        const int o = mo->owner[q];
        const Step& st = mo->merged.steps[q];
        const bool ok = o >= 0 ? launch_step(*mo->ops[o], st, phis[o]) : launch_step(mo->merged, st, nullptr);
        if (!ok) throw std::runtime_error("a step's launch shape does not fit it");
    }
    for (int i = 0; i < nops; ++i) {
        const size_t bytes = (size_t)batches[i] * nspecies * nts[i] * nx_ * 2 * 8;
        ISLE_HIP(hipMemcpyAsync(outs[i], mo->ops[i]->parts, bytes, hipMemcpyDeviceToDevice, stream_));
    }
}

// Block LU of Q = M_p M_h^T (getQLU, hubbardFermiMatrixExp.cpp:419-479, the Dia file
// repeats it), nt >= 3, as one cached operation: the off-diagonal blocks T+(t), T-(t) for all
// t first (a real matrix, signed, times a phase vector on its rows or columns), then the
// recursion, sequential in t:
//   d_0 = P;  u_0 = T-(0), l_0 = T+(1) d_0^{-1}, v_0 = T+(0), h_0 = T-(nt-1) d_0^{-1}
//   d_i = P - l_{i-1} u_{i-1};  l_i = T+(i+1) d_i^{-1}, h_i = -h_{i-1} u_{i-1} d_i^{-1},
//         v_i = -l_{i-1} v_{i-1}, u_i = T-(i)                                  (1 <= i <= nt-3)
//   d_{nt-2} = P - l_{nt-3} u_{nt-3};  u_{nt-2} = T-(nt-2) - l_{nt-3} v_{nt-3},
//         l_{nt-2} = (T+(nt-1) - h_{nt-3} u_{nt-3}) d_{nt-2}^{-1}
//   d_{nt-1} = P - l_{nt-2} u_{nt-2} - sum_i h_i v_i
// Each d_i^{-1} comes from one LU with partial pivoting (its log det kept on the device)
// and its triangular solves; products are complex GEMMs tuned per shape, with the four real
// products of the textbook form (the Gauss form's larger error in the imaginary parts,
// carried through the recursion and the final sum, moved log det Q by ~1e-8 relative at
// Nx=288, Nt=128, against ~3e-10 for the four-product form).  The h_i are
// stored side by side (one nx x (nt-2)nx matrix) and the v_i stacked, so the last sum is
// one GEMM with k = (nt-2) nx.
OpCache& Plan::get_qlu(int nt) {
    if (nt < 3) throw std::invalid_argument("the native QLU needs nt >= 3");
    auto key = std::make_tuple(7, nt, 0, 0, 0, 0);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = 1;
    op->phi_stride = 2L * nx_ * nt;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nh = nt - 2;  // h and v blocks
    // real factors of the off-diagonal blocks, and P, on the host
    std::vector<double> rp(nn), rm(nn), pm(nn, 0.0);
    if (hop_ == EXP) {
        // T+(t) = -+ exp(kappa - mu) diag(e^{i phi_{t-1}}),  T-(t) = -+ diag(e^{-i phi_t}) exp(sk kappa + mu)
        ISLE_HIP(hipMemcpy(rp.data(), e64_[PARTICLE]->p, nn * 8, hipMemcpyDeviceToHost));
        ISLE_HIP(hipMemcpy(rm.data(), e64_[HOLE]->p, nn * 8, hipMemcpyDeviceToHost));
        if (sk_ == -1) {
            for (int i = 0; i < n; ++i) pm[(long)i * n + i] = 2.0;
        } else {
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    const double a = rp[(long)i * n + k];
                    for (int j = 0; j < n; ++j) pm[(long)i * n + j] += a * rm[(long)k * n + j];
                }
            for (int i = 0; i < n; ++i) pm[(long)i * n + i] += 1.0;
        }
    } else {
        // T+(t) = +- diag(e^{i phi_{t-1}}) (sk kappa - (1 - mu)),  T-(t) = +- (kappa - (1 + mu)) diag(e^{-i phi_t})
        for (long e = 0; e < nn; ++e) {
            rp[e] = sk_ * kappa_[e];
            rm[e] = kappa_[e];
        }
        for (int i = 0; i < n; ++i) {
            rp[(long)i * n + i] -= 1.0 - mu_;
            rm[(long)i * n + i] -= 1.0 + mu_;
        }
        // P = (2 - mu^2) - (sk (1 + mu) + 1 - mu) kappa + sk kappa^2
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                const double a = kappa_[(long)i * n + k];
                for (int j = 0; j < n; ++j) pm[(long)i * n + j] += sk_ * a * kappa_[(long)k * n + j];
            }
        for (long e = 0; e < nn; ++e) pm[e] -= (sk_ * (1 + mu_) + 1 - mu_) * kappa_[e];
        for (int i = 0; i < n; ++i) pm[(long)i * n + i] += 2 - mu_ * mu_;
    }
    const bool expo = hop_ == EXP;
    double* R = b.d64(4 * nn);  // rp, -rp, rm, -rm
    {
        std::vector<double> host(4 * nn);
        for (long e = 0; e < nn; ++e) {
            host[e] = rp[e];
            host[nn + e] = -rp[e];
            host[2 * nn + e] = rm[e];
            host[3 * nn + e] = -rm[e];
        }
        ISLE_HIP(hipMemcpy(R, host.data(), 4 * nn * 8, hipMemcpyHostToDevice));
    }
    double* Pb = b.d64(2 * nn);  // P (imaginary part zero)
    {
        std::vector<double> host(2 * nn, 0.0);
        std::copy(pm.begin(), pm.end(), host.begin());
        ISLE_HIP(hipMemcpy(Pb, host.data(), 2 * nn * 8, hipMemcpyHostToDevice));
    }
    double* ph = b.d64(2 * 2 * (size_t)nt * n);  // phases e^{+i phi_{t-1}} and e^{-i phi_{t-1}}
    auto D = [&](int sgn, int t) {
        const double* base = ph + (sgn > 0 ? 0 : 2L * nt * n);
        return CMat<double>{const_cast<double*>(base) + (long)t * n, const_cast<double*>(base) + (long)nt * n + (long)t * n};
    };
    double* Tp = b.d64((size_t)nt * 2 * nn);
    double* Tm = b.d64((size_t)nt * 2 * nn);
    // With n % 32 == 0: [d_i | 1] (n x 2n), whose LU over the first n columns leaves L^{-1} P
    // in the right block; otherwise d_i alone and both triangular solves
    const bool aug = n % LU_NB == 0;
    const long ldd = aug ? 2L * n : (long)n;
    double* Dw = b.d64(2 * (size_t)n * ldd);  // factorised in place
    double* tmp = b.d64(2 * nn);
    double* dinv = b.d64((size_t)nt * 2 * nn);
    double* U = b.d64((size_t)(nt - 1) * 2 * nn);
    double* L = b.d64((size_t)(nt - 1) * 2 * nn);
    double* Vre = b.d64((size_t)nh * nn);
    double* Vim = b.d64((size_t)nh * nn);
    double* Hre = b.d64((size_t)nh * nn);  // n x (nh n), row stride nh n
    double* Him = b.d64((size_t)nh * nn);
    int* piv = b.di(n);
    double* luw = b.d64(lu_work_doubles(n, (int)ldd));
    op->logdet = b.d64(2 * (size_t)nt);
    op->extra = {dinv, U, L, Vre, Vim, Hre, Him};
    const long ldh = (long)nh * n;
    auto tp = [&](int t) { return cmat(Tp, t, nn); };
    auto tm = [&](int t) { return cmat(Tm, t, nn); };
    auto di = [&](int i) { return cmat(dinv, i, nn); };
    auto uu = [&](int i) { return cmat(U, i, nn); };
    auto ll = [&](int i) { return cmat(L, i, nn); };
    auto vv = [&](int i) { return CMat<double>{Vre + (long)i * nn, Vim + (long)i * nn}; };
    auto hh = [&](int i) { return CMat<double>{Hre + (long)i * n, Him + (long)i * n}; };  // row stride ldh
    const long dpl = (long)n * ldd;  // one plane of [d | 1] (or of d)
    const CMat<double> dw{Dw, Dw + dpl}, dr{Dw + n, Dw + dpl + n};  // d and (aug) the identity block
    const CMat<double> tw = cmat(tmp, 0, nn), pb = cmat(Pb, 0, nn);

    std::vector<PhaseItem> phs = {{0, 0, 1.0, D(1, 0).re, D(1, 0).im, nullptr, nullptr, n, nt},
                                  {0, 0, -1.0, D(-1, 0).re, D(-1, 0).im, nullptr, nullptr, n, nt}};
    b.add(K::PHASE, phs);
    {
        std::vector<ScaleItem<double>> sc;
        for (int t = 0; t < nt; ++t) {
            // signs: EXP T+ is + at t = 0, T- + at t = nt-1, - elsewhere; DIA the opposite
            const bool wrap_p = t == 0, wrap_m = t == nt - 1;
            const bool neg_p = expo ? !wrap_p : wrap_p, neg_m = expo ? !wrap_m : wrap_m;
            const double* ap = R + (neg_p ? nn : 0);
            const double* am = R + 2 * nn + (neg_m ? nn : 0);
            const CMat<double> dp = D(1, t), dm = D(-1, (t + 1) % nt);  // e^{i phi_{t-1}}, e^{-i phi_t}
            sc.push_back({ap, dp.re, dp.im, tp(t).re, tp(t).im, n, expo ? SCALE_COL : SCALE_ROW});
            sc.push_back({am, dm.re, dm.im, tm(t).re, tm(t).im, n, expo ? SCALE_ROW : SCALE_COL});
        }
        b.add(K::SCALE64, sc);
    }
    auto cc = [&](CMat<double> a, long lda, CMat<double> bb, long ldb, CMat<double> c, long ldc, int k) {
        GemmItem<double> g{};
        g.a_re = a.re;
        g.a_im = a.im;
        g.b_re = bb.re;
        g.b_im = bb.im;
        g.c_re = c.re;
        g.c_im = c.im;
        g.m = n;
        g.n = n;
        g.k = k;
        g.lda = (int)lda;
        g.ldb = (int)ldb;
        g.ldc = (int)ldc;
        return std::vector<GemmItem<double>>{g};
    };
    auto blk = [&](CMat<double> dst, long ldst, const CMat<double>* src, long lds, int mode) {
        std::vector<BlockItem> v = {{dst.re, dst.im, src ? src->re : nullptr, src ? src->im : nullptr, ldst, lds, n, n,
                                     mode}};
        b.add(K::BLOCK, v, 0, 0, 0, (int)nn);
    };
    // d_i = P and (aug) the identity beside it, in one launch
    auto load_d = [&]() {
        std::vector<BlockItem> v = {{dw.re, dw.im, pb.re, pb.im, ldd, n, n, n, 2}};
        if (aug) v.push_back({dr.re, dr.im, nullptr, nullptr, ldd, 0, n, n, 1});
        b.add(K::BLOCK, v, 0, 0, 0, (int)nn);
    };
    // dinv[i] = d^{-1} and log det d_i: the LU of [d | 1] over its first n columns leaves
    // L^{-1} P in the right block (the interchanges and trailing updates of the LU apply to it
    // on every CU), so only the backward solve with U remains (aug); else both solves
    auto invert = [&](int i) {
        LuItem li{dw.re, dw.im, piv, op->logdet + 2L * i, luw, n, (int)ldd};
        li.m = n;
        li.ncols = (int)ldd;
        li.kf = n;
        std::vector<LuItem> lu = {li};
        b.add(K::LU_SHAPE, lu, n, (int)ldd, n, 0);
        SolveItem si{};
        si.lu_re = dw.re;
        si.lu_im = dw.im;
        si.piv = piv;
        si.dinv = luw;
        if (aug) {  // X = U^{-1} (L^{-1} P), else X = A^{-1} I
            si.b_re = dr.re;
            si.b_im = dr.im;
            si.bsr = ldd;
            si.bsc = 1;
            si.mode = SOLVE_U;
        }
        si.x_re = di(i).re;
        si.x_im = di(i).im;
        si.xsr = n;
        si.xsc = 1;
        si.n = n;
        si.ld = (int)ldd;
        si.nrhs = n;
        std::vector<SolveItem> sv = {si};
        b.add(K::SOLVE, sv, n, n);
    };
    // i = 0
    load_d();
    invert(0);
    {
        const CMat<double> t0 = tm(0), p0 = tp(0);
        blk(uu(0), n, &t0, n, 2);
        b.add(K::GEMM64_CC4, cc(tp(1), n, di(0), n, ll(0), n, n), n, n);
        blk(vv(0), n, &p0, n, 2);
        b.add(K::GEMM64_CC4, cc(tm(nt - 1), n, di(0), n, hh(0), ldh, n), n, n);
    }
    for (int i = 1; i <= nt - 3; ++i) {
        load_d();
        b.add(K::GEMM64_CC4_SUB, cc(ll(i - 1), n, uu(i - 1), n, dw, ldd, n), n, n);
        invert(i);
        b.add(K::GEMM64_CC4, cc(tp(i + 1), n, di(i), n, ll(i), n, n), n, n);
        b.add(K::GEMM64_CC4, cc(hh(i - 1), ldh, uu(i - 1), n, tw, n, n), n, n);
        blk(hh(i), ldh, nullptr, 0, 0);
        b.add(K::GEMM64_CC4_SUB, cc(tw, n, di(i), n, hh(i), ldh, n), n, n);
        blk(vv(i), n, nullptr, 0, 0);
        b.add(K::GEMM64_CC4_SUB, cc(ll(i - 1), n, vv(i - 1), n, vv(i), n, n), n, n);
        const CMat<double> ti = tm(i);
        blk(uu(i), n, &ti, n, 2);
    }
    // i = nt - 2
    load_d();
    b.add(K::GEMM64_CC4_SUB, cc(ll(nt - 3), n, uu(nt - 3), n, dw, ldd, n), n, n);
    invert(nt - 2);
    {
        const CMat<double> t2 = tm(nt - 2), t1 = tp(nt - 1);
        blk(uu(nt - 2), n, &t2, n, 2);
        b.add(K::GEMM64_CC4_SUB, cc(ll(nt - 3), n, vv(nt - 3), n, uu(nt - 2), n, n), n, n);
        blk(tw, n, &t1, n, 2);
        b.add(K::GEMM64_CC4_SUB, cc(hh(nt - 3), ldh, uu(nt - 3), n, tw, n, n), n, n);
        b.add(K::GEMM64_CC4, cc(tw, n, di(nt - 2), n, ll(nt - 2), n, n), n, n);
    }
    // i = nt - 1
    load_d();
    b.add(K::GEMM64_CC4_SUB, cc(ll(nt - 2), n, uu(nt - 2), n, dw, ldd, n), n, n);
    b.add(K::GEMM64_CC4_SUB, cc(CMat<double>{Hre, Him}, ldh, CMat<double>{Vre, Vim}, n, dw, ldd, nh * n), n, n);
    invert(nt - 1);
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

void Plan::qlu(const double* phi, int nt, double* dinv, double* u, double* l, double* v, double* h,
               double* logdet_host) {
    OpCache& op = get_qlu(nt);
    run(op, phi);
    const int n = nx_;
    const long nn = (long)n * n;
    hipStream_t s = stream_;
    auto out = [&](const double* re, const double* im, long boff, long ld, int nblk, double* dst) {
        if (nblk <= 0) return;
        hipLaunchKernelGGL(blocks_to_interleaved_kernel, dim3(std::min<long>((nn + 255) / 256, 1024), nblk), dim3(256),
                           0, s, re, im, boff, ld, n, dst);
        ISLE_HIP(hipGetLastError());
    };
    const auto& e = op.extra;  // dinv, U, L, Vre, Vim, Hre, Him
    out(e[0], e[0] + nn, 2 * nn, n, nt, dinv);
    out(e[1], e[1] + nn, 2 * nn, n, nt - 1, u);
    out(e[2], e[2] + nn, 2 * nn, n, nt - 1, l);
    out(e[3], e[4], nn, n, nt - 2, v);
    out(e[5], e[6], n, (long)(nt - 2) * n, nt - 2, h);
    ISLE_HIP(hipMemcpyAsync(logdet_host, op.logdet, 2 * sizeof(double) * nt, hipMemcpyDeviceToHost, s));
    ISLE_HIP(hipStreamSynchronize(s));
}

// solveQ (hubbardFermiMatrixExp.cpp:494-537) on the blocks of Q's LU, as one cached operation
// per (nt, number of right-hand sides) whose workspace holds the blocks in split planes and the
// right-hand sides as columns:
//   forward   y_0 = b_0,  y_i = b_i - l_{i-1} y_{i-1} (i >= 1),  y_{nt-1} -= sum_{i<nt-2} h_i y_i
//   backward  x_{nt-1} = d_{nt-1}^{-1} y_{nt-1},  x_{nt-2} = d_{nt-2}^{-1} (y_{nt-2} - u_{nt-2} x_{nt-1}),
//             x_i = d_i^{-1} (y_i - u_i x_{i+1} - v_i x_{nt-1})  (i = nt-3 .. 0)
// with every product a complex GEMM of four real products (as in getQLU); the h_i side by side
// make the sum one GEMM with k = (nt-2) nx, and the v_i x_{nt-1} are one launch over all i.
OpCache& Plan::get_solveq(int nt, int nrhs) {
    if (nt < 1 || nrhs < 1) throw std::invalid_argument("solveQ needs nt >= 1 and at least one right-hand side");
    auto key = std::make_tuple(8, nt, nrhs, 0, 0, 0);
    auto found = cache_.find(key);
    if (found != cache_.end()) return *found->second;
    auto op = std::make_unique<OpCache>();
    op->nt = nt;
    op->batch = 1;
    op->phi_stride = 0;
    Builder b(*op);
    const int n = nx_;
    const long nn = (long)n * n;
    const int nu = std::max(nt - 1, 0), nh = std::max(nt - 2, 0);
    const long ldh = (long)std::max(nh, 1) * n;
    double* DI = b.d64((size_t)nt * 2 * nn);
    double* UU = b.d64((size_t)std::max(nu, 1) * 2 * nn);
    double* LL = b.d64((size_t)std::max(nu, 1) * 2 * nn);
    double* VV = b.d64((size_t)std::max(nh, 1) * 2 * nn);
    double* HH = b.d64((size_t)std::max(nh, 1) * 2 * nn);  // n x (nh n): re plane, then im plane
    const long plane = (long)nt * n * nrhs;                // right-hand sides as columns, row stride nrhs
    double* Y = b.d64(2 * (size_t)plane);
    double* X = b.d64(2 * (size_t)plane);
    op->extra = {DI, UU, LL, VV, HH, Y, X};
    auto blkm = [&](double* base, int k) { return cmat(base, k, nn); };
    auto col = [&](double* base, int t) {
        return CMat<double>{base + (long)t * n * nrhs, base + plane + (long)t * n * nrhs};
    };
    auto gi = [&](CMat<double> a, long lda, CMat<double> bb, CMat<double> c, int k) {
        GemmItem<double> g{};
        g.a_re = a.re;
        g.a_im = a.im;
        g.b_re = bb.re;
        g.b_im = bb.im;
        g.c_re = c.re;
        g.c_im = c.im;
        g.m = n;
        g.n = nrhs;
        g.k = k;
        g.lda = (int)lda;
        g.ldb = nrhs;
        g.ldc = nrhs;
        return g;
    };
    auto one = [&](K kind, GemmItem<double> g) {
        std::vector<GemmItem<double>> v = {g};
        b.add(kind, v, n, nrhs);
    };
    for (int i = 1; i < nt; ++i) one(K::GEMM64_CC4_SUB, gi(blkm(LL, i - 1), n, col(Y, i - 1), col(Y, i), n));
    if (nt > 2)
        one(K::GEMM64_CC4_SUB, gi(CMat<double>{HH, HH + (long)nh * nn}, ldh, col(Y, 0), col(Y, nt - 1), nh * n));
    one(K::GEMM64_CC4, gi(blkm(DI, nt - 1), n, col(Y, nt - 1), col(X, nt - 1), n));
    if (nt > 1) {
        one(K::GEMM64_CC4_SUB, gi(blkm(UU, nt - 2), n, col(X, nt - 1), col(Y, nt - 2), n));
        one(K::GEMM64_CC4, gi(blkm(DI, nt - 2), n, col(Y, nt - 2), col(X, nt - 2), n));
    }
    if (nt > 2) {
        std::vector<GemmItem<double>> vx;
        for (int i = 0; i < nh; ++i) vx.push_back(gi(blkm(VV, i), n, col(X, nt - 1), col(Y, i), n));
        b.add(K::GEMM64_CC4_SUB, vx, n, nrhs);
        for (int i = nt - 3; i >= 0; --i) {
            one(K::GEMM64_CC4_SUB, gi(blkm(UU, i), n, col(X, i + 1), col(Y, i), n));
            one(K::GEMM64_CC4, gi(blkm(DI, i), n, col(Y, i), col(X, i), n));
        }
    }
    b.finish();
    tune(*op);
    auto& ref = *op;
    cache_[key] = std::move(op);
    return ref;
}

void Plan::solve_q(int nt, int nrhs, const double* dinv, const double* u, const double* l, const double* v,
                   const double* h, const double* rhs, double* x) {
    OpCache& op = get_solveq(nt, nrhs);
    const int n = nx_;
    const long nn = (long)n * n, plane = (long)nt * n * nrhs;
    const int nh = std::max(nt - 2, 0);
    hipStream_t s = stream_;
    const auto& e = op.extra;  // DI, UU, LL, VV, HH, Y, X
    auto in = [&](const double* src, int nblk, double* re, double* im, long boff, long ld) {
        if (nblk <= 0) return;
        hipLaunchKernelGGL(interleaved_to_blocks_kernel, dim3(std::min<long>((nn + 255) / 256, 1024), nblk), dim3(256),
                           0, s, src, boff, ld, n, re, im);
        ISLE_HIP(hipGetLastError());
    };
    in(dinv, nt, e[0], e[0] + nn, 2 * nn, n);
    in(u, nt - 1, e[1], e[1] + nn, 2 * nn, n);
    in(l, nt - 1, e[2], e[2] + nn, 2 * nn, n);
    in(v, nt - 2, e[3], e[3] + nn, 2 * nn, n);
    in(h, nt - 2, e[4], e[4] + (long)nh * nn, n, (long)nh * n);  // h_i side by side
    const unsigned g = (unsigned)std::min<long>((plane + 255) / 256, 4096);
    hipLaunchKernelGGL(rhs_columns_kernel, dim3(g), dim3(256), 0, s, const_cast<double*>(rhs), e[5], e[5] + plane,
                       (long)nt * n, nrhs, 1);
    ISLE_HIP(hipGetLastError());
    run(op, nullptr);
    hipLaunchKernelGGL(rhs_columns_kernel, dim3(g), dim3(256), 0, s, x, e[6], e[6] + plane, (long)nt * n, nrhs, 0);
    ISLE_HIP(hipGetLastError());
}

// M x for rows x: per time slice one complex x real GEMM over all rows with the slice
// matrix S (EXP: exp(kappa - mu)..., DIA: K), between an elementwise prologue (the shifted,
// phase-scaled rows) and epilogue (back to interleaved rows); see applym_prep_kernel.
void Plan::applyM(const double* phi, int nt, int species, const double* x, long nrhs, double* out) {
    const int n = nx_;
    if (nt < 1 || nrhs < 1) throw std::invalid_argument("applyM needs nt >= 1 and at least one row");
    if (nrhs > (1L << 30) / ((long)nt * n)) throw std::invalid_argument("applyM: too many rows for one call");
    hipStream_t s = stream_;
    const long plane = nrhs * nt * (long)n;
    const size_t need = (size_t)4 * plane;  // two split-plane arrays
    if (need > apply_ws_doubles_) {
        apply_ws_.reset();
        apply_ws_.reset(new DevBuf(need * sizeof(double)));
        apply_ws_doubles_ = need;
    }
    double* ya = apply_ws_->as<double>();
    double* c = ya + 2 * plane;
    const bool exp_hop = hop_ == EXP;
    const double sgn = species == PARTICLE ? 1.0 : -1.0;
    const unsigned blocks = (unsigned)std::min<long>((plane + 255) / 256, 16384);
    hipLaunchKernelGGL(applym_prep_kernel, dim3(blocks), dim3(256), 0, s, x, phi, n, nt, nrhs, sgn, (int)exp_hop, ya,
                       c);
    ISLE_HIP(hipGetLastError());
    if (apply_items_n_ < nt) {
        apply_items_.reset();
        apply_items_.reset(new DevBuf((size_t)nt * sizeof(GemmItem<double>)));
        apply_items_n_ = nt;
    }
    const double* S = exp_hop ? e64_[species]->as<double>() : a64_[species]->as<double>();
    std::vector<GemmItem<double>> items(nt);
    for (int t = 0; t < nt; ++t) {
        GemmItem<double>& g = items[t];
        g = GemmItem<double>{};
        g.a_re = ya + (long)t * nrhs * n;
        g.a_im = ya + plane + (long)t * nrhs * n;
        g.b_re = S;
        g.c_re = c + (long)t * nrhs * n;
        g.c_im = c + plane + (long)t * nrhs * n;
        g.m = (int)nrhs;
        g.n = n;
        g.k = n;
        g.lda = g.ldb = g.ldc = n;
    }
    ISLE_HIP(hipMemcpyAsync(apply_items_->p, items.data(), nt * sizeof(GemmItem<double>), hipMemcpyHostToDevice, s));
    const auto* d_items = apply_items_->as<GemmItem<double>>();
    const bool ok = exp_hop ? launch_gemm<double, true, false, false, SUBTRACT>(d_items, nt, (int)nrhs, n, s, 0)
                            : launch_gemm<double, true, false, false, SCALE_NONE>(d_items, nt, (int)nrhs, n, s, 0);
    if (!ok) throw std::runtime_error("applyM: GEMM tile does not fit");
    hipLaunchKernelGGL(applym_finish_kernel, dim3(blocks), dim3(256), 0, s, x, phi, n, nt, nrhs, sgn, (int)exp_hop, c,
                       out);
    ISLE_HIP(hipGetLastError());
    // the descriptors above live on the host stack: wait before they go
    ISLE_HIP(hipStreamSynchronize(s));
}

double* Plan::solve_workspace(size_t doubles) {
    if (doubles > solve_ws_doubles_) {
        solve_ws_.reset();  // free before allocating the larger buffer
        solve_ws_.reset(new DevBuf(doubles * sizeof(double)));
        solve_ws_doubles_ = doubles;
    }
    return solve_ws_->as<double>();
}

// solveM (hubbardFermiMatrixExp.cpp:645-694, hubbardFermiMatrixDia.cpp:617-673).
void Plan::solveM(const double* phi, int nt, int species, const double* rhss, int nrhs, double* out, bool reuse) {
    const int n = nx_;
    const long nn = (long)n * n;
    hipStream_t s = stream_;
    const bool refactor = !(reuse && sf_.nt == nt && sf_.species == species && sf_.part);
    if (refactor) {
        sf_ = SolveFactor{};
        sf_.ph.reset(new DevBuf((size_t)2 * nt * n * sizeof(double)));
        sf_.part.reset(new DevBuf((size_t)nt * 2 * nn * sizeof(double)));
        sf_.mz.reset(new DevBuf((size_t)nt * 2 * nn * sizeof(double)));
        sf_.piv.reset(new DevBuf((size_t)nt * n * sizeof(int)));
        sf_.luw.reset(new DevBuf((size_t)nt * lu_work_doubles(n) * sizeof(double)));
    }
    double* ph = sf_.ph->as<double>();
    double* part = sf_.part->as<double>();
    double* Mz = sf_.mz->as<double>();  // matrices multiplying the rhs (EXP: a copy of the partial products)
    int* piv = sf_.piv->as<int>();
    double* luw_p = sf_.luw->as<double>();
    // right-hand-side buffers: Bt, Z (nt blocks of 2 n nrhs), Xl and a scratch block
    const size_t rblk = (size_t)2 * n * nrhs;
    double* Bt = solve_workspace((2 * (size_t)nt + 2) * rblk);
    double* Z = Bt + (size_t)nt * rblk;
    double* Xl = Z + (size_t)nt * rblk;
    double* Sc = Xl + rblk;
    auto D = [&](int t) { return CMat<double>{ph + (long)t * n, ph + (long)nt * n + (long)t * n}; };
    auto mP = [&](int t) { return cmat(part, t, nn); };
    auto mM = [&](int t) { return cmat(Mz, t, nn); };
    if (refactor) {
        OpCache op;
        op.nt = nt;
        op.phi_stride = 0;
        Builder b(op);
        std::vector<PhaseItem> phs{{0, 0, sigma_of(species), D(0).re, D(0).im, nullptr, nullptr, n, nt}};
        b.add(K::PHASE, phs);
        const double* A = a64_[species]->as<double>();
        std::vector<ScaleItem<double>> sc{{A, D(0).re, D(0).im, mP(0).re, mP(0).im, n, SCALE_ROW}};
        if (hop_ == DIA) sc.push_back({ident64_->as<double>(), D(0).re, D(0).im, mM(0).re, mM(0).im, n, SCALE_ROW});
        b.add(K::SCALE64, sc);
        for (int t = 1; t < nt; ++t) {
            std::vector<GemmItem<double>> gi{gitem<double>(mP(t - 1).re, mP(t - 1).im, A, nullptr, mP(t), D(t).re,
                                                           D(t).im, nullptr, nullptr, n)};
            b.add(K::GEMM64_CR_PRE, gi, n, n);
        }
        b.finish();
        run(op, phi);
        if (hop_ == DIA) {
            // W_t = part_{t-1} D_t (the product without the final K)
            std::vector<CScaleItem<double>> cs;
            for (int t = 1; t < nt; ++t)
                cs.push_back({mP(t - 1).re, mP(t - 1).im, D(t).re, D(t).im, mM(t).re, mM(t).im, n, SCALE_COL});
            if (!cs.empty()) {
                DevBuf dcs(cs.size() * sizeof(CScaleItem<double>));
                ISLE_HIP(hipMemcpy(dcs.p, cs.data(), cs.size() * sizeof(cs[0]), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(scale_complex_kernel<double>, dim3(ew_blocks(nn), cs.size()), dim3(EW_THREADS), 0,
                                   s, dcs.as<CScaleItem<double>>());
                ISLE_HIP(hipGetLastError());
            }
        } else {
            ISLE_HIP(hipMemcpyAsync(Mz, part, (size_t)nt * 2 * nn * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        // LU of every partial product (+1 on the last), in place; Mz keeps the products
        std::vector<LuItem> lu;
        for (int t = 0; t < nt; ++t)
            lu.push_back({mP(t).re, mP(t).im, piv + (long)t * n, nullptr, luw_p + t * lu_work_doubles(n), n, n});
        DevBuf dlu(lu.size() * sizeof(LuItem));
        ISLE_HIP(hipMemcpy(dlu.p, lu.data(), lu.size() * sizeof(LuItem), hipMemcpyHostToDevice));
        if (nt > 1) launch_lu(dlu.as<LuItem>(), nt - 1, n, 0, s);
        launch_lu(dlu.as<LuItem>() + (nt - 1), 1, n, 1, s);
        ISLE_HIP(hipStreamSynchronize(s));
        sf_.nt = nt;
        sf_.species = species;
    }
    const long blk = 2L * n * nrhs;
    // The slice products of the right-hand sides and the blocked triangular solves are launch
    // steps with tuned tiles (once per shape: the plan caches the choice).  Tuning runs them
    // on the workspace before it holds this call's data; the factorization is only read.
    OpCache og;
    og.nt = nt;
    {
        Builder bg(og);
        std::vector<GemmItem<double>> gi;
        for (int t = 0; t < nt; ++t) {
            GemmItem<double> g{};
            g.a_re = mM(t).re;
            g.a_im = mM(t).im;
            g.b_re = Bt + t * blk;
            g.b_im = Bt + t * blk + (long)n * nrhs;
            g.c_re = Z + t * blk;
            g.c_im = Z + t * blk + (long)n * nrhs;
            g.m = n;
            g.n = nrhs;
            g.k = n;
            g.lda = n;
            g.ldb = nrhs;
            g.ldc = nrhs;
            gi.push_back(g);
        }
        bg.add(K::GEMM64_CC, gi, n, nrhs);
        bg.finish();
        tune(og);
    }
    // many right-hand sides: the triangular solves as blocked GEMMs (rows gathered into
    // each LU's pivot order, then unit-lower and upper solves with the stored inverses of
    // the 32 x 32 diagonal blocks): o1 solves the last slice, o2 every slice
    const bool blocked = n % LU_NB == 0 && nrhs >= 64;
    OpCache o1, o2;
    o1.nt = o2.nt = nt;
    if (blocked) {
        auto lu_mat = [&](int t, double* rre, double* rim, double* xre, double* xim) {
            return TrsmMat{mP(t).re, mP(t).im, n, luw_p + t * lu_work_doubles(n), rre, rim, nrhs, xre, xim, nrhs};
        };
        auto rows = [&](double* p) { return std::make_pair(p, p + (long)n * nrhs); };
        Builder b1(o1);
        int* sig = b1.di((size_t)nt * n);  // o1 outlives o2, which reads it
        std::vector<SigmaItem> sv;
        for (int t = 0; t < nt; ++t) sv.push_back({piv + (long)t * n, sig + (long)t * n, n, n});
        b1.add(K::SIGMA, sv);
        auto z = rows(Z + (nt - 1) * blk), g0 = rows(Bt), g1 = rows(Sc), xl = rows(Xl);
        std::vector<RowGatherItem> rg{{sig + (long)(nt - 1) * n, z.first, z.second, g0.first, g0.second, nrhs, nrhs, n,
                                       nrhs}};
        b1.add(K::RGATHER, rg, n, 0, 0, nrhs);
        add_trsm(b1, 0, n, nrhs, {lu_mat(nt - 1, g0.first, g0.second, g1.first, g1.second)});
        add_trsm(b1, 1, n, nrhs, {lu_mat(nt - 1, g1.first, g1.second, xl.first, xl.second)});
        b1.finish();
        Builder b2(o2);
        std::vector<RowGatherItem> ra;
        std::vector<TrsmMat> tl, tu;
        for (int t = 0; t < nt; ++t) {
            auto zt = rows(Z + t * blk), bt = rows(Bt + t * blk);
            ra.push_back({sig + (long)t * n, zt.first, zt.second, bt.first, bt.second, nrhs, nrhs, n, nrhs});
            tl.push_back(lu_mat(t, bt.first, bt.second, zt.first, zt.second));
            tu.push_back(lu_mat(t, zt.first, zt.second, bt.first, bt.second));
        }
        b2.add(K::RGATHER, ra, n, 0, 0, nrhs);
        add_trsm(b2, 0, n, nrhs, tl);
        add_trsm(b2, 1, n, nrhs, tu);
        b2.finish();
        tune(o1);
        tune(o2);
    }
    const dim3 tgrid(cdiv(n, 32), cdiv(nrhs, 32), nt);
    hipLaunchKernelGGL(rhs_transpose_kernel, tgrid, dim3(256), 0, s, rhss, Bt, n, nt, nrhs);
    ISLE_HIP(hipGetLastError());
    run(og, phi);
    hipLaunchKernelGGL(prefix_sum_t_kernel, dim3(ew_blocks(blk)), dim3(EW_THREADS), 0, s, Z, blk, nt);
    ISLE_HIP(hipGetLastError());
    if (blocked) {
        run(o1, phi);  // x_{nt-1} = (1 + A^-1)^-1 z_{nt-1}
        if (nt > 1) {
            hipLaunchKernelGGL(sub_last_kernel, dim3(ew_blocks(blk * (nt - 1))), dim3(EW_THREADS), 0, s, Z, Xl, blk,
                               nt - 1);
            ISLE_HIP(hipGetLastError());
        }
        run(o2, phi);  // x_t = A_t^-1 (z_t - x_{nt-1}) (z_{nt-1} unchanged: solves x_{nt-1} again)
        hipLaunchKernelGGL(rhs_untranspose_kernel, tgrid, dim3(256), 0, s, Bt, out, n, nt, nrhs);
        ISLE_HIP(hipGetLastError());
        ISLE_HIP(hipStreamSynchronize(s));
        return;
    }
    ISLE_HIP(hipStreamSynchronize(s));
    auto sitem = [&](int t, const double* bre, long bsr, long bsc, double* xre, double* xim, long xsr, long xsc) {
        SolveItem si{};
        si.lu_re = mP(t).re;
        si.lu_im = mP(t).im;
        si.piv = piv + (long)t * n;
        si.dinv = luw_p + t * lu_work_doubles(n);
        si.b_re = bre;
        si.b_im = bre + (long)n * nrhs;
        si.bsr = bsr;
        si.bsc = bsc;
        si.x_re = xre;
        si.x_im = xim;
        si.xsr = xsr;
        si.xsc = xsc;
        si.n = n;
        si.ld = n;
        si.nrhs = nrhs;
        return si;
    };
    std::vector<SolveItem> s1{sitem(nt - 1, Z + (nt - 1) * blk, nrhs, 1, Xl, Xl + (long)n * nrhs, nrhs, 1)};
    DevBuf ds1(sizeof(SolveItem));
    ISLE_HIP(hipMemcpy(ds1.p, s1.data(), sizeof(SolveItem), hipMemcpyHostToDevice));
    launch_lu_solve(ds1.as<SolveItem>(), 1, n, nrhs, s);
    if (nt > 1) {
        hipLaunchKernelGGL(sub_last_kernel, dim3(ew_blocks(blk * (nt - 1))), dim3(EW_THREADS), 0, s, Z, Xl, blk, nt - 1);
        ISLE_HIP(hipGetLastError());
    }
    std::vector<SolveItem> sv;
    for (int t = 0; t < nt; ++t)
        sv.push_back(sitem(t, Z + t * blk, nrhs, 1, out + 2L * t * n, out + 2L * t * n + 1, 2, 2L * nt * n));
    // the last slice solves the original z_{nt-1} again (Z was not modified there)
    DevBuf dsv(sv.size() * sizeof(SolveItem));
    ISLE_HIP(hipMemcpy(dsv.p, sv.data(), sv.size() * sizeof(SolveItem), hipMemcpyHostToDevice));
    launch_lu_solve(dsv.as<SolveItem>(), nt, n, nrhs, s);
    ISLE_HIP(hipStreamSynchronize(s));
}

void Plan::F(int tp, const double* phi, int nt, int species, int inv, double* out) {
    const int n = nx_;
    const long nn = (long)n * n;
    if (tp < 0 || tp >= nt) throw std::out_of_range("time slice out of range");
    // sign of the phase (hubbardFermiMatrixExp.cpp:128-131, hubbardFermiMatrixDia.cpp:127-130)
    const bool plus = (species == PARTICLE && !inv) || (species == HOLE && inv);
    OpCache op;
    op.nt = nt;
    Builder b(op);
    double* ph = b.d64((size_t)2 * nt * n);
    double* m = b.d64((size_t)2 * nn);
    const double* D_re = ph + (long)tp * n;
    const double* D_im = ph + (long)nt * n + (long)tp * n;
    std::vector<PhaseItem> phs{{0, 0, plus ? 1.0 : -1.0, ph, ph + (long)nt * n, nullptr, nullptr, n, nt}};
    b.add(K::PHASE, phs);
    std::vector<ScaleItem<double>> sc;
    if (hop_ == EXP)
        sc.push_back({inv ? a64_[species]->as<double>() : e64_[species]->as<double>(), D_re, D_im, m, m + nn, n,
                      inv ? SCALE_ROW : SCALE_COL});
    else
        sc.push_back({ident64_->as<double>(), D_re, D_im, m, m + nn, n, SCALE_ROW});
    b.add(K::SCALE64, sc);
    b.finish();
    run(op, phi);
    hipLaunchKernelGGL(split_to_interleaved_kernel, dim3(ew_blocks(nn)), dim3(EW_THREADS), 0, stream_, m, m + nn, out,
                       nn);
    ISLE_HIP(hipGetLastError());
    ISLE_HIP(hipStreamSynchronize(stream_));
}

}  // namespace isle

// ===========================================================================
// C ABI
// ===========================================================================
namespace {
thread_local std::string g_last_error;

// Device scratch of the one-matrix LU entry points (isle_logdet, isle_invert,
// isle_invert_logdet): the split copy of the matrix, the LU's work buffer, pivots, the two
// launch descriptors and the log det, in one allocation at 256-byte aligned offsets.  Up to
// DENSE_CACHE_MAX bytes it is kept per device between calls (a hipMalloc / hipFree pair per
// call costs more than the LU of a small block, and hipFree waits for the device); a lease
// holds a lock for the whole call, since every call on a device shares that buffer.  Larger
// orders allocate for the call.  Every field is written by the call before it is read.
constexpr size_t DENSE_CACHE_MAX = size_t(64) << 20;
std::mutex dense_mu;
std::map<int, std::pair<void*, size_t>> dense_cache;  // device -> (buffer, bytes), kept to exit

// Device scratch of one dense LU / inverse call, held under dense_mu for the call.  The
// per-device cache is shared by every call: if the call throws after enqueueing work on
// `s`, the destructor waits for that work before the lock is released, so the next call
// (on any stream) cannot overlap kernels still writing the buffer.
struct DenseScratch {
    double *m, *w, *ld;
    int* piv;
    isle::LuItem* item;
    isle::SolveItem* sitem;
    DenseScratch(int n, hipStream_t s) : lock_(dense_mu), stream_(s), unwinding_(std::uncaught_exceptions()) {
        auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t nn = (size_t)n * n;
        size_t off[6], o = 0;
        off[0] = o; o = al(o + 2 * nn * sizeof(double));
        off[1] = o; o = al(o + isle::lu_work_doubles(n) * sizeof(double));
        off[2] = o; o = al(o + (size_t)n * sizeof(int));
        off[3] = o; o = al(o + sizeof(isle::LuItem));
        off[4] = o; o = al(o + sizeof(isle::SolveItem));
        off[5] = o; o = al(o + 2 * sizeof(double));
        char* base;
        if (o <= DENSE_CACHE_MAX) {
            int dev = 0;
            ISLE_HIP(hipGetDevice(&dev));
            auto& e = dense_cache[dev];
            if (e.second < o) {
                if (e.first) ISLE_HIP(hipFree(e.first));
                e = {nullptr, 0};
                ISLE_HIP(hipMalloc(&e.first, o));
                e.second = o;
            }
            base = static_cast<char*>(e.first);
        } else {
            own_ = std::make_unique<isle::DevBuf>(o);
            base = static_cast<char*>(own_->p);
        }
        m = reinterpret_cast<double*>(base + off[0]);
        w = reinterpret_cast<double*>(base + off[1]);
        piv = reinterpret_cast<int*>(base + off[2]);
        item = reinterpret_cast<isle::LuItem*>(base + off[3]);
        sitem = reinterpret_cast<isle::SolveItem*>(base + off[4]);
        ld = reinterpret_cast<double*>(base + off[5]);
    }

    ~DenseScratch() {
        if (std::uncaught_exceptions() > unwinding_) (void)hipStreamSynchronize(stream_);
    }

private:
    std::unique_lock<std::mutex> lock_;
    hipStream_t stream_;
    int unwinding_;
    std::unique_ptr<isle::DevBuf> own_;
};

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return ISLE_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return ISLE_ERR_VALUE;
    } catch (const std::out_of_range& e) {
        g_last_error = e.what();
        return ISLE_ERR_INDEX;
    } catch (const isle::HipError& e) {
        g_last_error = e.what();
        return ISLE_ERR_HIP;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return ISLE_ERR_RUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return ISLE_ERR_RUNTIME;
    }
}

isle::Plan* P(isle_hfm* h) {
    if (!h) throw std::invalid_argument("null fermion-matrix handle");
    return reinterpret_cast<isle::Plan*>(h);
}

void check_species(int s) {
    if (s != 0 && s != 1) throw std::invalid_argument("Unknown species");
}
}  // namespace

extern "C" {

const char* isle_last_error(void) { return g_last_error.c_str(); }

int isle_hfm_create(isle_hfm** out, int hopping, int nx, const double* kappa_tilde, double mu_tilde,
                    int sigma_kappa) {
    return guarded([&] {
        if (!out || !kappa_tilde) throw std::invalid_argument("null argument");
        *out = reinterpret_cast<isle_hfm*>(new isle::Plan(hopping, nx, kappa_tilde, mu_tilde, sigma_kappa));
    });
}

int isle_hfm_destroy(isle_hfm* h) {
    return guarded([&] { delete reinterpret_cast<isle::Plan*>(h); });
}

int isle_hfm_set_stream(isle_hfm* h, void* stream) {
    return guarded([&] { P(h)->set_stream(reinterpret_cast<hipStream_t>(stream)); });
}

int isle_hfm_set_recompute(isle_hfm* h, int k) {
    return guarded([&] { P(h)->set_recompute(k); });
}

int isle_hfm_set_tile(isle_hfm* h, int kind, int tile) {
    return guarded([&] { P(h)->set_tile(kind, tile); });
}

int isle_hfm_slice_matrix(isle_hfm* h, int species, int inv, double* out_host) {
    return guarded([&] {
        check_species(species);
        P(h)->slice_matrix(species, inv, out_host);
    });
}

int isle_hfm_F(isle_hfm* h, int tp, const double* phi, int nt, int species, int inv, double* out) {
    return guarded([&] {
        check_species(species);
        P(h)->F(tp, phi, nt, species, inv, out);
    });
}

int isle_logdetM(isle_hfm* h, const double* phi, int nt, int species, double* out_host) {
    return guarded([&] {
        check_species(species);
        double tmp[4];
        P(h)->logdetM(phi, nt, 1, 0, 0, 2, 0, tmp);
        out_host[0] = tmp[2 * species];
        out_host[1] = tmp[2 * species + 1];
    });
}

int isle_logdetM_batch(isle_hfm* h, const double* phi, int nt, int batch, int mode, int group, int nspecies,
                       int spin, double* out_host) {
    return guarded([&] {
        if (batch < 1 || (nspecies != 1 && nspecies != 2)) throw std::invalid_argument("bad batch/nspecies");
        if (mode < 0 || mode > 2) throw std::invalid_argument("mode must be 0, 1 or 2");
        P(h)->logdetM(phi, nt, batch, mode, group, nspecies, spin, out_host);
    });
}

int isle_force_parts(isle_hfm* h, const double* phi, int nt, int batch, int mode, int group, int nspecies, int spin,
                     double* out_dev) {
    return guarded([&] {
        if (batch < 1 || (nspecies != 1 && nspecies != 2)) throw std::invalid_argument("bad batch/nspecies");
        if (mode < 0 || mode > 3) throw std::invalid_argument("mode must be 0, 1, 2 or 3");
        P(h)->force_parts(phi, nt, batch, mode, group, nspecies, spin, out_dev);
    });
}

int isle_solveq(isle_hfm* h, int nt, int nrhs, const double* dinv, const double* u, const double* l,
                const double* v, const double* hh, const double* rhs, double* x) {
    return guarded([&] {
        if (nt < 1 || nrhs < 1) throw std::invalid_argument("isle_solveq needs nt >= 1 and nrhs >= 1");
        P(h)->solve_q(nt, nrhs, dinv, u, l, v, hh, rhs, x);
    });
}

int isle_qlu(isle_hfm* h, const double* phi, int nt, double* dinv, double* u, double* l, double* v, double* hh,
             double* logdet_host) {
    return guarded([&] {
        if (nt < 3) throw std::invalid_argument("isle_qlu needs nt >= 3");
        P(h)->qlu(phi, nt, dinv, u, l, v, hh, logdet_host);
    });
}

int isle_force_parts_multi(isle_hfm* h, int nops, const double* const* phis, const int* nts, const int* batches,
                           int mode, const int* groups, int nspecies, int spin, double* const* outs) {
    return guarded([&] {
        if (nops < 1 || (nspecies != 1 && nspecies != 2)) throw std::invalid_argument("bad nops/nspecies");
        if (mode < 0 || mode > 3) throw std::invalid_argument("mode must be 0, 1, 2 or 3");
        P(h)->force_parts_multi(nops, phis, nts, batches, mode, groups, nspecies, spin, outs);
    });
}

int isle_solveM(isle_hfm* h, const double* phi, int nt, int species, const double* rhss, int nrhs, double* out) {
    return guarded([&] {
        check_species(species);
        if (nrhs < 1) throw std::invalid_argument("need at least one right hand side");
        P(h)->solveM(phi, nt, species, rhss, nrhs, out);
    });
}

int isle_solveM_ex(isle_hfm* h, const double* phi, int nt, int species, const double* rhss, int nrhs, double* out,
                   int reuse) {
    return guarded([&] {
        check_species(species);
        if (nrhs < 1) throw std::invalid_argument("need at least one right hand side");
        P(h)->solveM(phi, nt, species, rhss, nrhs, out, reuse != 0);
    });
}

int isle_applyM(isle_hfm* h, const double* phi, int nt, int species, const double* x, long nrhs, double* out) {
    return guarded([&] {
        check_species(species);
        P(h)->applyM(phi, nt, species, x, nrhs, out);
    });
}

int isle_profile_kernels(isle_hfm* h, const double* phi, int op, int nt, int batch, int group, int nspecies,
                         int spin, int reps, double* out_host, int nkinds) {
    return guarded([&] {
        if (op < 0 || op > 5 || reps < 1) throw std::invalid_argument("bad profile request");
        P(h)->profile(phi, op, nt, batch, group, nspecies, spin, reps, out_host, nkinds);
    });
}

int isle_hfm_sync(isle_hfm* h) {
    return guarded([&] { P(h)->sync(); });
}

int isle_combine_force(const double* fp, const double* fh, double* out, long n, int mode, int accumulate,
                       void* stream) {
    return guarded([&] {
        hipLaunchKernelGGL(isle::combine_force_kernel, dim3(isle::ew_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                           fp, fh, out, n, 0L, 0L, mode, accumulate);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_combine_force_batch(const double* fp, const double* fh, double* out, long n, int batch, long in_stride,
                             long out_stride, int mode, int accumulate, void* stream) {
    return guarded([&] {
        if (batch < 1 || batch > 65535) throw std::invalid_argument("batch must be in 1..65535");
        hipLaunchKernelGGL(isle::combine_force_kernel, dim3(isle::ew_blocks(n), batch), dim3(256), 0,
                           (hipStream_t)stream, fp, fh, out, n, in_stride, out_stride, mode, accumulate);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_gauge_force(const double* phi, double* out, long n, double utilde, int accumulate, void* stream) {
    return guarded([&] {
        hipLaunchKernelGGL(isle::gauge_force_kernel, dim3(isle::ew_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                           phi, out, n, 1.0 / utilde, accumulate);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_field_sums(const double* phi, long n, long stride, int batch, double* out_host, void* stream) {
    return guarded([&] {
        isle::DevBuf d((size_t)batch * 4 * 8);
        hipLaunchKernelGGL(isle::sums_kernel, dim3(batch), dim3(256), 0, (hipStream_t)stream, phi, n, stride,
                           d.as<double>());
        ISLE_HIP(hipGetLastError());
        ISLE_HIP(hipMemcpyAsync(out_host, d.p, (size_t)batch * 4 * 8, hipMemcpyDeviceToHost, (hipStream_t)stream));
        ISLE_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int isle_leapfrog_kick(double* pi, const double* force, const double* phi, long n, double eps, double utilde,
                       void* stream) {
    return guarded([&] {
        const double iu = (phi && utilde > 0) ? 1.0 / utilde : 0.0;
        hipLaunchKernelGGL(isle::kick_kernel, dim3(isle::ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, pi, force,
                           iu > 0 ? phi : nullptr, n, eps, iu);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_leapfrog_drift(double* phi, const double* pi, long n, double eps, void* stream) {
    return guarded([&] {
        hipLaunchKernelGGL(isle::drift_kernel, dim3(isle::ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, phi, pi, n,
                           eps);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_cscale(const double* in, double* out, long n, double cr, double ci, void* stream) {
    return guarded([&] {
        hipLaunchKernelGGL(isle::cscale_vec_kernel, dim3(isle::ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, in,
                           out, n, cr, ci);
        ISLE_HIP(hipGetLastError());
    });
}

int isle_logdet(const double* a, int n, double* out_host, void* stream) {
    return guarded([&] {
        if (n < 1) throw std::invalid_argument("empty matrix");
        if (n > isle::LU_TALL_MAX) throw std::invalid_argument("matrix too large for the LU kernel (n > 16383)");
        hipStream_t s = (hipStream_t)stream;
        const long nn = (long)n * n;
        DenseScratch sc(n, s);
        hipLaunchKernelGGL(isle::interleaved_to_split_kernel, dim3(isle::ew_blocks(nn)), dim3(256), 0, s, a, sc.m,
                           sc.m + nn, nn);
        ISLE_HIP(hipGetLastError());
        isle::LuItem it{sc.m, sc.m + nn, sc.piv, sc.ld, sc.w, n, n};
        ISLE_HIP(hipMemcpyAsync(sc.item, &it, sizeof(it), hipMemcpyHostToDevice, s));
        isle::launch_lu(sc.item, 1, n, 0, s);
        ISLE_HIP(hipMemcpyAsync(out_host, sc.ld, 16, hipMemcpyDeviceToHost, s));
        ISLE_HIP(hipStreamSynchronize(s));
    });
}

// A^{-1} by one LU with partial pivoting; with logdet_host set, also log det A from the same LU.
static int invert_impl(const double* a, int n, double* out, double* logdet_host, void* stream) {
    return guarded([&] {
        if (n < 1) throw std::invalid_argument("empty matrix");
        if (n > isle::LU_TALL_MAX) throw std::invalid_argument("matrix too large for the LU kernel (n > 16383)");
        hipStream_t s = (hipStream_t)stream;
        const long nn = (long)n * n;
        DenseScratch sc(n, s);
        hipLaunchKernelGGL(isle::interleaved_to_split_kernel, dim3(isle::ew_blocks(nn)), dim3(256), 0, s, a, sc.m,
                           sc.m + nn, nn);
        ISLE_HIP(hipGetLastError());
        isle::LuItem it{sc.m, sc.m + nn, sc.piv, logdet_host ? sc.ld : nullptr, sc.w, n, n};
        ISLE_HIP(hipMemcpyAsync(sc.item, &it, sizeof(it), hipMemcpyHostToDevice, s));
        isle::launch_lu(sc.item, 1, n, 0, s);
        isle::SolveItem si{};
        si.lu_re = sc.m;
        si.lu_im = sc.m + nn;
        si.piv = sc.piv;
        si.dinv = sc.w;
        si.x_re = out;
        si.x_im = out + 1;
        si.xsr = 2L * n;
        si.xsc = 2;
        si.n = n;
        si.ld = n;
        si.nrhs = n;
        ISLE_HIP(hipMemcpyAsync(sc.sitem, &si, sizeof(si), hipMemcpyHostToDevice, s));
        isle::launch_lu_solve(sc.sitem, 1, n, n, s);
        if (logdet_host) ISLE_HIP(hipMemcpyAsync(logdet_host, sc.ld, 16, hipMemcpyDeviceToHost, s));
        ISLE_HIP(hipStreamSynchronize(s));
    });
}

// ---- log det of one matrix without allocating: the scratch of the LU is a static device
// array (one copy per GPU, held until the process ends), found per device with
// hipGetSymbolAddress on the first call there.  Calls in one process are serialised by a
// lock and return after their stream is synchronised; calls that overlap on one GPU from
// other threads without that lock would share the array and race.
static constexpr int NOALLOC_MAX_N = 512;
static constexpr size_t noalloc_align(size_t b) { return (b + 255) & ~size_t(255); }
static constexpr size_t NOALLOC_WORK =
    (size_t)isle::LU_DINV * ((NOALLOC_MAX_N + isle::LU_NB - 1) / isle::LU_NB) + 2 * (size_t)isle::LU_NB * NOALLOC_MAX_N +
    2 * isle::LU_NB + 2;  // lu_work_doubles(NOALLOC_MAX_N)
static constexpr size_t NOALLOC_OFF_W = noalloc_align(2 * sizeof(double) * NOALLOC_MAX_N * NOALLOC_MAX_N);
static constexpr size_t NOALLOC_OFF_PIV = NOALLOC_OFF_W + noalloc_align(NOALLOC_WORK * sizeof(double));
static constexpr size_t NOALLOC_OFF_ITEM = NOALLOC_OFF_PIV + noalloc_align(NOALLOC_MAX_N * sizeof(int));
static constexpr size_t NOALLOC_OFF_LD = NOALLOC_OFF_ITEM + noalloc_align(sizeof(isle::LuItem));
static constexpr size_t NOALLOC_BYTES = NOALLOC_OFF_LD + noalloc_align(2 * sizeof(double));
static constexpr int NOALLOC_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex noalloc_mu;
static unsigned char* noalloc_base[NOALLOC_DEVICES];  // device address of `scratch` per device

// The current device's copy of `scratch` (caller holds noalloc_mu).
static unsigned char* noalloc_scratch() {
    int dev = 0;
    ISLE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_DEVICES) throw std::out_of_range("device index beyond the scratch table");
    if (!noalloc_base[dev]) {
        void* p = nullptr;
        ISLE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        noalloc_base[dev] = static_cast<unsigned char*>(p);
    }
    return noalloc_base[dev];
}

// Waits for the stream when the call unwinds after enqueueing work, before the lock goes.
struct NoallocDrain {
    hipStream_t s;
    int unwinding = std::uncaught_exceptions();
    ~NoallocDrain() {
        if (std::uncaught_exceptions() > unwinding) (void)hipStreamSynchronize(s);
    }
};

int isle_logdet_noalloc(const double* a, int n, double* out_host, void* stream) {
    return guarded([&] {
        if (n < 1) throw std::invalid_argument("empty matrix");
        if (n > NOALLOC_MAX_N)
            throw std::invalid_argument("isle_logdet_noalloc: n = " + std::to_string(n) +
                                        " exceeds 512, the order its static scratch holds");
        if (isle::lu_work_doubles(n) > NOALLOC_WORK) throw std::logic_error("LU work beyond the static scratch");
        hipStream_t s = (hipStream_t)stream;
        const long nn = (long)n * n;
        std::lock_guard<std::mutex> lock(noalloc_mu);
        unsigned char* base = noalloc_scratch();
        NoallocDrain drain{s};
        double* m = reinterpret_cast<double*>(base);
        double* w = reinterpret_cast<double*>(base + NOALLOC_OFF_W);
        int* piv = reinterpret_cast<int*>(base + NOALLOC_OFF_PIV);
        isle::LuItem* item = reinterpret_cast<isle::LuItem*>(base + NOALLOC_OFF_ITEM);
        double* ld = reinterpret_cast<double*>(base + NOALLOC_OFF_LD);
        hipLaunchKernelGGL(isle::interleaved_to_split_kernel, dim3(isle::ew_blocks(nn)), dim3(256), 0, s, a, m,
                           m + nn, nn);
        ISLE_HIP(hipGetLastError());
        isle::LuItem it{m, m + nn, piv, ld, w, n, n};
        ISLE_HIP(hipMemcpyAsync(item, &it, sizeof(it), hipMemcpyHostToDevice, s));
        isle::launch_lu(item, 1, n, 0, s);
        ISLE_HIP(hipMemcpyAsync(out_host, ld, 16, hipMemcpyDeviceToHost, s));
        ISLE_HIP(hipStreamSynchronize(s));
    });
}

int isle_invert(const double* a, int n, double* out, void* stream) {
    return invert_impl(a, n, out, nullptr, stream);
}

int isle_invert_logdet(const double* a, int n, double* out, double* logdet_host, void* stream) {
    if (!logdet_host) return invert_impl(a, 0, out, nullptr, stream);  // reports the bad argument
    return invert_impl(a, n, out, logdet_host, stream);
}

}  // extern "C"
