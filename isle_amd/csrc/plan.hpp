// Fermion-matrix plan: device copies of the slice matrices, cached per-operation
// descriptor sets and workspaces, and the step lists that implement logdetM,
// solveM and the DIRECT_SINGLE force in plain FP64 and in mixed precision.
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "cyclic.hpp"
#include "elementwise.hpp"
#include "gemm.hpp"
#include "lu.hpp"

namespace isle {

enum Species : int { PARTICLE = 0, HOLE = 1 };
enum Hopping : int { DIA = 0, EXP = 1 };

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) : bytes(b) {
        if (b) ISLE_HIP(hipMalloc(&p, b));
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Kinds of recorded launches.
enum class K : int {
    GEMM64_RC_PRE,      // real A x complex B, B pre-scaled along k        (f64)
    GEMM64_RC_ROW,      // real A x complex B, rows post-scaled            (f64)
    GEMM64_CR_PRE,      // complex A x real B, A pre-scaled along k        (f64)
    GEMM64_CC,          // complex x complex                               (f64)
    GEMM32_RC_ROW,      // (f32)
    GEMM32_RC_PRE,      // (f32)
    SCALE64, SCALE32,   // chain start diag(d) A or A diag(d)
    LU,                 // factorise (flag: + I), logdet
    SOLVE,              // X = A^{-1} B with stored LU
    DIAG64, DIAG32,     // diag(L R)
    COPY64, CVT32TO64, CVT64TO32,
    PHASE,              // phase vectors from the runtime phi
    TCVT64TO32,         // transposed copy with conversion
    DIAG64X32,          // diag(L R), FP64 left factor, FP32 right factor
    GEMM64_CC_SUB,      // C -= A B, complex x complex                     (f64)
    LU_SHAPE,           // LU of the first kf columns of m x ncols matrices (max_m = m, max_n = ncols,
                        // flag = kf, aux = identity-row bound nt)
    BLOCK,              // block assembly (zero / identity / copy / negated copy)
    SIGMA,              // row permutation of stacked right-hand sides
    GATHER,             // permuted split of stacked right-hand sides
    TBLOCK,             // signed transposed block extraction
    TRSM,               // one product of a blocked triangular solve (flag 1: C -= A B)
    GEMM32_RC_PRE_DIAG, // GEMM32_RC_PRE with the force diagonal's partial row sums fused (f32)
    DREDUCE,            // sum of those partial row sums
    RGATHER,            // rows of right-hand sides in the pivot order of an LU (max_m = rows, aux = columns)
    GEMM32_RC_DIAG,     // f32 real x complex, no pre-scaling, force diagonal's row sums fused and the
                        // stored rows scaled by the next slice's phase
    TSCALE64TO32,       // transposed copy with conversion and complex row scaling
    GEMM32_RCK_DIAG,    // f32 real x complex on k-contiguous slabs (operand and product stored
                        // transposed), force diagonal's row sums fused, rows scaled by the next phase
    COLSCALE64TO32,     // copy with conversion and complex column scaling
    LU_SHAPE_DEFER,     // LU_SHAPE whose columns beyond kf take the row interchanges only
    GEMM64_CC4,         // complex x complex with four real products       (f64)
    GEMM64_CC4_SUB,     // C -= A B, the same                              (f64)
    KIND_END
};

struct Step {
    K kind;
    const void* items;   // device descriptor array
    int count;           // number of items
    int max_m, max_n;    // launch geometry (GEMM: max m, n; LU/SOLVE: max n, max rhs)
    int flag;            // LU: add identity
    double flops;        // algorithmic real flops of the launch (GEMMs: 2mnk per real product)
    int tile = 0;        // GEMM tile shape chosen by the autotuner (TILE_64 / TILE_32)
    int aux = 0;         // LU_SHAPE: nt; BLOCK / GATHER: largest element count per item
    size_t ibytes = 0;   // bytes per descriptor
    size_t off = 0;      // offset of the descriptors in the operation's host blob
};

// A cached operation: descriptor blob + workspaces + step list.
struct OpCache {
    std::vector<std::unique_ptr<DevBuf>> bufs;
    std::vector<Step> steps;
    std::vector<char> blob;    // host copy of the descriptors (for merged operations)
    std::vector<double*> extra;  // operation-specific result buffers (the QLU blocks)
    long phi_stride = 0;       // doubles between consecutive fields of the batch
    double* logdet = nullptr;  // device [batch][nspecies][2]
    double* parts = nullptr;   // device [batch][nspecies][nx*nt] interleaved
    int nspecies = 0;
    int batch = 0;
    int nt = 0;
};

class Plan {
public:
    Plan(int hopping, int nx, const double* kappa_tilde, double mu_tilde, int sigma_kappa);

    int nx() const { return nx_; }
    int hopping() const { return hop_; }
    void set_stream(hipStream_t s) { stream_ = s; }
    // Recompute interval k of the mixed force's suffix chains (0: keep them all); applies to
    // force operations built afterwards (see get_force_mixed).
    void set_recompute(int k) {
        if (k < 0) throw std::invalid_argument("recompute interval must be >= 0");
        recompute_ = k;
    }
    hipStream_t stream() const { return stream_; }
    // Pins the launch shape of every step of kind `kind` (a K value) to `tile` instead of the
    // autotuner's choice, in operations built before and after; tile -1 returns to the tuned
    // shape.  A pinned shape that does not fit a step makes its operation throw when run.
    // HIP graphs captured earlier keep the shapes they recorded.
    void set_tile(int kind, int tile);

    // log det M for a batch of fields (consecutive, nx*nt complex each) and nspecies
    // species (1: particle only, 2: particle and hole).  mode 0 follows the reference
    // algorithm step by step; mode 1 forms FP64 products of groups of `group` slices
    // in parallel first; mode 2 factorises the block-cyclic matrix of those group products
    // (structured elimination with partial pivoting, no product of all groups is formed).
    // spin != 0 evaluates at aux = -i*phi.  out_host: [batch][nspecies][2].
    void logdetM(const double* phi, int nt, int batch, int mode, int group, int nspecies, int spin,
                 double* out_host);
    // DIRECT_SINGLE force terms (without the -i prefactor) for each species, written
    // interleaved to out_dev [batch][nspecies][nx*nt].  mode 0: plain FP64 reference
    // algorithm; mode 1: FP32 local chains over groups of `group` slices on MFMA plus
    // an FP64 combination of the group products; mode 2: FP64 local chains with the
    // combination taken from the structured factorisation of the block-cyclic matrix.
    void force_parts(const double* phi, int nt, int batch, int mode, int group, int nspecies, int spin,
                     double* out_dev);
    // force_parts for several batches of fields at once (nops of them, batch i with nts[i]
    // slices, batches[i] fields at phis[i], group groups[i], parts to outs[i]): their
    // operations' steps are aligned by kind and every aligned step is one launch over the
    // descriptors of all of them (the phase steps, which bind the fields, stay separate).
    void force_parts_multi(int nops, const double* const* phis, const int* nts, const int* batches, int mode,
                           const int* groups, int nspecies, int spin, double* const* outs);
    // Block LU of Q = M_p M_h^T (getQLU, nt >= 3): the stacks d_i^{-1} (nt), u_i, l_i (nt-1),
    // v_i, h_i (nt-2) written interleaved to dinv, u, l, v, h (device [k][nx][nx] complex128)
    // and log det d_i to logdet_host (nt x 2); one synchronisation at the end.
    void qlu(const double* phi, int nt, double* dinv, double* u, double* l, double* v, double* h,
             double* logdet_host);
    // x with Q x = rhs from the blocks of Q's LU (interleaved complex128 device stacks as
    // getQLU returns them); rhs, x: nrhs x (nt * nx) rows.  nt >= 1.
    void solve_q(int nt, int nrhs, const double* dinv, const double* u, const double* l, const double* v,
                 const double* h, const double* rhs, double* x);
    // reuse: phi, nt and species equal those of the previous call (its factorization is kept)
    void solveM(const double* phi, int nt, int species, const double* rhss, int nrhs, double* out,
                bool reuse = false);
    void F(int tp, const double* phi, int nt, int species, int inv, double* out);
    // M(phi) of `species` applied to nrhs rows x (nrhs x nx*nt interleaved complex) -> out
    void applyM(const double* phi, int nt, int species, const double* x, long nrhs, double* out);
    void slice_matrix(int species, int inv, double* out_host);
    void sync() { ISLE_HIP(hipStreamSynchronize(stream_)); }
    // Time every launch of a cached force (op 0/1/4) or logdet (op 2/3/5) operation with HIP
    // events, `reps` times.  Per launch kind k (enum K): out[4k..4k+3] =
    // (launches per operation, ms per operation, algorithmic flops per operation, 0).
    void profile(const double* phi, int op, int nt, int batch, int group, int nspecies, int spin, int reps,
                 double* out, int nkinds);

    double logdet_A(int species) const { return logdet_a_[species]; }

private:
    int hop_, nx_;
    double mu_;
    int sk_;
    std::vector<double> kappa_;
    hipStream_t stream_ = nullptr;
    int recompute_ = 0;
    std::map<int, int> forced_;  // kind -> pinned tile
    // inverse-form slice matrices A_s (EXP: exp(-k+mu), exp(-s k-mu); DIA: K_p, K_h)
    std::unique_ptr<DevBuf> a64_[2], a32_[2];
    // EXP non-inverse matrices exp(k-mu), exp(s k+mu)
    std::unique_ptr<DevBuf> e64_[2];
    double logdet_a_[2] = {0, 0};
    std::map<std::tuple<int, int, int, int, int, int>, std::unique_ptr<OpCache>> cache_;

    std::unique_ptr<DevBuf> ident64_;
    // solveM's right-hand-side buffers (GBs for many sources), kept between calls and grown
    // when a call needs more: a hipMalloc of that size costs far more than the solve
    std::unique_ptr<DevBuf> solve_ws_;
    // solveM's factorization of the last field: phases, the matrices that multiply the
    // right-hand sides, the partial products' LUs (in place), pivots and LU work buffers
    struct SolveFactor {
        int nt = -1, species = -1;
        std::unique_ptr<DevBuf> ph, mz, part, piv, luw;
    };
    SolveFactor sf_;
    // applyM's split-plane operands and GEMM descriptors, kept and grown like solve_ws_
    std::unique_ptr<DevBuf> apply_ws_, apply_items_;
    size_t apply_ws_doubles_ = 0;
    int apply_items_n_ = 0;
    size_t solve_ws_doubles_ = 0;
    double* solve_workspace(size_t doubles);

    void build_exp(const std::vector<double>& x, DevBuf& out);
    OpCache& get_force_plain(int nt, int batch, int nspecies, int spin);
    OpCache& get_force_mixed(int nt, int batch, int group, int nspecies, int spin);
    OpCache& get_logdet_plain(int nt, int batch, int nspecies, int spin);
    OpCache& get_logdet_grouped(int nt, int batch, int group, int nspecies, int spin);
    OpCache& get_force_stable(int nt, int batch, int group, int nspecies, int spin, bool right32 = false);
    OpCache& get_logdet_stable(int nt, int batch, int group, int nspecies, int spin);
    OpCache& get_qlu(int nt);
    OpCache& get_solveq(int nt, int nrhs);
    void run(OpCache& op, const double* phi);
    bool launch_step(OpCache& op, const Step& st, const double* phi);
    // Autotuner: time every tile shape of each GEMM step once per (kind, m, n, items) and
    // keep the fastest for that step.
    void tune(OpCache& op);
    std::map<std::tuple<int, int, int, int>, int> tuned_;
    // a merged operation (force_parts_multi): its own steps (descriptors in merged.bufs) and,
    // per step, the operation whose field it binds (-1: none) and that operation
    struct MultiOp {
        OpCache merged;
        std::vector<int> owner;
        std::vector<OpCache*> ops;
    };
    std::map<std::vector<long>, std::unique_ptr<MultiOp>> multi_;
    OpCache& get_force(int mode, int nt, int batch, int group, int nspecies, int spin);
    std::unique_ptr<MultiOp> merge_ops(const std::vector<OpCache*>& ops);
};

}  // namespace isle
