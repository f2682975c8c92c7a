/*
 * isle_hfa.h -- C ABI of the gfx950 Hubbard fermion-action library (libisle_hfa.so).
 *
 * Plain pointers and sizes only.  "device" pointers are HIP device allocations;
 * complex numbers are complex128, interleaved (re, im).  Fields phi are time-major:
 * element (t, x) at index t*nx + x.  Batched entry points take `batch` fields stored
 * back to back (stride nx*nt complex).  Every function returns ISLE_OK (0) or an
 * error code; isle_last_error() gives the message of the last failure on the thread.
 * Work is ordered on the stream set with isle_hfm_set_stream (default: null stream).
 *
 * Each entry point replaces the reference interface cited beside it
 * (paths relative to the evanberkowitz/isle source tree).
 */
#ifndef ISLE_HFA_H
#define ISLE_HFA_H

#ifdef __cplusplus
extern "C" {
#endif

#define ISLE_OK 0
#define ISLE_ERR_VALUE 1   /* maps to ValueError   (std::invalid_argument) */
#define ISLE_ERR_INDEX 2   /* maps to IndexError   (std::out_of_range)     */
#define ISLE_ERR_RUNTIME 3 /* maps to RuntimeError (std::runtime_error)    */
#define ISLE_ERR_HIP 4     /* HIP runtime failure                          */

#define ISLE_HOPPING_DIA 0 /* HFAHopping::DIA, src/isle/cpp/action/hubbardFermiAction.hpp:16 */
#define ISLE_HOPPING_EXP 1 /* HFAHopping::EXP                                              */
#define ISLE_PARTICLE 0    /* Species::PARTICLE, src/isle/cpp/species.hpp:10-13            */
#define ISLE_HOLE 1        /* Species::HOLE                                                */

typedef struct isle_hfm isle_hfm; /* opaque fermion-matrix plan */

/* Message of the last error on this thread. */
const char* isle_last_error(void);

/* HubbardFermiMatrixExp/Dia(kappaTilde, muTilde, sigmaKappa)
 *   src/isle/cpp/bind/bind_hubbardFermiMatrix.cpp:26, hubbardFermiMatrixExp.cpp:52-68,
 *   hubbardFermiMatrixDia.cpp:49-62.
 * kappa_tilde: host, nx*nx row-major, symmetric. */
int isle_hfm_create(isle_hfm** out, int hopping, int nx, const double* kappa_tilde, double mu_tilde,
                    int sigma_kappa);
int isle_hfm_destroy(isle_hfm* h);
/* stream: a hipStream_t (e.g. torch.cuda.current_stream().cuda_stream) */
int isle_hfm_set_stream(isle_hfm* h, void* stream);
/* Memory schedule of the mixed-precision force (no reference counterpart; the reference keeps
 * every left product, hubbardFermiAction.cpp:30-40): k > 0 keeps the FP64 suffix chains only
 * every k slices inside each group and recomputes the rest segment by segment (k must divide
 * the group size and nx % 32 == 0, else every product is kept); 0 keeps them all.  Applies to
 * force operations first used after the call. */
int isle_hfm_set_recompute(isle_hfm* h, int k);
/* Launch-shape pinning (no reference counterpart): every step of launch kind `kind` (the kinds
 * of isle_profile_kernels) uses tile shape `tile` instead of the autotuner's choice, in the
 * plan's operations built before and after the call; tile -1 returns to the tuned shape.  An
 * operation with a pinned shape that does not fit one of its steps fails when run (return 3).
 * For reproducible timings and for tests that cover every shape (FP32 chain products with the
 * force diagonal fused, kind 30: tiles 0-3; the register-staged shapes with one LDS buffer of
 * the complex products, kinds 3 and 18: tiles 25-27). */
int isle_hfm_set_tile(isle_hfm* h, int kind, int tile);
int isle_hfm_sync(isle_hfm* h);

/* HubbardFermiMatrixExp::expKappa(species, inv) / HubbardFermiMatrixDia::K(species)
 *   hubbardFermiMatrixExp.cpp:76-93, hubbardFermiMatrixDia.cpp:96-103.
 * out_host: nx*nx real row-major (for DIA, inv is ignored and K is returned). */
int isle_hfm_slice_matrix(isle_hfm* h, int species, int inv, double* out_host);

/* HubbardFermiMatrix{Exp,Dia}::F(tp, phi, species, inv)
 *   bind_hubbardFermiMatrix.cpp:29-30, hubbardFermiMatrixExp.cpp:119-141, hubbardFermiMatrixDia.cpp:118-131.
 * out: device, nx*nx complex128 row-major. */
int isle_hfm_F(isle_hfm* h, int tp, const double* phi, int nt, int species, int inv, double* out);

/* logdetM(hfm, phi, species) -> complex, Im in [-pi, pi]
 *   bind_hubbardFermiMatrix.cpp:51-52, hubbardFermiMatrixExp.cpp:615-625, hubbardFermiMatrixDia.cpp:571-597.
 * Plain FP64, the reference algorithm.  out_host: 2 doubles. */
int isle_logdetM(isle_hfm* h, const double* phi, int nt, int species, double* out_host);

/* Batched logdetM for the action's eval (hubbardFermiAction.cpp:230-272).
 * mode 0: reference algorithm in FP64; mode 1: FP64 products of groups of `group` slices
 * formed in parallel first; mode 2: those group products combined by the pivoted
 * elimination of their block-cyclic matrix (stable where the product of all slices is not
 * representable in FP64; needs nx % 32 == 0 when nt/group > 1).
 * nspecies 1 (particle) or 2 (particle, hole); spin != 0 evaluates at aux = -i*phi (spin
 * basis).  out_host: [batch][nspecies][2]. */
int isle_logdetM_batch(isle_hfm* h, const double* phi, int nt, int batch, int mode, int group, int nspecies,
                       int spin, double* out_host);

/* forceDirectSinglePart(hfm, phi, K, species) for each species (without the -i factor)
 *   src/isle/cpp/action/hubbardFermiAction.cpp:16-60.
 * mode 0: plain FP64 reference algorithm; mode 1: FP32 chains over groups of `group`
 * slices on MFMA with an FP64 combination of the group products; mode 2: FP64 chains over
 * the groups with the combination taken from the elimination of mode 2 of logdetM_batch
 * (same size limits); mode 3: mode 2 with the right chains in FP32 on MFMA (two groups or
 * more).
 * out_dev: device [batch][nspecies][nx*nt] complex128. */
int isle_force_parts(isle_hfm* h, const double* phi, int nt, int batch, int mode, int group, int nspecies,
                     int spin, double* out_dev);

/* isle_force_parts for nops batches of fields of one plan at once (no reference counterpart:
 * the reference evaluates one field at a time; for ensembles of chains of several time
 * extents on one lattice).  Batch i: batches[i] fields of nts[i] slices at phis[i] (device),
 * group groups[i], parts to outs[i] (device [batches[i]][nspecies][nx*nts[i]]).  The batches'
 * operations run merged: steps of one kind and launch shape share one launch.  Two batches
 * may not have the same (nt, batch, group). */
int isle_force_parts_multi(isle_hfm* h, int nops, const double* const* phis, const int* nts, const int* batches,
                           int mode, const int* groups, int nspecies, int spin, double* const* outs);

/* solveM(hfm, phi, species, rhss): solves M x = b for every row b of rhss
 *   bind_hubbardFermiMatrix.cpp:53-55, hubbardFermiMatrixExp.cpp:645-694, hubbardFermiMatrixDia.cpp:617-673.
 * rhss, out: device [nrhs][nx*nt] complex128. */
int isle_solveM(isle_hfm* h, const double* phi, int nt, int species, const double* rhss, int nrhs, double* out);

/* M(phi) applied to each of nrhs rows: the product that the reference forms with the sparse
 * matrix of HubbardFermiMatrixExp::M / HubbardFermiMatrixDia::M (hubbardFermiMatrixExp.cpp:
 * 150-166, hubbardFermiMatrixDia.cpp:140-160; bind_hubbardFermiMatrix.cpp:30-31), used for the
 * residuals of solveM.  x, out: device nrhs x (nx*nt) complex128 rows (time-major, as solveM's
 * right-hand sides); phi: device nx*nt complex128. */
int isle_applyM(isle_hfm* h, const double* phi, int nt, int species, const double* x, long nrhs, double* out);

/* Block LU of Q = M_p M_h^T (getQLU, hubbardFermiMatrixExp.cpp:419-479 for nt > 2; the Dia
 * file repeats it; the QLU struct is hubbardFermiMatrixExp.hpp / .cpp:282-360), nt >= 3, as one
 * cached device operation: the off-diagonal blocks, then the recursion with one LU per
 * diagonal block and complex GEMMs for the products; the log dets of the d_i stay on the
 * device until the end; the block products use four real products each (the three-product
 * form's rounding, amplified by the recursion, would move log det Q by ~1e-8 relative at
 * nx=288, nt=128).  phi: device nx*nt complex128.  Outputs (device, interleaved
 * complex128, row-major nx x nx blocks): dinv [nt], u and l [nt-1], v and h [nt-2];
 * logdet_host: 2 nt doubles (re, im of log det d_i).  Synchronises the handle's stream once. */
int isle_qlu(isle_hfm* h, const double* phi, int nt, double* dinv, double* u, double* l, double* v, double* hh,
             double* logdet_host);

/* solveQ (hubbardFermiMatrixExp.cpp:494-537; bind_hubbardFermiMatrix.cpp:47-49) on the blocks of
 * Q's LU as isle_qlu returns them (device interleaved complex128: dinv [nt], u and l [nt-1], v and
 * h [nt-2] blocks of nx*nx): the forward substitution with l and h, then the backward one with
 * dinv, u and v, every block product a complex GEMM on the device.  rhs, x: device nrhs x
 * (nx*nt) complex128 rows (time-major).  nt >= 1. */
int isle_solveq(isle_hfm* h, int nt, int nrhs, const double* dinv, const double* u, const double* l,
                const double* v, const double* hh, const double* rhs, double* x);

/* solveM with the factorization kept between calls (no counterpart in the reference: for
 * callers that solve many right-hand sides against one field, such as iterative refinement).
 * reuse = 0: as isle_solveM (the handle keeps the factorization: partial products and their
 * LUs, 2 nt nx^2 complex128).  reuse = 1: the caller promises that phi is the field of the
 * previous solveM on this handle, unchanged; with the same nt and species the factorization
 * is reused (otherwise recomputed). */
int isle_solveM_ex(isle_hfm* h, const double* phi, int nt, int species, const double* rhss, int nrhs, double* out,
                   int reuse);

/* Combination of the species terms into the fermion force, n complex elements
 *   (hubbardFermiAction.cpp:243-272): mode 0: -i(fp - fh); 1: -i(fp - conj fp); 2: fh - fp. */
int isle_combine_force(const double* fp, const double* fh, double* out, long n, int mode, int accumulate,
                       void* stream);
/* The same for a batch of fields in one launch (no counterpart in the reference, which
 * evaluates one field at a time): field b's parts start at fp/fh + b*in_stride and its output
 * at out + b*out_stride (complex elements).  batch <= 65535. */
int isle_combine_force_batch(const double* fp, const double* fh, double* out, long n, int batch, long in_stride,
                             long out_stride, int mode, int accumulate, void* stream);

/* HubbardGaugeAction::force = -phi/utilde (hubbardGaugeAction.cpp:11-13). */
int isle_gauge_force(const double* phi, double* out, long n, double utilde, int accumulate, void* stream);

/* sum(phi) and the bilinear (phi, phi) of `batch` vectors of n complex elements with
 * the given stride (complex elements); out_host [batch][4] = (Re s1, Im s1, Re s2, Im s2).
 *   hubbardGaugeAction.cpp:7-9 (eval), hubbardFermiMatrixExp.cpp:610 (sum of phi). */
int isle_field_sums(const double* phi, long n, long stride, int batch, double* out_host, void* stream);

/* Leapfrog kick  pi += eps*Re(force) [ - eps*Re(phi)/utilde if phi != NULL and utilde > 0 ]
 * and drift phi += eps*pi on n complex elements (src/isle/cpp/integrator.cpp:20-32). */
int isle_leapfrog_kick(double* pi, const double* force, const double* phi, long n, double eps, double utilde,
                       void* stream);
int isle_leapfrog_drift(double* phi, const double* pi, long n, double eps, void* stream);

/* out = c * in for n complex elements (the spin basis' aux = -i*phi, hubbardFermiAction.cpp:261). */
int isle_cscale(const double* in, double* out, long n, double cr, double ci, void* stream);

/* Instrumentation (no counterpart in the reference, which only has timeit benchmarks,
 * benchmarks/logdet.py:46,85): replays a cached operation `reps` times with a HIP event
 * after every launch.  op: 0 plain force, 1 mixed force, 2 plain logdet, 3 grouped logdet,
 * 4 stable force, 5 stable logdet.
 * For launch kind k < nkinds: out_host[4k..4k+3] = (launches per operation, ms per
 * operation, algorithmic real flops per operation, 0).  Kinds: 0 GEMM64 real x complex
 * (pre-scaled), 1 GEMM64 real x complex (row-scaled), 2 GEMM64 complex x real, 3 GEMM64
 * complex x complex, 4 GEMM32 real x complex (row-scaled), 5 GEMM32 real x complex
 * (pre-scaled), 6-7 chain starts, 8 LU, 9 LU solve, 10-11 diagonal extraction, 12-14
 * copies, 15 phases, 16 transposed copy, 17 diagonal extraction FP64 x FP32, 18 GEMM64
 * complex x complex subtracted, 19 LU of a block-cyclic elimination step, 20 block
 * assembly, 21 row permutation, 22 permuted gather, 23 transposed block extraction, 24 blocked
 * triangular solve (GEMM launches), 25 GEMM32 real x complex with the force diagonal's row sums
 * fused, 26 sum of those row sums, 27 rows of right-hand sides in pivot order (solveM), 28 GEMM32
 * real x complex with the row sums fused and the stored rows scaled by the next slice's phase,
 * 29 transposed FP64 -> FP32 copy with complex row scaling, 30 GEMM32 real x complex on
 * k-contiguous slabs with the row sums fused (product stored transposed), 31 FP64 -> FP32 copy
 * with complex column scaling, 32 LU of a block-cyclic elimination step whose columns beyond
 * the factorised ones are eliminated by a blocked triangular solve and one GEMM, 33 GEMM64
 * complex x complex with four real products (block LU of Q), 34 the same, subtracted. */
int isle_profile_kernels(isle_hfm* h, const double* phi, int op, int nt, int batch, int group, int nspecies,
                         int spin, int reps, double* out_host, int nkinds);

/* logdet(matrix): LU with partial pivoting, sum log U_ii + i*pi*(swap parity), Im in [-pi, pi]
 *   src/isle/cpp/math.hpp:265-334 (bound as isle.logdet).  a: device n*n complex128. */
int isle_logdet(const double* a, int n, double* out_host, void* stream);

/* log det of one matrix as isle_logdet, without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the LU's scratch is a static device array of about
 * 5 MiB per GPU, held until the process ends; n must be <= 512 (larger n is rejected with
 * ISLE_ERR_VALUE before anything is enqueued).  Calls in one process are serialised and return
 * after `stream` is synchronised; they share that GPU's array, so overlapping calls from
 * outside this lock would race.  a: device n*n complex128; out_host: 2 doubles. */
int isle_logdet_noalloc(const double* a, int n, double* out_host, void* stream);

/* invert(matrix): getrf + getri equivalent (src/isle/cpp/math.hpp:222-226).
 * a, out: device n*n complex128 (a is not modified). */
int isle_invert(const double* a, int n, double* out, void* stream);

/* A^{-1} and log det A (Im in [-pi, pi]) from one LU; no single reference counterpart.  It
 * serves the QLU recursion of Q = M_p M_h^T, which inverts every diagonal block d_i
 * (src/isle/cpp/hubbardFermiMatrixExp.cpp:420-500) and later sums -ilogdet(d_i^{-1})
 * (:526-541); here log det d_i comes from the LU that produced d_i^{-1}.
 * a, out: device n*n complex128; logdet_host: 2 doubles (required). */
int isle_invert_logdet(const double* a, int n, double* out, double* logdet_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ISLE_HFA_H */
