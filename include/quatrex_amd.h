/*
 * quatrex_amd — C ABI of the MI355X (gfx950) batched RGF + open-boundary path.
 *
 * Plain C: opaque handle, plain device pointers, sizes and strides.  Every
 * array argument is DEVICE memory unless marked "host".  Complex numbers are
 * interleaved (re, im) doubles (complex128) in all arrays that cross the
 * boundary; the `precision` argument selects the arithmetic used inside.
 * Calls are asynchronous on the given HIP stream (hipStream_t passed as
 * void*, NULL = default stream) and never throw.  Return codes: 0 ok,
 * >0 warning, <0 error; qx_last_error() gives the message.
 *
 * Block-tridiagonal (BTD) stacks use the storage of the reference's DSDBCOO
 * container (src/qttools/datastructures/dsdbcoo.py:23-1008): one array of
 * shape (n_stack, stack_stride) whose fully populated blocks are contiguous
 * row-major slices.  qx_btd_layout gives the element offset of each block
 * inside one stack entry; an offset of -1 marks a block that is not stored
 * (read as zero, never written).
 */
#ifndef QUATREX_AMD_H
#define QUATREX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QX_ABI_VERSION 3

/* status codes */
#define QX_OK 0
#define QX_NOT_CONVERGED 1   /* warning: iteration limit reached            */
#define QX_FP64_FALLBACK 2   /* warning: mixed mode redid block inverses in  */
                             /* FP64 (Newton-Schulz residual check failed)   */
#define QX_EINVAL (-1)       /* invalid argument                            */
#define QX_ENOMEM (-2)       /* device allocation failed                    */
#define QX_EHIP (-3)         /* HIP runtime error                           */
#define QX_ESINGULAR (-4)    /* exactly singular block met by an inversion  */

/* arithmetic selector */
#define QX_PREC_C128 0       /* complex128 on FP64 MFMA                      */
#define QX_PREC_C64 1        /* complex64 on FP32 MFMA, FP64 in/out          */
#define QX_PREC_MIXED 2      /* RGF only: complex64 products, FP64 block     */
                             /* inverses (env QX_MIXED_INVERSE=c64: complex64 */
                             /* inverses with one FP64 Newton-Schulz step)    */
#define QX_PREC_C128_3M 3    /* complex128 on FP64 MFMA with the 3M product   */
                             /* split (3/4 of the matrix-core work, rounding  */
                             /* error a small factor above the 4M split)      */

/* complex128 storage and inverses; block products on int8 matrix cores by
 * modular arithmetic (Ozaki scheme II) with n pairwise coprime moduli,
 * 8 <= n <= 20: every line of a product's operands is scaled to integers of
 * T bits (T = 42 for 12 moduli at k <= 256), so the relative error per
 * product is ~2^-T of the line maxima.  Accepted by qx_zgemm_batched and the
 * RGF calls. */
#define QX_PREC_OZAKI2(n) (16 + (n))
/* complex128 storage and inverses; block products on bf16 matrix cores from
 * s 7-bit digit slices of every operand line (2 <= s <= 7, the arithmetic of
 * qx_zgemm_ozaki): relative error per product ~2^(-7 s).  Accepted by
 * qx_zgemm_batched and the RGF calls. */
#define QX_PREC_OZAKI_BF16(s) (40 + (s))

/* GEMM operand ops */
#define QX_OP_N 0
#define QX_OP_C 2            /* conjugate transpose                          */

/* selected_solve flags */
#define QX_SOLVE_RETARDED 1u         /* write diagonal G^R blocks              */
#define QX_SOLVE_CURRENT 2u          /* Meir-Wingreen layer current            */
#define QX_SOLVE_RETARDED_OFFDIAG 4u /* also write off-diagonal G^R blocks     */
#define QX_SOLVE_SKIP_LOWER_L 8u     /* lesser output is symmetric-stored      */
#define QX_SOLVE_SKIP_LOWER_G 16u    /* greater output is symmetric-stored     */

typedef struct qx_handle qx_handle;

typedef struct {
  int32_t n_blocks;
  int32_t reserved;
  int64_t n_stack;             /* matrices addressed by this call            */
  int64_t stack_stride;        /* elements between consecutive matrices      */
  const int32_t* block_sizes;  /* host [n_blocks]                             */
  const int64_t* diag;         /* host [n_blocks]    offsets of (i, i)        */
  const int64_t* upper;        /* host [n_blocks-1]  offsets of (i, i+1)      */
  const int64_t* lower;        /* host [n_blocks-1]  offsets of (i+1, i)      */
} qx_btd_layout;

/* Library version of this header (QX_ABI_VERSION). */
int qx_abi_version(void);

/* Handle bound to HIP device `device`; owns a growable workspace. */
int qx_create(qx_handle** handle, int device);
int qx_destroy(qx_handle* handle);
const char* qx_last_error(const qx_handle* handle);

/* Workspace the RGF calls need for `layout` (bytes), for sizing budgets. */
int qx_workspace_query(qx_handle* handle, const qx_btd_layout* layout,
                       int precision, size_t* bytes);

/* Kernel timing (no counterpart in the reference; replaces its NVTX/timer
 * profiler, src/qttools/profiling/profiler.py:187-429, for the kernels of
 * this path).  Between qx_timer_start and qx_timer_stop every GEMM and
 * inverse launch made through `handle` is bracketed by HIP events on its
 * stream.  qx_timer_stop waits for the device and writes host stats[8] =
 * {gemm launches, gemm ms summed over launches, gemm flops, gemm ms of wall
 * time covered by at least one gemm launch, then the same four for the
 * inverse} (flops = algorithmic 8 m n k per complex product and 8 n^3 per
 * complex inverse).  `stream` of qx_timer_start anchors the time origin. */
int qx_timer_start(qx_handle* handle, void* stream);
int qx_timer_stop(qx_handle* handle, void* stream, double* stats);

/* Wait for `stream`, then report (and clear) device-side errors recorded
 * since the last check: QX_ESINGULAR if an inversion met a zero pivot, else
 * QX_FP64_FALLBACK if a QX_PREC_MIXED call with complex64 block inverses
 * (QX_MIXED_INVERSE=c64) replaced one whose Newton-Schulz residual
 * ||I - M X0||_F was not below the tolerance (env QX_MIXED_TOL, default 1e-2)
 * by the FP64 pivoted inverse. */
int qx_sync_check(qx_handle* handle, void* stream);

/* RGF selected inversion.  Replaces RGF.selected_inv
 * (src/qttools/greens_function_solver/rgf.py:25-113) for one batch of the
 * stack.  obc_r: host array [n_blocks] of device pointers to
 * (n_stack, b_i, b_i) contiguous boundary self-energies or NULL entries
 * (OBCBlocks.retarded, solver.py:12-28); may itself be NULL. */
int qx_rgf_selected_inv(qx_handle* handle, void* stream,
                        const qx_btd_layout* la, const void* a,
                        const void* const* obc_r,
                        const qx_btd_layout* lx, void* x, int precision);

/* RGF selected solve X^{<>} = A^{-1} S^{<>} A^{-H}.  Replaces
 * RGF.selected_solve (rgf.py:115-414) for one batch of the stack: G^<, G^>
 * blocks (diagonal anti-Hermitian-projected, lower = -(upper)^H unless the
 * SKIP_LOWER flag is set), G^R diagonal (QX_SOLVE_RETARDED) and optionally
 * off-diagonal blocks, and the layer current (n_stack, n_blocks-1) complex
 * when QX_SOLVE_CURRENT is set.  obc_*: OBCBlocks.retarded/lesser/greater.
 * corner (nullable, no counterpart in the reference): receives the retarded
 * corner block G^R_{0,N-1}, (n_stack, b_0, b_{N-1}) contiguous, from the
 * forward recursion g_{0,j} = -g_{0,j-1} A_{j-1,j} x^R_jj; it gives the
 * transmission Tr[Gamma_L G^R_{0,N-1} Gamma_R (G^R_{0,N-1})^H].
 * density (nullable): (n_stack, sum of block sizes) doubles, receives
 * Im diag(G^<) as each diagonal block is finished (the quantity of
 * quatrex/core/observables.py:57-86 without overlap), without a second pass
 * over the output.
 * ldos (nullable): same layout, receives the electron local density of
 * states -Im diag(G^R) / (2 pi) (quatrex/core/scba.py:572-576 without
 * overlap) as each diagonal G^R block is finished, whether or not
 * QX_SOLVE_RETARDED is set. */
int qx_rgf_selected_solve(qx_handle* handle, void* stream,
                          const qx_btd_layout* la, const void* a,
                          const qx_btd_layout* lsl, const void* sigma_lesser,
                          const qx_btd_layout* lsg, const void* sigma_greater,
                          const void* const* obc_r, const void* const* obc_l,
                          const void* const* obc_g,
                          const qx_btd_layout* lxl, void* x_lesser,
                          const qx_btd_layout* lxg, void* x_greater,
                          const qx_btd_layout* lxr, void* x_retarded,
                          void* current, void* corner, double* density,
                          double* ldos, uint32_t flags, int precision);

/* Batched complex GEMM C = alpha op(A) op(B) + beta C, op in {N, C}.
 * Replaces the batched `@` of the reference (cuBLAS zgemm strided-batched).
 * alpha/beta: host [2] (re, im). */
int qx_zgemm_batched(qx_handle* handle, void* stream, int64_t batch, int m,
                     int n, int k, int op_a, int op_b, const double* alpha,
                     const void* a, int lda, int64_t stride_a, const void* b,
                     int ldb, int64_t stride_b, const double* beta, void* c,
                     int ldc, int64_t stride_c, int precision);

/* Batched inverse with partial pivoting, X = inv(A - S) (S nullable).
 * Replaces linalg.inv (src/qttools/kernels/linalg/inv.py:24-40).  n <= 512. */
int qx_inv_batched(qx_handle* handle, void* stream, int64_t batch, int n,
                   const void* a, int lda, int64_t stride_a, const void* s,
                   int lds, int64_t stride_s, void* x, int ldx,
                   int64_t stride_x, int precision);

/* Batched inverse that allocates no device memory (no counterpart in the
 * reference: for callers that cannot allocate, such as code under HIP graph
 * capture or tight serving loops).  Same result as qx_inv_batched, computed
 * by the same kernels, with its scratch (ping-pong buffer, pivots, status
 * flags) in a fixed array of QX_NOALLOC_SCRATCH_BYTES per GPU instead of the
 * handle's workspace: batch * (16 n^2 + 12 n + 4) bytes plus alignment must
 * fit (e.g. 63 matrices of 256, 15 of 512), else QX_EINVAL and nothing runs.
 * info (nullable, device, batch ints) receives a nonzero flag per exactly
 * singular matrix.  All calls on one GPU share its array: calls that overlap
 * in time on one GPU race.  No kernel timing. */
#define QX_NOALLOC_SCRATCH_BYTES (64u << 20)
int qx_inv_batched_noalloc(qx_handle* handle, void* stream, int64_t batch, int n,
                           const void* a, int lda, int64_t stride_a, const void* s,
                           int lds, int64_t stride_s, void* x, int ldx,
                           int64_t stride_x, int* info, int precision);

/* Sancho-Rubio surface Green's function on (n, b, b) stacks.  Replaces
 * SanchoRubio.__call__ (src/qttools/boundary_conditions/obc/sancho_rubio.py:33-86):
 * batch-global stop when max_e ||(|alpha|+|beta|)||_F / 2 < tol.  Writes
 * the iteration count to *iterations (host).  Returns QX_NOT_CONVERGED when
 * max_iterations is reached (x_ii is still written). */
int qx_sancho_rubio(qx_handle* handle, void* stream, int64_t n, int b,
                    const void* a_ii, const void* a_ij, const void* a_ji,
                    void* x_ii, int max_iterations, double tol,
                    int* iterations);

/* Beyn contour moments.  Replaces Beyn._contour_integrate
 * (src/qttools/nevp/beyn.py:177-235) with operator_inverse
 * (src/qttools/kernels/operator.py:113-179): for the n_quad nodes
 * z = r exp(2 pi i (k + 1.5) / n_quad), k = 0..n_quad-1 (beyn.py:198-202),
 * on the outer (r_o) and inner (r_i) circles,
 *   q0 = sum_k w (z_o T(z_o)^{-1} - z_i T(z_i)^{-1}),
 *   q1 = sum_k w (z_o^2 T(z_o)^{-1} - z_i^2 T(z_i)^{-1}),  w = 1/n_quad,
 * T(z) = a_ji / z + a_ii + a_ij z.  All (n, b, b) contiguous. */
int qx_beyn_moments(qx_handle* handle, void* stream, int64_t n, int b,
                    int n_quad, double r_o, double r_i, const void* a_ji,
                    const void* a_ii, const void* a_ij, void* q0, void* q1,
                    int precision);

/* Batched thin SVD a = u diag(s) v^H of (n, rows, cols) matrices, cols <=
 * rows, cols <= 512, by one-sided Jacobi rotations: u (n, rows, cols) and
 * v (n, cols, cols) contiguous, s (n, cols) doubles in descending order.
 * Replaces the device branch of linalg.svd in Beyn's projection
 * (src/qttools/nevp/beyn.py:96-145 with project_compute_location other than
 * "numpy").  Stops after a sweep in which every column pair is orthogonal to
 * |g_p^H g_q| <= tol |g_p| |g_q|, or after max_sweeps; sweeps (nullable,
 * device, n ints) receives the sweeps used per matrix. */
int qx_svd_batched(qx_handle* handle, void* stream, int64_t n, int rows, int cols,
                   const void* a, int lda, int64_t stride_a, void* u, double* s, void* v,
                   int max_sweeps, double tol, int* sweeps);

/* Reflected-mode classification of the Spectral OBC, fused into one pass
 * over the eigenpairs.  Replaces the elementwise part of
 * src/qttools/boundary_conditions/obc/spectral.py:198-320
 * (_find_reflected_modes, no injection vectors) and :160-196
 * (_compute_dE_dk): for every eigenpair (ws[e, j], column j of vs[e]), given
 * the products av_ji = a_ji vs, av_ii = a_ii vs, av_ij = a_ij vs,
 *   residual = || av_ji w^-1 + av_ii + av_ij w || / || v ||  (/ |w| when
 *              residual_normalization),
 *   de_dk    = i (v^H av_ji) w^-1 - i (v^H av_ij) w,  k = -i log w,
 * and mask = 1 for a propagating mode (|Im k| < min_decay, Re dE/dk < 0,
 * |Re dE/dk| / (|Im dE/dk| + eps) > min_propagation) or a decaying one
 * (Im k < -min_decay, or < -eta_decay when the ratio is not above
 * min_propagation; Im k > -max_decay) whose residual is below
 * residual_tolerance; NaN and infinities in dE/dk and k count as zero.
 * ws (n, m), vs and the products (n, b, m) contiguous complex128; mask
 * (n, m) bytes; residual (n, m) doubles and de_dk (n, m) complex128 are
 * nullable. */
int qx_spectral_modes(qx_handle* handle, void* stream, int64_t n, int b, int m,
                      const void* ws, const void* vs, const void* av_ji,
                      const void* av_ii, const void* av_ij, double min_decay,
                      double max_decay, double min_propagation,
                      double residual_tolerance, double eta_decay,
                      int residual_normalization, unsigned char* mask,
                      double* residual, void* de_dk);

/* Contact (boundary) self-energies of one lead.  Replaces the assembly of
 * src/quatrex/electron/solver.py:324-337 (left) and :370-386 (right):
 *   sigma_r = m10 g m01,  Gamma = i (sigma_r - sigma_r^H),
 *   sigma_l = i f Gamma,  sigma_g = i (f - 1) Gamma
 * (scale_stack, src/qttools/utils/stack_utils.py:6-30), with one occupation
 * f per matrix (occupancies: n doubles in device memory).  All matrices
 * (n, b, b) contiguous. */
int qx_contact_self_energy(qx_handle* handle, void* stream, int64_t n, int b,
                           const void* m10, const void* g, const void* m01,
                           const double* occupancies, void* sigma_r,
                           void* sigma_l, void* sigma_g);

/* Ozaki-style split GEMM (no counterpart in the reference; an alternative
 * arithmetic for the products of rgf.py): C = alpha op(A) op(B) + beta C for
 * complex128 data, computed on bf16 matrix cores from `slices` (2..7)
 * 7-bit digit slices of every row of op(A) and column of op(B), with exact
 * fp32 accumulation of each digit product; the relative error is a small
 * multiple of 2^(-7 slices) (6e-6 at 3 slices, 5e-14 at 7).
 * Same argument conventions as qx_zgemm_batched; k <= 4096.  Uses the
 * handle's workspace for the digits. */
int qx_zgemm_ozaki(qx_handle* handle, void* stream, int64_t batch, int m, int n,
                   int k, int op_a, int op_b, const double* alpha,
                   const void* a, int lda, int64_t stride_a, const void* b,
                   int ldb, int64_t stride_b, const double* beta, void* c,
                   int ldc, int64_t stride_c, int slices);

/* GEMM kernel choices for qx_set_tuning's gemm_tile.  The register-direct
 * kernels (complex128 arithmetic only; no LDS, no workgroup barriers) are
 * named by their workgroup tile; 64 and 128 select the LDS-staged kernel with
 * 64 x 64 or 128 x 128 tiles (all precisions). */
#define QX_GEMM_AUTO 0
#define QX_GEMM_REG_64x64 1
#define QX_GEMM_REG_64x128 2
#define QX_GEMM_REG_128x64 3
#define QX_GEMM_LDS_64 64
#define QX_GEMM_LDS_128 128

/* Execution tuning of a handle (no counterpart in the reference): streams =
 * concurrent sub-batches of one RGF call (1..8, default 4, env
 * QX_RGF_STREAMS); gemm_tile = one of QX_GEMM_* (env QX_ZGEMM_TILE,
 * QX_ZGEMM_KERNEL and QX_ZGEMM_REG override it).  Used by the autotuner and
 * by callers that run several handles side by side. */
int qx_set_tuning(qx_handle* handle, int streams, int gemm_tile);

/* System matrix on the DSDBCOO nnz pattern.  Replaces
 * ElectronSolver._assemble_system_matrix (src/quatrex/electron/solver.py:
 * 390-407 with _btd_subtract, :26-50) for stacks whose operands share the
 * pattern of A:  a[e, k] = z_e s[k] + t[k] - sigma[e, k], where z_e = E_e +
 * i eta (n_stack complex), s the overlap (or identity) values and t the
 * static part -H - diag(V) on the pattern (nnz complex each, nullable = 0),
 * sigma the retarded scattering self-energy (nullable), stack strides
 * sigma_stride / a_stride. */
int qx_assemble_system_matrix(qx_handle* handle, void* stream, int64_t n_stack,
                              int64_t nnz, const void* z, const void* s,
                              const void* t, const void* sigma,
                              int64_t sigma_stride, void* a, int64_t a_stride);

#ifdef __cplusplus
}
#endif

#endif /* QUATREX_AMD_H */
