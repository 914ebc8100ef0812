// Internal declarations shared by the quatrex_amd HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/quatrex_amd.h"

namespace qx {

// Complex128 element as stored in every array that crosses the C ABI.
using z2 = double2;

// A stack of dense row-major matrices: element (r, c) of matrix e sits at
// p[e * stride + r * ld + c].  p == nullptr denotes the zero matrix.
struct Mat {
  z2* p = nullptr;
  int64_t stride = 0;
  int ld = 0;
  Mat() = default;
  Mat(z2* p_, int64_t s_, int ld_) : p(p_), stride(s_), ld(ld_) {}
  Mat(const void* p_, int64_t s_, int ld_)
      : p(const_cast<z2*>(static_cast<const z2*>(p_))), stride(s_), ld(ld_) {}
  bool zero() const { return p == nullptr; }
};

struct Cplx {
  double re, im;
};

// Pivot key of the batched inverses: a larger key means a larger |re| + |im|,
// ties go to the smaller row r (r < 512 in the low 9 bits).  The magnitude
// enters through the upper word of its double (sign 0, 11 exponent bits, the
// top mantissa bits), which orders non-negative doubles over the whole range;
// after the row bits, 11 mantissa bits remain, so candidates within a
// relative 2^-11 of each other count as equal.  A key below 512 (magnitude
// zero or below ~2^-1033) marks a zero pivot.
__device__ __forceinline__ unsigned pivot_key(double re, double im, int r) {
  const double m = fabs(re) + fabs(im);
  return ((unsigned)__double2hiint(m) & 0xFFFFFE00u) | (unsigned)(511 - r);
}

// 1 / (ar + i ai) = (ar - i ai) / (ar^2 + ai^2) for the pivots of the
// batched inverses.  Both parts are first scaled by the same power of two
// (exactly) so that the larger magnitude lies in [0.5, 1): the sum of squares
// d is then in [0.25, 2) and nothing overflows or underflows before the final
// rescale.  1 / d is the hardware reciprocal estimate refined by two Newton
// steps y <- y + y (1 - d y), each of which squares the relative error, so y
// is within a few ulp of 1 / d.  The dependent chain is ~10 operations
// against ~25 for three IEEE divisions (Smith's algorithm).  A zero pivot
// gives inf / nan, as a division by zero would; callers flag it separately.
__device__ __forceinline__ void complex_recip(double ar, double ai, double& qr, double& qi) {
  const int e = __builtin_amdgcn_frexp_exp(fmax(fabs(ar), fabs(ai)));
  const double sr = __builtin_amdgcn_ldexp(ar, -e), si = __builtin_amdgcn_ldexp(ai, -e);
  const double d = fma(sr, sr, si * si);
  double y = __builtin_amdgcn_rcp(d);
  y = fma(y, fma(-d, y, 1.0), y);
  y = fma(y, fma(-d, y, 1.0), y);
  qr = __builtin_amdgcn_ldexp(sr * y, -e);
  qi = __builtin_amdgcn_ldexp(-si * y, -e);
}

// Status carried up to the C ABI.
struct Error {
  int code;
  std::string msg;
};

#define QX_HIP_CHECK(expr)                                                   \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::qx::Error{QX_EHIP, std::string(#expr) + ": " +                 \
                                     hipGetErrorString(e_)};                 \
  } while (0)

#define QX_REQUIRE(cond, msg)                                                \
  do {                                                                       \
    if (!(cond)) throw ::qx::Error{QX_EINVAL, (msg)};                        \
  } while (0)

// ---- kernels (launchers) --------------------------------------------------

// Batched X = inv(A - S) for blocks n <= 32, one wave per matrix
// (rgf_small.hip); returns false, having done nothing, for larger n.
bool inv_small(hipStream_t st, int64_t batch, int n, const Mat& A, const Mat& S, const Mat& X,
               int* info, const int* mask = nullptr);

// Fused selected inversion for stacks of small blocks (rgf_small.hip);
// returns false, having done nothing, when the call is not eligible.
bool rgf_small_inv(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                   const void* const* obc_r, const qx_btd_layout* lx, void* x, int prec);

// Fused selected solve for stacks of small blocks (rgf_small.hip); returns
// false, having done nothing, when the call is not eligible.
bool rgf_small_solve(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                     const qx_btd_layout* lsl, const void* sl, const qx_btd_layout* lsg,
                     const void* sg, const void* const* obc_r, const void* const* obc_l,
                     const void* const* obc_g, const qx_btd_layout* lxl, void* xl,
                     const qx_btd_layout* lxg, void* xg, const qx_btd_layout* lxr, void* xr,
                     void* current, void* corner, double* density, double* ldos,
                     uint32_t flags, int prec);

// C = alpha op(A) op(B) + beta C for `batch` matrices; m x n output, inner k.
void zgemm(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
           Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& C,
           int precision);
// C = alpha op(A) op(B) + beta Cin (Cin may equal C; a zero Cin counts as 0).
// A non-zero Ct (n x m, aliasing no operand) also receives -C^H.
void zgemm_in(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
              Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& Cin,
              const Mat& C, int precision, const Mat& Ct = Mat());
// C = alpha op(A) op(B) + beta (D + E - E^H), n x n: the Hermitian-part sum
// (ew_herm_add) formed in the product's epilogue, rounded as that kernel
// rounds it.  C aliases neither D nor E.
void zgemm_hin(hipStream_t s, int64_t batch, int n, int k, int opA, int opB, Cplx alpha,
               const Mat& A, const Mat& B, Cplx beta, const Mat& D, const Mat& E, const Mat& C,
               int precision);
// Operands of one product C = alpha op(A) op(B) + beta (Cin + E - E^H), E
// optional (zero Mat), with the optional second output Ct = -C^H.
struct GemmOperands {
  Mat A, B, Cin, C, Ct, E;
};
// Two products of one shape, ops, alpha and beta (x and y, batch matrices
// each) as a single launch of 2 batch matrices when every operand of x is
// present exactly when its counterpart in y is (strides and leading
// dimensions may differ; an operand shared by both is passed twice), op(A)
// op(B) is not identically zero, and neither product reads or
// writes an output of the other; otherwise x, then y, as two launches.  Each
// element is computed as by the single-product call.
void zgemm_pair(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                Cplx alpha, Cplx beta, const GemmOperands& x, const GemmOperands& y,
                int precision);
// Three products the same way (x, y, z: 3 batch matrices in one launch).
void zgemm_triple(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                  Cplx alpha, Cplx beta, const GemmOperands& x, const GemmOperands& y,
                  const GemmOperands& z, int precision);

// Ozaki-split GEMM on bf16 matrix cores: C = alpha op(A) op(B) + beta Cin
// from `slices` 7-bit digit slices per operand (2..7); ws holds the digits.
size_t ozaki_workspace_bytes(int64_t batch, int m, int n, int k, int slices);
void zgemm_ozaki(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                 Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& Cin,
                 const Mat& C, int slices, void* ws, size_t ws_bytes);

// Modular-arithmetic (Ozaki scheme II) GEMM on int8 matrix cores for
// complex128 data with nmod moduli (8..20); ws holds the residues.
size_t oz2_workspace_bytes(int64_t batch, int m, int n, int k, int nmod);
void zgemm_oz2(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB, Cplx alpha,
               Cplx beta, const GemmOperands& x, int nmod, void* ws, size_t ws_bytes);
// Number of moduli of an Ozaki-II precision code, or 0 for any other code.
inline int oz2_moduli(int precision) {
  return (precision >= QX_PREC_OZAKI2(8) && precision <= QX_PREC_OZAKI2(20)) ? precision - 16 : 0;
}
// Number of digit slices of a bf16 Ozaki precision code, or 0.
inline int ozbf_slices(int precision) {
  return (precision >= QX_PREC_OZAKI_BF16(2) && precision <= QX_PREC_OZAKI_BF16(7))
             ? precision - 40
             : 0;
}
// Workspace of the split products (Ozaki II residues or bf16 digits) of one
// launch of batch products of at most m x n x k, or 0 for other precisions.
inline size_t split_workspace_bytes(int precision, int64_t batch, int m, int n, int k) {
  if (const int nmod = oz2_moduli(precision)) return oz2_workspace_bytes(batch, m, n, k, nmod);
  if (const int sl = ozbf_slices(precision)) return ozaki_workspace_bytes(batch, m, n, k, sl);
  return 0;
}
// Residue workspace of the calling thread for Ozaki-II products (set by the
// RGF drivers per sub-batch, by qx_zgemm_batched per call).
struct OzWorkspace {
  void* p = nullptr;
  size_t bytes = 0;
};
extern thread_local OzWorkspace active_oz_ws;

// Batched thin SVD A = U diag(S) V^H of d x m matrices (m <= d, m <= 512) by
// one-sided Jacobi rotations; U (d x m), V (m x m) contiguous row-major, S
// descending with stride sS; ws holds the rotated columns.
size_t svd_workspace_bytes(int64_t n, int d, int m);
void svd_tall(hipStream_t s, int64_t n, int d, int m, const Mat& A, const Mat& U, double* S,
              int64_t sS, const Mat& V, int max_sweeps, double tol, void* ws, size_t ws_bytes,
              int* sweeps);
// Spectral OBC mode classification (csrc/spectral.hip): mask[e, j] = 1 for
// the reflected modes; res and dk (nullable) receive the residuals and
// dE/dk.  ws (n, m); vs and the products a_ji V, a_ii V, a_ij V (n, b, m).
void spectral_modes(hipStream_t s, int64_t n, int b, int m, const z2* ws, const z2* vs,
                    const z2* am, const z2* a0, const z2* ap, double min_decay,
                    double max_decay, double min_propagation, double residual_tolerance,
                    double eta_decay, bool residual_normalization, unsigned char* mask,
                    double* res, z2* dk);

// X = inv(A - S) with partial pivoting (S may be zero).  info[e] is set to a
// nonzero value for an exactly singular matrix.  n <= kInvMaxN.  The scratch
// holds, per matrix, an n x n complex buffer (buf + e * buf_stride) and 3 n
// ints (ints + e * int_stride); A may alias X, or be zinv_input(n, X, scr);
// X may not alias buf.
constexpr int kInvMaxN = 512;
struct InvScratch {
  z2* buf;
  int64_t buf_stride;
  int* ints;
  int64_t int_stride;
};
// A non-null mask restricts the inverse to the matrices with mask[e] != 0;
// the others (X included) are left untouched.
void zinv(hipStream_t s, int64_t batch, int n, const Mat& A, const Mat& S,
          const Mat& X, int* info, int precision, const InvScratch& scr,
          const int* mask = nullptr);
// The buffer an inverse into X starts from (X itself or the scratch buffer):
// a matrix formed there and passed as A (S zero) is inverted into X without
// the initial copy.
Mat zinv_input(int n, const Mat& X, const InvScratch& scr);

// Elementwise helpers on batched m x n blocks.
// Y = a*A + b*B            (A, B may be zero matrices)
void ew_axpby(hipStream_t s, int64_t batch, int m, int n, Cplx a, const Mat& A,
              Cplx b, const Mat& B, const Mat& Y);
// Y = scale * (X - X^H)    (square)
// Contact self-energies: L = -f (S - S^H), G = (1 - f)(S - S^H), one f per
// matrix (device array).
void ew_contact_sigma(hipStream_t s, int64_t batch, int n, const Mat& S, const double* f,
                      const Mat& L, const Mat& G);
// A[e * sA + k] = z_e S_k + T_k - Sig[e * sSig + k] (S, T, Sig nullable = 0).
void ew_assemble(hipStream_t s, int64_t batch, int64_t nnz, const z2* z, const z2* S,
                 const z2* T, const z2* Sig, int64_t sSig, z2* A, int64_t sA);
// out[e * out_stride + r] = scale * Im X_e[r][r], r < n.
void ew_diag_imag(hipStream_t s, int64_t batch, int n, const Mat& X, double* out,
                  int64_t out_stride, double scale = 1.0);
// Newton-Schulz safety check of a complex64 block inverse X0 of M: R holds
// I - M X0 (formed in FP64).  Per matrix, flagged[e] = !(||R||_F < tol)
// (NaN-safe); a flagged matrix gets R = 0 (so the Newton-Schulz update
// leaves X unchanged) and info[e] |= 2.
void ew_ns_check(hipStream_t s, int64_t batch, int n, const Mat& R, double tol, int* flagged,
                 int* info);
// Y = I - Y for a stack of n x n matrices.
void ew_ident_sub(hipStream_t s, int64_t batch, int n, const Mat& Y);
// Y = D + scale (X - X^H) (Y may alias D, not X).
void ew_herm_add(hipStream_t s, int64_t batch, int n, const Mat& D, double scale,
                 const Mat& X, const Mat& Y);
void ew_antiherm(hipStream_t s, int64_t batch, int n, double scale,
                 const Mat& X, const Mat& Y);
// Y0 = scale (X0 - X0^H) and Y1 = scale (X1 - X1^H) as one launch (in order,
// as two launches, if an output of one is an operand of the other).
void ew_antiherm_pair(hipStream_t s, int64_t batch, int n, double scale, const Mat& X0,
                      const Mat& Y0, const Mat& X1, const Mat& Y1);
// Y += T^H - T             (square)
void ew_add_hdiff(hipStream_t s, int64_t batch, int n, const Mat& T,
                  const Mat& Y);
// Y = -X^H                 (X is m x n, Y is n x m)
void ew_neg_h(hipStream_t s, int64_t batch, int m, int n, const Mat& X,
              const Mat& Y);
// out[e * out_stride] = tr(X Y - Z W)   (square; complex)
void trace_diff(hipStream_t s, int64_t batch, int n, const Mat& X, const Mat& Y,
                const Mat& Z, const Mat& W, z2* out, int64_t out_stride);
// out[e] = || |A| + |B| ||_F   (square)
void absnorm2(hipStream_t s, int64_t batch, int n, const Mat& A, const Mat& B,
              double* out);
// Beyn: T = a_ji / z + a_ii + a_ij z for each (e, q) node; result stacks
// ordered e-major then node.
void contour_operator(hipStream_t s, int64_t batch, int nq, int b,
                      const Mat& aji, const Mat& aii, const Mat& aij,
                      const z2* z, const Mat& T);
// q0 += sum_q w0[q] Tinv[e, q]; q1 += sum_q w1[q] Tinv[e, q]
void contour_accumulate(hipStream_t s, int64_t batch, int nq, int b,
                        const Mat& Tinv, const z2* w0, const z2* w1,
                        const Mat& q0, const Mat& q1);

// Opt-in kernel timing: while a handle's profiler is active, the GEMM and
// inverse launchers bracket every launch with HIP events on the launch
// stream and record the algorithmic flops of the launch.
struct KernelTimer {
  enum Kind { GEMM = 0, INV = 1, NKIND = 2 };
  struct Rec {
    hipEvent_t a, b;
    int kind;
    double flops;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> recs;
  size_t next = 0;
  hipEvent_t t0 = nullptr;  // recorded on the caller's stream at start
  std::mutex mu;  // launches may come from several host threads
  hipEvent_t get();
  void begin(hipStream_t s, hipEvent_t* ev);
  void end(hipStream_t s, hipEvent_t a, int kind, double flops);
  ~KernelTimer();
};
extern thread_local KernelTimer* active_timer;
// GEMM tile shape of the calling handle: 0 automatic, 64 or 128.
extern thread_local int active_gemm_tile;

// Looks up this device's copy of the fixed scratch array of
// qx_inv_batched_noalloc (once per device; called by qx_create).
void noalloc_init(int device);

// Splits a batch of S matrices into up to `handle->parts` contiguous
// sub-batches and runs f(stream, e0, n) for each concurrently: sub-batch 0 on
// the caller's stream `st` in this thread, the others on handle-owned streams
// from helper host threads.  Events order all sub-batches after the work
// already queued on `st`, and `st` waits for all of them before returning.
void run_split(qx_handle* h, hipStream_t st, int64_t S,
               const std::function<void(hipStream_t, int64_t, int64_t)>& f);
// Index of the sub-batch the calling thread runs inside run_split (0 on the
// caller's thread).
extern thread_local int split_part;

}  // namespace qx

// Handle: device, workspace and sticky error message.
struct qx_handle {
  int device = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  int* info = nullptr;  // device error flags: bit 0 singular, bit 1 FP64 fallback
  size_t info_len = 0;
  int* mask = nullptr;  // per-matrix selection flags (device)
  size_t mask_len = 0;
  std::string last_error;
  qx::KernelTimer* timer = nullptr;
  int parts = 4;                      // concurrent sub-batches of one RGF call
  int gemm_tile = 0;                  // GEMM tile shape: 0 automatic, 64, 128
  std::vector<hipStream_t> streams;   // helper streams (default priority)
  std::vector<hipEvent_t> join;       // one per helper stream
  hipEvent_t fork = nullptr;
  void* ensure_ws(size_t bytes);
  int* ensure_info(size_t n);
  int* ensure_mask(size_t n);
};
