// Host drivers of the RGF recursions (selected inversion / selected solve).
//
// Restates src/qttools/greens_function_solver/rgf.py:25-414 as a sequence of
// batched device kernels on one stream: every block product is one launch of
// the MFMA GEMM over all matrices of the batch, every block inverse one
// launch of the batched Gauss-Jordan inverse.  Inputs and outputs stay in
// the caller's DSDBCOO-layout arrays (fully populated blocks are contiguous
// row-major slices, addressed with stack stride = nnz); the forward-sweep
// blocks live in a handle-owned workspace.  Products with a block that is not
// stored (e.g. Sigma_{i,i+1} of a block-diagonal self-energy) are skipped:
// they are exactly zero.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <future>

#include "qx_internal.h"

namespace qx {

namespace {

const Cplx kOne{1.0, 0.0}, kMinus{-1.0, 0.0}, kZero{0.0, 0.0};

// One BTD stack in DSDBCOO layout, optionally restricted to the matrices
// [e0, e0 + n) of the call (sub-batch).
struct Btd {
  const qx_btd_layout* L = nullptr;
  z2* base = nullptr;
  Btd() = default;
  Btd(const qx_btd_layout* l, const void* p) : L(l), base(const_cast<z2*>(static_cast<const z2*>(p))) {}
  Btd sub(int64_t e0) const {
    Btd b = *this;
    if (b.base) b.base += e0 * L->stack_stride;
    return b;
  }
  int bs(int i) const { return L->block_sizes[i]; }
  Mat at(int64_t off, int ld) const {
    if (!L || !base || off < 0) return Mat();
    return Mat(base + off, L->stack_stride, ld);
  }
  Mat diag(int i) const { return L ? at(L->diag[i], bs(i)) : Mat(); }
  Mat upper(int i) const { return L ? at(L->upper[i], bs(i + 1)) : Mat(); }  // (i, i+1)
  Mat lower(int i) const { return L ? at(L->lower[i], bs(i)) : Mat(); }      // (i+1, i)
};

void check_layout(const qx_btd_layout* l, const qx_btd_layout* ref, const char* what) {
  if (!l) return;
  QX_REQUIRE(l->n_blocks == ref->n_blocks, std::string(what) + ": block count differs from A");
  QX_REQUIRE(l->n_stack == ref->n_stack, std::string(what) + ": stack size differs from A");
  for (int i = 0; i < l->n_blocks; ++i)
    QX_REQUIRE(l->block_sizes[i] == ref->block_sizes[i],
               std::string(what) + ": block sizes differ from A");
}

// Workspace of equally sized matrix stacks ("slots"), each S x bmax^2;
// e0 selects the sub-batch.
struct Slots {
  z2* base;
  int64_t S;
  int64_t cap;  // elements per matrix
  int64_t e0;
  Mat get(int k, int ld) const { return Mat(base + (int64_t)k * S * cap + e0 * cap, cap, ld); }
  // inverse scratch: complex buffer in slot k, pivot ints in slot k + 1
  InvScratch inv(int k) const {
    return InvScratch{base + (int64_t)k * S * cap + e0 * cap, cap,
                      reinterpret_cast<int*>(base + (int64_t)(k + 1) * S * cap + e0 * cap),
                      4 * cap};
  }
};

// Per-block boundary self-energies, (S, b, b) contiguous, or absent.
struct Obc {
  const void* const* arr;
  int64_t e0;
  Mat get(int i, int b) const {
    if (!arr || !arr[i]) return Mat();
    return Mat(static_cast<const z2*>(arr[i]) + e0 * (int64_t)b * b, (int64_t)b * b, b);
  }
};

// dst = a + sign * s (either may be zero); dst is a workspace matrix.
void fold(hipStream_t st, int64_t S, int m, int n, const Mat& a, double sign,
          const Mat& s, const Mat& dst) {
  ew_axpby(st, S, m, n, kOne, a, Cplx{sign, 0.0}, s, dst);
}

// Arithmetic of the block products: QX_PREC_MIXED runs them in complex64,
// QX_PREC_OZAKI2(n) on int8 and QX_PREC_OZAKI_BF16(s) on bf16 matrix cores
// (the block inverses stay FP64).
int gemm_precision(int prec) {
  QX_REQUIRE(prec == QX_PREC_C128 || prec == QX_PREC_C64 || prec == QX_PREC_MIXED ||
                 prec == QX_PREC_C128_3M || oz2_moduli(prec) > 0 || ozbf_slices(prec) > 0,
             "unknown precision selector");
  return prec == QX_PREC_MIXED ? QX_PREC_C64 : prec;
}

// Split-product workspace (Ozaki II residues or bf16 digits) of one matrix
// for products of blocks <= bmax.
size_t oz2_bytes_per_matrix(int prec, int bmax) {
  const size_t b = split_workspace_bytes(prec, 1, bmax, bmax, bmax);
  return (b + 255) / 256 * 256;
}

// Sets the calling thread's Ozaki-II residue workspace to the part of
// sub-batch [e0, e0 + S) and clears it on scope exit.
struct OzSubBatch {
  OzSubBatch(int prec, void* region, int bmax, int64_t e0, int64_t S) {
    const size_t per = oz2_bytes_per_matrix(prec, bmax);
    if (per)
      active_oz_ws = OzWorkspace{static_cast<char*>(region) + e0 * per, (size_t)S * per};
  }
  ~OzSubBatch() { active_oz_ws = OzWorkspace{}; }
};

// In-place block inverse X <- inv(X) in the arithmetic `prec` selects.
// QX_PREC_MIXED inverts in complex64, then applies one Newton-Schulz step
// X <- X + X (I - M X): the residual I - M X is formed by the FP64 GEMM (its
// entries are of the order of the FP32 rounding error and would be lost in
// FP32), the small correction X R by the FP32 GEMM.  Mc and R are n x n
// temporaries.
// X = inv(M) for M in Ain, which is X or zinv_input(n, X, scr): the products
// that form M write it there, so the inverse starts without a copy.
//
// Safety net of the mixed mode: since I - M X1 = (I - M X0)^2 for the
// Newton-Schulz iterate X1, ||R0||_F = ||I - M X0||_F < tol bounds the
// residual of X1 by tol^2.  A matrix whose complex64 inverse misses that
// (or is not finite: entries outside the FP32 range, a pivot lost to FP32
// rounding) is inverted again by the FP64 pivoted inverse, selected per
// matrix on the device through `mask`, and flagged with info bit 1
// (reported by qx_sync_check as QX_FP64_FALLBACK).
// Block inverses of the mixed arithmetic: the FP64 inverse (default), or,
// with QX_MIXED_INVERSE=c64, the complex64 inverse refined by one FP64
// Newton-Schulz step with the FP64 re-inversion of flagged blocks.  The
// inverse kernels are bound by their pivot chains, not by arithmetic, so the
// complex64 inverse is barely faster than the FP64 one (523 vs 584 us for
// 32 x 256^2) and the refinement costs more than it saves.
bool mixed_inverse_c64() {
  const char* e = std::getenv("QX_MIXED_INVERSE");
  return e && std::string(e) == "c64";
}

void block_inv(hipStream_t st, int64_t S, int n, const Mat& Ain, const Mat& X, int* info,
               int prec, const InvScratch& scr, const Mat& Mc, const Mat& R, int* mask) {
  if (prec != QX_PREC_MIXED || !mixed_inverse_c64()) {
    zinv(st, S, n, Ain, Mat(), X, info, prec == QX_PREC_C64 ? QX_PREC_C64 : QX_PREC_C128, scr);
    return;
  }
  static const double tol = [] {
    const char* e = std::getenv("QX_MIXED_TOL");
    const double v = e ? std::atof(e) : 0.0;
    return v > 0.0 ? v : 1e-2;
  }();
  ew_axpby(st, S, n, n, kOne, Ain, kZero, Mat(), Mc);  // M
  zinv(st, S, n, Ain, Mat(), X, info, QX_PREC_C64, scr);
  zgemm(st, S, n, n, n, QX_OP_N, QX_OP_N, kOne, Mc, X, kZero, R, QX_PREC_C128);
  ew_ident_sub(st, S, n, R);  // R = I - M X0
  ew_ns_check(st, S, n, R, tol, mask, info);
  zinv(st, S, n, Mc, Mat(), X, info, QX_PREC_C128, scr, mask);  // flagged: FP64 inverse
  ew_axpby(st, S, n, n, kOne, X, kZero, Mat(), Mc);  // X0 (FP64 inverse where flagged)
  // X0 + X0 R; a flagged matrix has R = 0 and keeps its FP64 inverse
  zgemm(st, S, n, n, n, QX_OP_N, QX_OP_N, kOne, Mc, R, kOne, X, QX_PREC_C64);
}

int block_limits(const qx_btd_layout* la, const char* what) {
  int bmax = 1;
  for (int i = 0; i < la->n_blocks; ++i) {
    QX_REQUIRE(la->block_sizes[i] >= 1 && la->block_sizes[i] <= kInvMaxN,
               std::string(what) + ": block sizes must be in [1, 512]");
    bmax = std::max(bmax, (int)la->block_sizes[i]);
  }
  return bmax;
}

}  // namespace

// Bytes of the slot stacks (before the Ozaki-II residue region).
size_t rgf_slot_bytes(const qx_btd_layout* la, bool solve, int bmax) {
  const int64_t slots = solve ? 3 * (int64_t)la->n_blocks + 26 : (int64_t)la->n_blocks + 10;  // see slot maps
  return (size_t)(slots * la->n_stack * (int64_t)bmax * bmax) * sizeof(z2);
}

size_t rgf_workspace_bytes(const qx_btd_layout* la, bool solve, int prec) {
  int bmax = 1;
  for (int i = 0; i < la->n_blocks; ++i) bmax = std::max(bmax, (int)la->block_sizes[i]);
  size_t bytes = rgf_slot_bytes(la, solve, bmax);
  // Ozaki-II products: residues of one product per matrix of the call
  const size_t per = oz2_bytes_per_matrix(prec, bmax);
  bytes += (size_t)la->n_stack * per + (per ? 512 : 0);
  // the fused small-block kernels keep 6 N + 23 (solve) or 2 N + 3 (inverse)
  // slots of 32 x 32 per energy
  if (la->n_blocks <= 16 && bmax <= 32) {
    const int64_t fs = solve ? 6 * (int64_t)la->n_blocks + 23 : 2 * (int64_t)la->n_blocks + 3;
    bytes = std::max(bytes, (size_t)(fs * la->n_stack * 1024) * sizeof(z2));
  }
  return bytes;
}

void rgf_selected_inv(qx_handle* h, hipStream_t st0, const qx_btd_layout* la,
                      const void* a, const void* const* obc_r,
                      const qx_btd_layout* lx, void* x, int prec) {
  QX_REQUIRE(la && lx && a && x, "selected_inv: null argument");
  QX_REQUIRE(la->n_blocks >= 1, "selected_inv: need at least one block");
  check_layout(lx, la, "x");
  const int N = la->n_blocks;
  const int64_t Stot = la->n_stack;
  if (Stot == 0) return;
  const int bmax = block_limits(la, "selected_inv");
  gemm_precision(prec);  // validates the selector
  if (rgf_small_inv(h, st0, la, a, obc_r, lx, x, prec)) return;
  z2* ws = static_cast<z2*>(h->ensure_ws(rgf_workspace_bytes(la, false, prec)));
  int* info_all = h->ensure_info(Stot);
  int* mask_all = prec == QX_PREC_MIXED ? h->ensure_mask(Stot) : nullptr;
  const Btd A0(la, a), X0(lx, x);
  char* oz_region = reinterpret_cast<char*>(ws) + rgf_slot_bytes(la, false, bmax);

  const int gp = gemm_precision(prec);
  // QX_RGF_PAIR=0 (read per call): one launch per backward product
  const char* pair_env = std::getenv("QX_RGF_PAIR");
  const bool pair_on = !(pair_env && pair_env[0] == '0');
  run_split(h, st0, Stot, [&](hipStream_t st, int64_t e0, int64_t S) {
    const OzSubBatch oz(prec, oz_region, bmax, e0, S);
    const Btd A = A0.sub(e0), X = X0.sub(e0);
    const Slots W{ws, Stot, (int64_t)bmax * bmax, e0};
    const Obc OR{obc_r, e0};
    int* info = info_all + e0;
    int* mask = mask_all ? mask_all + e0 : nullptr;
    std::vector<int> xd(N);
    for (int i = 0; i < N; ++i) xd[i] = i;
    int spare = N;
    const int t1 = N + 1, t2 = N + 2, t3 = N + 3;
    const InvScratch scr = W.inv(N + 4);  // N + 6, N + 7: Newton-Schulz temporaries
    auto bs = [&](int i) { return A.bs(i); };

    // Forward sweep (rgf.py:70-92).
    {
      const Mat X0 = W.get(xd[0], bs(0)), M0 = zinv_input(bs(0), X0, scr);
      fold(st, S, bs(0), bs(0), A.diag(0), -1.0, OR.get(0, bs(0)), M0);
      block_inv(st, S, bs(0), M0, X0, info, prec, scr, W.get(N + 6, bs(0)), W.get(N + 7, bs(0)),
                mask);
    }
    for (int i = 0; i + 1 < N; ++i) {
      const int j = i + 1, bi = bs(i), bj = bs(j);
      const Mat T1 = W.get(t1, bi);
      zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, A.lower(i), W.get(xd[i], bi), kZero, T1, gp);
      const Mat Xj = W.get(xd[j], bj), Mj = zinv_input(bj, Xj, scr);
      Mat Ajj = A.diag(j);
      if (!OR.get(j, bj).zero()) {
        fold(st, S, bj, bj, A.diag(j), -1.0, OR.get(j, bj), Mj);
        Ajj = Mj;
      }
      zgemm_in(st, S, bj, bj, bi, QX_OP_N, QX_OP_N, kMinus, T1, A.upper(i), kOne, Ajj, Mj, gp);
      block_inv(st, S, bj, Mj, Xj, info, prec, scr, W.get(N + 6, bj), W.get(N + 7, bj), mask);
    }
    // rgf.py:94-95
    if (!X.diag(N - 1).zero())
      ew_axpby(st, S, bs(N - 1), bs(N - 1), kOne, W.get(xd[N - 1], bs(N - 1)), kZero, Mat(),
               X.diag(N - 1));

    // Backward sweep (rgf.py:97-110).
    for (int i = N - 2; i >= 0; --i) {
      const int j = i + 1, bi = bs(i), bj = bs(j);
      const Mat Xii = W.get(xd[i], bi), Xjj = W.get(xd[j], bj);
      const Mat T1 = W.get(t1, bi), T2 = W.get(t2, bj);
      Mat Xji = X.lower(i);
      if (Xji.zero()) Xji = W.get(t3, bi);
      const Mat Xij = X.upper(i);
      if (pair_on && bi == bj) {
        // the x_ji and x_ij chains below, each product pair as one launch
        const Mat none;
        zgemm_pair(st, S, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
                   {Xjj, A.lower(i), T1, T1, none, none}, {Xii, A.upper(i), T2, T2, none, none},
                   gp);
        if (!Xij.zero())
          zgemm_pair(st, S, bi, bi, bi, QX_OP_N, QX_OP_N, kMinus, kZero,
                     {T1, Xii, Xji, Xji, none, none}, {T2, Xjj, Xij, Xij, none, none}, gp);
        else
          zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kMinus, T1, Xii, kZero, Xji, gp);
      } else {
        // x_ji = -x_jj a_ji x_ii
        zgemm(st, S, bj, bi, bj, QX_OP_N, QX_OP_N, kOne, Xjj, A.lower(i), kZero, T1, gp);
        zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kMinus, T1, Xii, kZero, Xji, gp);
        // x_ij = -x_ii a_ij x_jj
        zgemm(st, S, bi, bj, bi, QX_OP_N, QX_OP_N, kOne, Xii, A.upper(i), kZero, T2, gp);
        if (!Xij.zero())
          zgemm(st, S, bi, bj, bj, QX_OP_N, QX_OP_N, kMinus, T2, Xjj, kZero, Xij, gp);
      }
      // x_ii <- x_ii - x_ii a_ij x_ji
      const Mat Z = W.get(spare, bi);
      zgemm_in(st, S, bi, bi, bj, QX_OP_N, QX_OP_N, kMinus, T2, Xji, kOne, Xii, Z, gp);
      std::swap(spare, xd[i]);
      if (!X.diag(i).zero())
        ew_axpby(st, S, bi, bi, kOne, W.get(xd[i], bi), kZero, Mat(), X.diag(i));
    }
  });
}

void rgf_selected_solve(qx_handle* h, hipStream_t st0, const qx_btd_layout* la,
                        const void* a, const qx_btd_layout* lsl, const void* sl,
                        const qx_btd_layout* lsg, const void* sg,
                        const void* const* obc_r, const void* const* obc_l,
                        const void* const* obc_g, const qx_btd_layout* lxl,
                        void* xl, const qx_btd_layout* lxg, void* xg,
                        const qx_btd_layout* lxr, void* xr, void* current,
                        void* corner, double* density, double* ldos, uint32_t flags,
                        int prec) {
  QX_REQUIRE(la && a, "selected_solve: null A");
  QX_REQUIRE(lxl && xl && lxg && xg, "selected_solve: null lesser/greater output");
  const bool want_r = (flags & QX_SOLVE_RETARDED) != 0;
  const bool want_r_off = (flags & QX_SOLVE_RETARDED_OFFDIAG) != 0;
  const bool want_cur = (flags & QX_SOLVE_CURRENT) != 0;
  QX_REQUIRE(!want_r || (lxr && xr), "selected_solve: retarded output requested but null");
  QX_REQUIRE(!want_cur || current, "selected_solve: current requested but null");
  check_layout(lsl, la, "sigma_lesser");
  check_layout(lsg, la, "sigma_greater");
  check_layout(lxl, la, "x_lesser");
  check_layout(lxg, la, "x_greater");
  if (want_r) check_layout(lxr, la, "x_retarded");
  const int N = la->n_blocks;
  const int64_t Stot = la->n_stack;
  QX_REQUIRE(N >= 1, "selected_solve: need at least one block");
  if (Stot == 0) return;
  const int bmax = block_limits(la, "selected_solve");
  if (rgf_small_solve(h, st0, la, a, lsl, sl, lsg, sg, obc_r, obc_l, obc_g, lxl, xl, lxg, xg,
                      want_r ? lxr : nullptr, want_r ? xr : nullptr, want_cur ? current : nullptr,
                      corner, density, ldos, flags, prec))
    return;
  z2* ws = static_cast<z2*>(h->ensure_ws(rgf_workspace_bytes(la, true, prec)));
  int* info_all = h->ensure_info(Stot);
  int* mask_all = prec == QX_PREC_MIXED ? h->ensure_mask(Stot) : nullptr;
  char* oz_region = reinterpret_cast<char*>(ws) + rgf_slot_bytes(la, true, bmax);
  const Btd A0(la, a), SL0(lsl, sl), SG0(lsg, sg), XL0(lxl, xl), XG0(lxg, xg),
      XR0(want_r ? lxr : nullptr, want_r ? xr : nullptr);

  const int gp = gemm_precision(prec);
  // The odd sub-batches start after sub-batch 0's first block inverse, so
  // that the sub-batches' latency-bound inverse phases do not coincide and
  // the GEMMs of one fill the GPU during the inverse of another
  // (QX_RGF_STAGGER=0, read per call, starts all together).
  const char* stg_env = std::getenv("QX_RGF_STAGGER");
  const bool stagger = !(stg_env && stg_env[0] == '0');
  const char* pair_env = std::getenv("QX_RGF_PAIR");
  const bool pair_on = !(pair_env && pair_env[0] == '0');
  hipEvent_t stg_ev = nullptr;
  if (stagger) QX_HIP_CHECK(hipEventCreateWithFlags(&stg_ev, hipEventDisableTiming));
  std::promise<void> stg_set;
  std::shared_future<void> stg_ready = stg_set.get_future().share();
  run_split(h, st0, Stot, [&](hipStream_t st, int64_t e0, int64_t S) {
  const OzSubBatch oz(prec, oz_region, bmax, e0, S);
  const int part = split_part;
  // sub-batch 0 releases the others once its first inverse is enqueued, or
  // when it leaves early with an error
  struct Release {
    std::promise<void>* p;
    ~Release() {
      if (p) p->set_value();
    }
  } release{stagger && part == 0 ? &stg_set : nullptr};
  if (stagger && part % 2 == 1) {
    stg_ready.wait();
    QX_HIP_CHECK(hipStreamWaitEvent(st, stg_ev, 0));
  }
  Btd A = A0.sub(e0), SL = SL0.sub(e0), SG = SG0.sub(e0), XL = XL0.sub(e0), XG = XG0.sub(e0),
      XR = XR0.sub(e0);
  const Slots W{ws, Stot, (int64_t)bmax * bmax, e0};
  const Obc OR{obc_r, e0}, OL{obc_l, e0}, OG{obc_g, e0};
  int* info = info_all + e0;
  int* mask = mask_all ? mask_all + e0 : nullptr;
  z2* cur = current ? static_cast<z2*>(current) + e0 * (N - 1) : nullptr;
  // density Im diag(G^<): row of n_orb per matrix, block i at orb_off[i]
  std::vector<int64_t> orb_off(N + 1, 0);
  for (int i = 0; i < N; ++i) orb_off[i + 1] = orb_off[i] + A.bs(i);
  const int64_t n_orb = orb_off[N];
  double* dens = density ? density + e0 * n_orb : nullptr;
  // LDOS -Im diag(G^R) / (2 pi), same layout, as each G^R_ii is finished
  double* lds = ldos ? ldos + e0 * n_orb : nullptr;
  constexpr double kLdos = -0.15915494309189533577;  // -1 / (2 pi)
  // corner block G^R_{0,N-1}: ping-pong slots + product temp, b0 x b_j
  // matrices with leading dimension bmax
  const int b0 = A.bs(0), bl = A.bs(N - 1);
  int gc_cur = 3 * N + 17, gc_nxt = 3 * N + 18;
  const int gc_tmp = 3 * N + 19;
  auto bs = [&](int i) { return A.bs(i); };

  // slot bookkeeping: forward blocks (rotated in the backward sweep) + temps
  std::vector<int> sr(N), sl_(N), sg_(N);
  for (int i = 0; i < N; ++i) {
    sr[i] = i;
    sl_[i] = N + i;
    sg_[i] = 2 * N + i;
  }
  int spare = 3 * N;
  const int T1s = 3 * N + 1, T3s = 3 * N + 3, T4s = 3 * N + 4, T5s = 3 * N + 5,
            T6s = 3 * N + 6, B1s = 3 * N + 7, B2s = 3 * N + 8, B3s = 3 * N + 9, B4s = 3 * N + 10,
            C1s = 3 * N + 11, C2s = 3 * N + 12, C3s = 3 * N + 13, C4s = 3 * N + 14;
  const InvScratch scr = W.inv(3 * N + 15);
  const int NS1 = 3 * N + 20, NS2 = 3 * N + 21;  // Newton-Schulz temporaries
  // G^> temporaries, so that each G^< product and its G^> counterpart can run
  // as one launch (QX_RGF_PAIR=0, read per call: one launch each)
  const int T3g = 3 * N + 22, T4g = 3 * N + 23, T5g = 3 * N + 24, T6g = 3 * N + 25;
  int spare_g = T6g;

  struct Q {
    Btd* sig;
    Btd* out;
    std::vector<int>* slot;
    const Obc* obc;
    bool skip_lower;
    int T3, T4, T5, T6;  // temporaries of this quantity
    int* spare;          // free slot of the backward sweep, rotated with slot
  };
  Q qs[2] = {{&SL, &XL, &sl_, &OL, (flags & QX_SOLVE_SKIP_LOWER_L) != 0, T3s, T4s, T5s, T6s,
              &spare},
             {&SG, &XG, &sg_, &OG, (flags & QX_SOLVE_SKIP_LOWER_G) != 0, T3g, T4g, T5g, T6g,
              &spare_g}};
  // prod(h0, h1, ...): product x (if h0) and product y (if h1) of one shape,
  // typically a G^< product and its G^> counterpart; one launch when both
  // are wanted and pairing is on
  auto one = [&](int m, int n, int k, int opA, int opB, Cplx alpha, Cplx beta,
                 const GemmOperands& o) {
    if (o.E.zero())
      zgemm_in(st, S, m, n, k, opA, opB, alpha, o.A, o.B, beta, o.Cin, o.C, gp, o.Ct);
    else
      zgemm_hin(st, S, n, k, opA, opB, alpha, o.A, o.B, beta, o.Cin, o.E, o.C, gp);
  };
  auto prod = [&](bool h0, bool h1, int m, int n, int k, int opA, int opB, Cplx alpha,
                  Cplx beta, const GemmOperands& x, const GemmOperands& y) {
    if (h0 && h1 && pair_on) {
      zgemm_pair(st, S, m, n, k, opA, opB, alpha, beta, x, y, gp);
      return;
    }
    if (h0) one(m, n, k, opA, opB, alpha, beta, x);
    if (h1) one(m, n, k, opA, opB, alpha, beta, y);
  };
  auto slot = [&](int qi, int blk, int b) { return W.get((*qs[qi].slot)[blk], b); };
  const Mat none;

  // ---- block 0 (rgf.py:207-231) ----
  {
    const int b0 = bs(0);
    const Mat X0 = W.get(sr[0], b0), M0 = zinv_input(b0, X0, scr);
    fold(st, S, b0, b0, A.diag(0), -1.0, OR.get(0, b0), M0);
    block_inv(st, S, b0, M0, X0, info, prec, scr, W.get(NS1, b0), W.get(NS2, b0), mask);
    if (release.p) {
      QX_HIP_CHECK(hipEventRecord(stg_ev, st));
      release.p->set_value();
      release.p = nullptr;
    }
    if (corner) ew_axpby(st, S, b0, b0, kOne, X0, kZero, Mat(), W.get(gc_cur, bmax));
    Mat T3[2], T6[2], X[2];
    for (int qi = 0; qi < 2; ++qi) {
      const Q& q = qs[qi];
      T3[qi] = W.get(q.T3, b0);
      T6[qi] = W.get(q.T6, b0);
      X[qi] = slot(qi, 0, b0);
      fold(st, S, b0, b0, q.sig->diag(0), 1.0, q.obc->get(0, b0), T3[qi]);
    }
    prod(true, true, b0, b0, b0, QX_OP_N, QX_OP_N, kOne, kZero,
         {X0, T3[0], T6[0], T6[0], none, none}, {X0, T3[1], T6[1], T6[1], none, none});
    prod(true, true, b0, b0, b0, QX_OP_N, QX_OP_C, kOne, kZero,
         {T6[0], X0, X[0], X[0], none, none}, {T6[1], X0, X[1], X[1], none, none});
  }

  // ---- forward sweep (rgf.py:234-295) ----
  for (int i = 0; i + 1 < N; ++i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const Mat Xri = W.get(sr[i], bi), Xrj = W.get(sr[j], bj);
    const Mat T1 = W.get(T1s, bi);
    // a_ji xr_ii; with pairing on, in one launch with a_ji x^<_ii and
    // a_ji x^>_ii of the Σ side below (inputs of the previous step only)
    const bool aji3 = pair_on;
    if (aji3) {
      const Mat Aji = A.lower(i), none;
      const Mat t4l = W.get(qs[0].T4, bi), t4g = W.get(qs[1].T4, bi);
      const Mat xl = slot(0, i, bi), xg = slot(1, i, bi);
      zgemm_triple(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
                   {Aji, Xri, T1, T1, none, none}, {Aji, xl, t4l, t4l, none, none},
                   {Aji, xg, t4g, t4g, none, none}, gp);
    } else {
      zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, A.lower(i), Xri, kZero, T1, gp);
    }
    // xr_jj = inv(a_jj - obc - a_ji xr_ii a_ij), the argument formed where
    // the inverse starts
    const Mat Mj = zinv_input(bj, Xrj, scr);
    Mat Ajj = A.diag(j);
    if (!OR.get(j, bj).zero()) {
      fold(st, S, bj, bj, A.diag(j), -1.0, OR.get(j, bj), Mj);
      Ajj = Mj;
    }
    zgemm_in(st, S, bj, bj, bi, QX_OP_N, QX_OP_N, kMinus, T1, A.upper(i), kOne, Ajj, Mj, gp);
    block_inv(st, S, bj, Mj, Xrj, info, prec, scr, W.get(NS1, bj), W.get(NS2, bj), mask);
    if (corner) {
      // g_{0,j} = -(g_{0,i} a_ij) x^R_jj
      zgemm(st, S, b0, bj, bi, QX_OP_N, QX_OP_N, kOne, W.get(gc_cur, bmax), A.upper(i), kZero,
            W.get(gc_tmp, bmax), gp);
      zgemm(st, S, b0, bj, bj, QX_OP_N, QX_OP_N, kMinus, W.get(gc_tmp, bmax), Xrj, kZero,
            W.get(gc_nxt, bmax), gp);
      std::swap(gc_cur, gc_nxt);
    }
    {
      // sigma_jj + obc + a_ji x_ii a_ji^H + u^H - u,  u = a_ji xr_ii sigma_ij
      Mat Sjj[2], T3[2], T4[2], T5[2], T6[2], Xi[2], Xj[2], sij[2];
      for (int qi = 0; qi < 2; ++qi) {
        const Q& q = qs[qi];
        T3[qi] = W.get(q.T3, bj);
        T4[qi] = W.get(q.T4, bi);
        T5[qi] = W.get(q.T5, bj);
        T6[qi] = W.get(q.T6, bj);
        Xi[qi] = slot(qi, i, bi);
        Xj[qi] = slot(qi, j, bj);
        sij[qi] = q.sig->upper(i);
        Sjj[qi] = q.sig->diag(j);
        if (!q.obc->get(j, bj).zero()) {
          fold(st, S, bj, bj, q.sig->diag(j), 1.0, q.obc->get(j, bj), T3[qi]);
          Sjj[qi] = T3[qi];
        }
      }
      const Mat Aji = A.lower(i);
      if (!aji3)
        prod(true, true, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
             {Aji, Xi[0], T4[0], T4[0], none, none}, {Aji, Xi[1], T4[1], T4[1], none, none});
      prod(true, true, bj, bj, bi, QX_OP_N, QX_OP_C, kOne, kOne,
           {T4[0], Aji, Sjj[0], T3[0], none, none}, {T4[1], Aji, Sjj[1], T3[1], none, none});
      prod(!sij[0].zero(), !sij[1].zero(), bj, bj, bi, QX_OP_N, QX_OP_N, kOne, kZero,
           {T1, sij[0], T5[0], T5[0], none, none}, {T1, sij[1], T5[1], T5[1], none, none});
      for (int qi = 0; qi < 2; ++qi)
        if (!sij[qi].zero()) ew_add_hdiff(st, S, bj, T5[qi], T3[qi]);
      prod(true, true, bj, bj, bj, QX_OP_N, QX_OP_N, kOne, kZero,
           {Xrj, T3[0], T6[0], T6[0], none, none}, {Xrj, T3[1], T6[1], T6[1], none, none});
      prod(true, true, bj, bj, bj, QX_OP_N, QX_OP_C, kOne, kZero,
           {T6[0], Xrj, Xj[0], Xj[0], none, none}, {T6[1], Xrj, Xj[1], Xj[1], none, none});
    }
  }

  if (corner)
    ew_axpby(st, S, b0, bl, kOne, W.get(gc_cur, bmax), kZero, Mat(),
             Mat(static_cast<z2*>(corner) + e0 * (int64_t)b0 * bl, (int64_t)b0 * bl, bl));

  // ---- last diagonal blocks (rgf.py:297-305) ----
  {
    const int bn = bs(N - 1);
    if (dens) ew_diag_imag(st, S, bn, W.get(sl_[N - 1], bn), dens + orb_off[N - 1], n_orb);
    const Mat d0 = qs[0].out->diag(N - 1), d1 = qs[1].out->diag(N - 1);
    if (pair_on && !d0.zero() && !d1.zero()) {
      ew_antiherm_pair(st, S, bn, 0.5, slot(0, N - 1, bn), d0, slot(1, N - 1, bn), d1);
    } else {
      for (auto& q : qs)
        if (!q.out->diag(N - 1).zero())
          ew_antiherm(st, S, bn, 0.5, W.get((*q.slot)[N - 1], bn), q.out->diag(N - 1));
    }
    if (want_r && !XR.diag(N - 1).zero())
      ew_axpby(st, S, bn, bn, kOne, W.get(sr[N - 1], bn), kZero, Mat(), XR.diag(N - 1));
    if (lds) ew_diag_imag(st, S, bn, W.get(sr[N - 1], bn), lds + orb_off[N - 1], n_orb, kLdos);
  }

  // ---- backward sweep (rgf.py:308-404) ----
  // gr[i]: full G^R_ii once computed (the caller's output block when it is
  // written, else a workspace slot)
  std::vector<Mat> gr(N);
  gr[N - 1] = W.get(sr[N - 1], bs(N - 1));
  for (int i = N - 2; i >= 0; --i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const Mat Xri = W.get(sr[i], bi), Xrj = gr[j];
    const Mat Aij = A.upper(i), Aji = A.lower(i);
    const Mat B1 = W.get(B1s, bj), B2 = W.get(B2s, bi), B3 = W.get(B3s, bj), B4 = W.get(B4s, bi);
    // B1 = xr_ii a_ij, B2 = xr_jj a_ji, B3 = B1 xr_jj (xr_ii a_ij xr_jj),
    // B4 = B1 B2 (... a_ji): two pairs of one shape when bi == bj
    if (bi == bj) {
      prod(true, true, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
           {Xri, Aij, B1, B1, none, none}, {Xrj, Aji, B2, B2, none, none});
      prod(true, true, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
           {B1, Xrj, B3, B3, none, none}, {B1, B2, B4, B4, none, none});
    } else {
      zgemm(st, S, bi, bj, bi, QX_OP_N, QX_OP_N, kOne, Xri, Aij, kZero, B1, gp);
      zgemm(st, S, bj, bi, bj, QX_OP_N, QX_OP_N, kOne, Xrj, Aji, kZero, B2, gp);
      zgemm(st, S, bi, bj, bj, QX_OP_N, QX_OP_N, kOne, B1, Xrj, kZero, B3, gp);
      zgemm(st, S, bi, bi, bj, QX_OP_N, QX_OP_N, kOne, B1, B2, kZero, B4, gp);
    }

    if (want_cur) {
      // Meir-Wingreen layer current from the forward (left-connected) blocks
      // of layer i and the full blocks of layer j (rgf.py:381-400).
      const Mat C1 = W.get(C1s, bi), C2 = W.get(C2s, bj), C3 = W.get(C3s, bj), C4 = W.get(C4s, bj);
      zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, Aji, Xri, kZero, C1, gp);  // a_ji xr_ii
      Mat tl[2] = {C2, C3};
      for (int qi = 0; qi < 2; ++qi) {
        const Q& q = qs[qi];
        const Mat T4 = W.get(T4s, bi);
        zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kOne, Aji, W.get((*q.slot)[i], bi), kZero, T4, gp);
        zgemm(st, S, bj, bj, bi, QX_OP_N, QX_OP_C, kOne, T4, Aji, kZero, tl[qi], gp);
        const Mat sij = q.sig->upper(i);
        if (!sij.zero()) {
          zgemm(st, S, bj, bj, bi, QX_OP_N, QX_OP_N, kOne, C1, sij, kZero, C4, gp);
          ew_add_hdiff(st, S, bj, C4, tl[qi]);
        }
      }
      // tr(sigma_greater_tilde xl_jj - xg_jj sigma_lesser_tilde)
      trace_diff(st, S, bj, C3, W.get(sl_[j], bj), W.get(sg_[j], bj), C2,
                 cur + i, N - 1);
    }

    {
      // G^< temporaries T1, T5, T4 (x_ij when the output has no upper
      // block); G^> ones in its own slots
      Mat Xi[2], Xj[2], T1[2], T5[2], Z[2], Oij[2], low[2], sij[2];
      bool hs[2];
      for (int qi = 0; qi < 2; ++qi) {
        const Q& q = qs[qi];
        Xi[qi] = slot(qi, i, bi);
        Xj[qi] = slot(qi, j, bj);
        T1[qi] = W.get(qi == 0 ? T1s : q.T3, bi);
        T5[qi] = W.get(q.T5, bj);
        Z[qi] = W.get(*q.spare, bi);
        sij[qi] = q.sig->upper(i);
        hs[qi] = !sij[qi].zero();
        Oij[qi] = q.out->upper(i);
        if (Oij[qi].zero()) Oij[qi] = W.get(q.T4, bj);
        low[qi] = (!q.skip_lower) ? q.out->lower(i) : Mat();
      }
      // temp_1x = B4 x_ii - xr_ii s_ij B3^H
      prod(true, true, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, kZero,
           {B4, Xi[0], T1[0], T1[0], none, none}, {B4, Xi[1], T1[1], T1[1], none, none});
      prod(hs[0], hs[1], bi, bj, bi, QX_OP_N, QX_OP_N, kOne, kZero,
           {Xri, sij[0], T5[0], T5[0], none, none}, {Xri, sij[1], T5[1], T5[1], none, none});
      prod(hs[0], hs[1], bi, bi, bj, QX_OP_N, QX_OP_C, kMinus, kOne,
           {T5[0], B3, T1[0], T1[0], none, none}, {T5[1], B3, T1[1], T1[1], none, none});
      // x_ij = -temp_2x - x_ii B2^H + xr_ii s_ij xr_jj^H,  temp_2x = B1 x_jj
      prod(true, true, bi, bj, bj, QX_OP_N, QX_OP_N, kMinus, kZero,
           {B1, Xj[0], Oij[0], Oij[0], none, none}, {B1, Xj[1], Oij[1], Oij[1], none, none});
      // x_ii' = x_ii + (temp_1x - temp_1x^H) + temp_2x B1^H, the first two
      // terms formed in the epilogue of this product
      prod(true, true, bi, bi, bj, QX_OP_N, QX_OP_C, kMinus, kOne,
           {Oij[0], B1, Xi[0], Z[0], none, T1[0]}, {Oij[1], B1, Xi[1], Z[1], none, T1[1]});
      // the GEMM that completes x_ij also writes the lower block -x_ij^H
      prod(true, true, bi, bj, bi, QX_OP_N, QX_OP_C, kMinus, kOne,
           {Xi[0], B2, Oij[0], Oij[0], hs[0] ? none : low[0], none},
           {Xi[1], B2, Oij[1], Oij[1], hs[1] ? none : low[1], none});
      prod(hs[0], hs[1], bi, bj, bj, QX_OP_N, QX_OP_C, kOne, kOne,
           {T5[0], Xrj, Oij[0], Oij[0], low[0], none}, {T5[1], Xrj, Oij[1], Oij[1], low[1], none});
      const Mat d0 = qs[0].out->diag(i), d1 = qs[1].out->diag(i);
      const bool proj2 = pair_on && !d0.zero() && !d1.zero();
      if (proj2) ew_antiherm_pair(st, S, bi, 0.5, Z[0], d0, Z[1], d1);
      for (int qi = 0; qi < 2; ++qi) {
        const Q& q = qs[qi];
        std::swap(*q.spare, (*q.slot)[i]);
        if (!proj2 && !q.out->diag(i).zero()) ew_antiherm(st, S, bi, 0.5, Z[qi], q.out->diag(i));
        if (dens && qi == 0) ew_diag_imag(st, S, bi, Z[qi], dens + orb_off[i], n_orb);
      }
    }

    // retarded: xr_ii <- xr_ii + B4 xr_ii ; off-diagonal blocks on request
    if (want_r && want_r_off) {
      if (!XR.upper(i).zero()) ew_axpby(st, S, bi, bj, kMinus, B3, kZero, Mat(), XR.upper(i));
      if (!XR.lower(i).zero())
        zgemm(st, S, bj, bi, bi, QX_OP_N, QX_OP_N, kMinus, B2, Xri, kZero, XR.lower(i), gp);
    }
    const Mat out_r = want_r ? XR.diag(i) : Mat();
    if (!out_r.zero()) {
      // straight into the output block, which the next step reads
      zgemm_in(st, S, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, B4, Xri, kOne, Xri, out_r, gp);
      gr[i] = out_r;
    } else {
      const Mat Z = W.get(spare, bi);
      zgemm_in(st, S, bi, bi, bi, QX_OP_N, QX_OP_N, kOne, B4, Xri, kOne, Xri, Z, gp);
      std::swap(spare, sr[i]);
      gr[i] = Z;
    }
    if (lds) ew_diag_imag(st, S, bi, gr[i], lds + orb_off[i], n_orb, kLdos);
  }
  });
  // released once the waits enqueued on it have been satisfied
  if (stg_ev) QX_HIP_CHECK(hipEventDestroy(stg_ev));
}

}  // namespace qx
