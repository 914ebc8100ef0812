// Batched complex GEMM on CDNA4 matrix cores (gfx950).
//
// C = alpha * op(A) * op(B) + beta * C, op in {N, C (conjugate transpose)},
// for a strided stack of matrices.  This is the workhorse behind every `@`
// of the RGF / Sancho-Rubio / Beyn recursions.
//
// Design
// - One workgroup owns a BM x BN complex tile of one matrix of the stack and
//   is a WM x WN grid of waves; each wave owns a (BM/WM) x (BN/WN) sub-tile
//   made of 16 x 16 MFMA tiles, with one accumulator set per real product.
//   Two shapes are built: 64 x 64 with 2 x 2 waves (small matrices, many
//   workgroups) and 128 x 128 with 4 x 2 waves (half the L2 -> LDS traffic
//   per flop; used when the stack has enough 128 x 128 tiles to fill the
//   chip).  QX_ZGEMM_TILE=64|128 forces one of them.
// - The complex product runs on real MFMAs through one of two splits:
//     4M:  Cr += Ar Br - Ai Bi,  Ci += Ar Bi + Ai Br        (4 real MFMAs)
//     3M:  P1 += Ar Br,  P2 += Ai Bi,  P3 += (Ar + Ai)(Br + Bi),
//          Cr = P1 - P2,  Ci = P3 - P1 - P2                 (3 real MFMAs)
//   3M saves a quarter of the matrix-core work; its imaginary part carries
//   a cancellation, so its rounding error is a small constant factor above
//   4M.  QX_PREC_C128 uses 4M, QX_PREC_C128_3M and QX_PREC_C64 use 3M
//   (QX_ZGEMM_SPLIT=3|4 overrides all of them).
// - Operands are staged through LDS as real planes (re, im, and re + im for
//   3M), k-major, so that one MFMA operand fetch is 16 consecutive values.
//   The conjugate transpose is folded into the global->LDS copy.
// - The next K slice is prefetched into registers while the current one is
//   multiplied (single LDS buffer, two barriers per slice).
// - blockIdx is remapped so that the tiles of one matrix land on the same
//   XCD and share its L2.
//
// Precision: PREC_C128 runs v_mfma_f64_16x16x4_f64; PREC_C64 converts the
// complex128 operands to float on the LDS store and runs
// v_mfma_f32_16x16x4_f32, converting back on the store of C.
//
// MFMA 16x16x4 lane layout (checked on gfx950 hardware): lane l supplies
// A[row = l % 16][k = l / 16] and B[k = l / 16][col = l % 16].  Result
// element e of lane l is D[l / 16 + 4 e][l % 16] for the f64 instruction and
// D[4 (l / 16) + e][l % 16] for the f32 instruction.
#include <cstdlib>
#include <string>

#include "qx_internal.h"

namespace qx {
namespace {

// K slab per LDS stage and row padding (doubles); BK = 32 and PAD = 2 or 4
// measured no faster at 256^3.
constexpr int BK = 16, PAD = 1;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
  int m, n, k;
  double ar, ai, br, bi;
  const z2* A;
  int64_t sA;
  int lda;
  const z2* B;
  int64_t sB;
  int ldb;
  z2* C;
  int64_t sC;
  int ldc;
  const z2* Cin;  // beta term source (== C for the in-place form)
  int64_t sCin;
  int ldcin;
  int tilesM, tilesN;
  int beta_zero;
  z2* Ct;  // optional second output Ct = -C^H (n x m), or null
  int64_t sCt;
  int ldct;
  // optional: the beta term is Cin + E - E^H (m == n), E n x n
  const z2* E;
  int64_t sE;
  int lde;
  // Up to three products of one shape in one launch: matrices half <= e <
  // 2 half take every operand from the second set (base, matrix stride,
  // leading dimension) at index e - half, matrices e >= 2 half from the
  // third at e - 2 half.  Sets a launch does not use equal the first.
  int64_t half;
  const z2 *A2, *B2, *Cin2, *E2;
  z2 *C2, *Ct2;
  int64_t sA2, sB2, sC2, sCin2, sCt2, sE2;
  int lda2, ldb2, ldc2, ldcin2, ldct2, lde2;
  const z2 *A3, *B3, *Cin3, *E3;
  z2 *C3, *Ct3;
  int64_t sA3, sB3, sC3, sCin3, sCt3, sE3;
  int lda3, ldb3, ldc3, ldcin3, ldct3, lde3;
};

// Operands of matrix e of a launch (see GemmArgs::half); uniform over a
// workgroup.
struct GemmPtrs {
  const z2 *A, *B, *Cin, *E;
  z2 *C, *Ct;
  int lda, ldb, ldc, ldcin, ldct, lde;
};
__device__ __forceinline__ GemmPtrs gemm_ptrs(const GemmArgs& g, int64_t e) {
  const int grp = e >= 2 * g.half ? 2 : (e >= g.half ? 1 : 0);
  const int64_t l = e - grp * g.half;
  // field of the group's set
#define QX_SEL(f1, f2, f3) (grp == 2 ? (f3) : grp == 1 ? (f2) : (f1))
  GemmPtrs p;
  p.A = QX_SEL(g.A + l * g.sA, g.A2 + l * g.sA2, g.A3 + l * g.sA3);
  p.B = QX_SEL(g.B + l * g.sB, g.B2 + l * g.sB2, g.B3 + l * g.sB3);
  p.C = QX_SEL(g.C + l * g.sC, g.C2 + l * g.sC2, g.C3 + l * g.sC3);
  p.Cin = QX_SEL(g.Cin + l * g.sCin, g.Cin2 + l * g.sCin2, g.Cin3 + l * g.sCin3);
  p.E = g.E ? QX_SEL(g.E + l * g.sE, g.E2 + l * g.sE2, g.E3 + l * g.sE3) : nullptr;
  p.Ct = g.Ct ? QX_SEL(g.Ct + l * g.sCt, g.Ct2 + l * g.sCt2, g.Ct3 + l * g.sCt3) : nullptr;
  p.lda = QX_SEL(g.lda, g.lda2, g.lda3);
  p.ldb = QX_SEL(g.ldb, g.ldb2, g.ldb3);
  p.ldc = QX_SEL(g.ldc, g.ldc2, g.ldc3);
  p.ldcin = QX_SEL(g.ldcin, g.ldcin2, g.ldcin3);
  p.ldct = QX_SEL(g.ldct, g.ldct2, g.ldct3);
  p.lde = QX_SEL(g.lde, g.lde2, g.lde3);
#undef QX_SEL
  return p;
}

// beta-term source of element (row, col) of matrix e: Cin, or Cin + E - E^H
// with the additions in the order of the Hermitian-part kernel
// ((E - E^H) + Cin), so that the fused form rounds like the separate pass.
__device__ __forceinline__ z2 beta_source(const GemmPtrs& P, int row, int col) {
  const z2 c = P.Cin[(int64_t)row * P.ldcin + col];
  if (!P.E) return c;
  const z2 a = P.E[(int64_t)row * P.lde + col], b = P.E[(int64_t)col * P.lde + row];
  return make_double2((a.x - b.x) + c.x, (a.y + b.y) + c.y);
}

// Hardware dispatches workgroup ids round-robin over the 8 XCDs.  Return a
// logical id such that each XCD receives one contiguous logical range.
__device__ inline int xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx, rem = nblocks % nx;
  const int x = bid % nx, j = bid / nx;
  return x * per + min(x, rem) + j;
}

template <typename T>
struct MfmaTraits;

template <>
struct MfmaTraits<double> {
  typedef d4 acc_t;
  static __device__ inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of result element e of lane l inside a 16x16 tile
  static __device__ inline int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};

template <>
struct MfmaTraits<float> {
  typedef f4 acc_t;
  static __device__ inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

template <typename T, int OPA, int OPB, int SPLIT, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2)))
void zgemm_kernel(GemmArgs g) {
  typedef MfmaTraits<T> MT;
  typedef typename MT::acc_t acc_t;
  constexpr int NT = 64 * WM * WN;              // threads
  constexpr int TM = BM / (16 * WM), TN = BN / (16 * WN);  // MFMA tiles per wave
  constexpr int RA = BM * BK / NT, RB = BN * BK / NT;      // staged elements per thread
  static_assert(TM >= 1 && TN >= 1 && RA * NT == BM * BK && RB * NT == BN * BK, "tile shape");
  constexpr int NP = SPLIT == 3 ? 3 : 2;  // LDS planes per operand
  __shared__ T As[NP][BK][BM + PAD];
  __shared__ T Bs[NP][BK][BN + PAD];

  const int id = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles = g.tilesM * g.tilesN;
  const int64_t e = id / tiles;
  const int t = id - (int)(e * tiles);
  const int m0 = (t / g.tilesN) * BM, n0 = (t % g.tilesN) * BN;
  const GemmPtrs P = gemm_ptrs(g, e);
  const z2* __restrict__ A = P.A;
  const z2* __restrict__ B = P.B;
  z2* C = P.C;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w / WN, wn = w % WN;

  // 4M: acc1 = Cr, acc2 = Ci.  3M: acc1 = P1, acc2 = P2, acc3 = P3.
  acc_t acc1[TM][TN], acc2[TM][TN], acc3[SPLIT == 3 ? TM : 1][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc1[i][j] = acc_t{0, 0, 0, 0};
      acc2[i][j] = acc_t{0, 0, 0, 0};
      if constexpr (SPLIT == 3) acc3[i][j] = acc_t{0, 0, 0, 0};
    }

  z2 ra[RA], rb[RB];
  const z2 zero = make_double2(0.0, 0.0);

  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int idx = tid + NT * r;
      if (OPA == QX_OP_N) {  // A is m x k row-major
        const int row = idx / BK, col = idx % BK;
        const int gm = m0 + row, gk = k0 + col;
        ra[r] = zero;
        if (gm < g.m && gk < g.k) ra[r] = A[(int64_t)gm * P.lda + gk];
      } else {  // A stored k x m, op = conj transpose
        const int kk = idx / BM, mm = idx % BM;
        const int gm = m0 + mm, gk = k0 + kk;
        ra[r] = zero;
        if (gm < g.m && gk < g.k) ra[r] = A[(int64_t)gk * P.lda + gm];
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int idx = tid + NT * r;
      if (OPB == QX_OP_N) {  // B is k x n row-major
        const int kk = idx / BN, nn = idx % BN;
        const int gn = n0 + nn, gk = k0 + kk;
        rb[r] = zero;
        if (gn < g.n && gk < g.k) rb[r] = B[(int64_t)gk * P.ldb + gn];
      } else {  // B stored n x k, op = conj transpose
        const int nn = idx / BK, kk = idx % BK;
        const int gn = n0 + nn, gk = k0 + kk;
        rb[r] = zero;
        if (gn < g.n && gk < g.k) rb[r] = B[(int64_t)gn * P.ldb + gk];
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int idx = tid + NT * r;
      if (OPA == QX_OP_N) {
        const int row = idx / BK, col = idx % BK;
        const T re = (T)ra[r].x, im = (T)ra[r].y;
        As[0][col][row] = re;
        As[1][col][row] = im;
        if constexpr (SPLIT == 3) As[NP - 1][col][row] = re + im;
      } else {
        const int kk = idx / BM, mm = idx % BM;
        const T re = (T)ra[r].x, im = (T)(-ra[r].y);
        As[0][kk][mm] = re;
        As[1][kk][mm] = im;
        if constexpr (SPLIT == 3) As[NP - 1][kk][mm] = re + im;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int idx = tid + NT * r;
      if (OPB == QX_OP_N) {
        const int kk = idx / BN, nn = idx % BN;
        const T re = (T)rb[r].x, im = (T)rb[r].y;
        Bs[0][kk][nn] = re;
        Bs[1][kk][nn] = im;
        if constexpr (SPLIT == 3) Bs[NP - 1][kk][nn] = re + im;
      } else {
        const int nn = idx / BK, kk = idx % BK;
        const T re = (T)rb[r].x, im = (T)(-rb[r].y);
        Bs[0][kk][nn] = re;
        Bs[1][kk][nn] = im;
        if constexpr (SPLIT == 3) Bs[NP - 1][kk][nn] = re + im;
      }
    }
  };

  const int nk = (g.k + BK - 1) / BK;
  load(0);
  store();
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      T a_r[TM], a_i[TM], a_s[TM], b_r[TN], b_i[TN], b_s[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int mm = wm * TM * 16 + i * 16 + (lane & 15);
        a_r[i] = As[0][kr][mm];
        a_i[i] = As[1][kr][mm];
        if constexpr (SPLIT == 3) a_s[i] = As[NP - 1][kr][mm];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int nn = wn * TN * 16 + j * 16 + (lane & 15);
        b_r[j] = Bs[0][kr][nn];
        b_i[j] = Bs[1][kr][nn];
        if constexpr (SPLIT == 3) b_s[j] = Bs[NP - 1][kr][nn];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (SPLIT == 3) {
            acc1[i][j] = MT::mma(a_r[i], b_r[j], acc1[i][j]);
            acc2[i][j] = MT::mma(a_i[i], b_i[j], acc2[i][j]);
            acc3[i][j] = MT::mma(a_s[i], b_s[j], acc3[i][j]);
          } else {
            acc1[i][j] = MT::mma(a_r[i], b_r[j], acc1[i][j]);
            acc1[i][j] = MT::mma(-a_i[i], b_i[j], acc1[i][j]);
            acc2[i][j] = MT::mma(a_r[i], b_i[j], acc2[i][j]);
            acc2[i][j] = MT::mma(a_i[i], b_r[j], acc2[i][j]);
          }
        }
    }
    __syncthreads();
    if (kt + 1 < nk) {
      store();
      __syncthreads();
    }
  }

  // Epilogue: C = alpha * acc + beta * C.
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm * TM * 16 + i * 16 + MT::row(lane, q);
        const int col = n0 + wn * TN * 16 + j * 16 + (lane & 15);
        if (row < g.m && col < g.n) {
          double vr, vi;
          if constexpr (SPLIT == 3) {
            const double p1 = (double)acc1[i][j][q], p2 = (double)acc2[i][j][q];
            vr = p1 - p2;
            vi = (double)acc3[i][j][q] - p1 - p2;
          } else {
            vr = (double)acc1[i][j][q];
            vi = (double)acc2[i][j][q];
          }
          // explicit fma: the same rounding whatever the compiler contracts
          double outr = fma(g.ar, vr, -(g.ai * vi)), outi = fma(g.ar, vi, g.ai * vr);
          z2* dst = C + (int64_t)row * P.ldc + col;
          if (!g.beta_zero) {
            const z2 c = beta_source(P, row, col);
            outr += fma(g.br, c.x, -(g.bi * c.y));
            outi += fma(g.br, c.y, g.bi * c.x);
          }
          *dst = make_double2(outr, outi);
          if (P.Ct) P.Ct[(int64_t)col * P.ldct + row] = make_double2(-outr, outi);
        }
      }
}

// Register-direct complex128 kernel: no LDS and no barriers.
//
// Each wave loads its MFMA fragments straight from global memory (through
// L1/L2) into a ring of D register slots and multiplies k-step kk while the
// loads for k-steps kk+1 .. kk+D-1 are in flight (one k-step = 4 values of
// k).  The four waves of a workgroup (2 x 2) never wait for each other, so a
// SIMD's matrix core is not idled by workgroup barriers.  Wave tile is
// (16 TM) x (16 TN); workgroup tile (32 TM) x (32 TN).
//
// Fragment layout of v_mfma_f64_16x16x4_f64: lane l supplies
// op(A)[row l % 16][k l / 16] and op(B)[k l / 16][col l % 16].
// op = C (conjugate transpose) reads the stored element transposed and
// negates its imaginary part through the instruction's neg modifier (bits
// 0, 1, 2 of the last builtin operand negate the A, B, C sources).
//
// CHECK = false requires m % BM == 0, n % BN == 0 and k % (4 D) == 0.
// CHECK = true clamps every address into the matrix (rows of A past m and
// columns of B past n only feed outputs that are not stored); the pipelined
// loop then covers whole groups of D full k-steps and a short tail loop the
// rest, with both operands zeroed past k.  No value is selected on the
// pipelined path, so its loads are never waited for early.
//
// The arithmetic per output element (order of the k sum, 4M / 3M split, the
// epilogue's explicit fma) is the same as zgemm_kernel's, so both kernels
// give bit-identical results.
//
// T = float (complex64 arithmetic on complex128 storage): the fragments are
// loaded as complex128 and rounded to float right before their k-step, the
// conjugation is an explicit negation (the f32 instruction's last operand is
// a lane-broadcast pattern, not a neg modifier), and the 3M sums are formed
// in float, as zgemm_kernel<float> forms them on its LDS store; the result
// rows follow the f32 layout.
template <typename T, int OPA, int OPB, int SPLIT, int TM, int TN, int D, bool CHECK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void zgemm_reg_kernel(GemmArgs g) {
  constexpr bool F64 = sizeof(T) == 8;
  typedef typename MfmaTraits<T>::acc_t acc_t;
  constexpr int BM = 32 * TM, BN = 32 * TN;
  // neg-modifier masks: s_a = -1 when op(A) = C (imag part negated), s_b ditto
  constexpr bool CA = OPA == QX_OP_C, CB = OPB == QX_OP_C;
  const int id = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles = g.tilesM * g.tilesN;
  const int64_t e = id / tiles;
  const int t = id - (int)(e * tiles);
  const int m0 = (t / g.tilesN) * BM, n0 = (t % g.tilesN) * BN;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w >> 1, wn = w & 1;
  const GemmPtrs P = gemm_ptrs(g, e);
  const z2* __restrict__ A = P.A;
  const z2* __restrict__ B = P.B;
  const int r16 = lane & 15, k4 = lane >> 4;

  // pa[i] / pb[j]: this lane's element at k = 0 (k index k4 added below)
  const z2* pa[TM];
  const z2* pb[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    int row = m0 + wm * 16 * TM + i * 16 + r16;
    if (CHECK) row = min(row, g.m - 1);
    pa[i] = CA ? A + row : A + (int64_t)row * P.lda;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    int col = n0 + wn * 16 * TN + j * 16 + r16;
    if (CHECK) col = min(col, g.n - 1);
    pb[j] = CB ? B + (int64_t)col * P.ldb : B + col;
  }
  const int64_t ka = CA ? P.lda : 1, kb_ = CB ? 1 : P.ldb;  // element step per k
  const int nk = (g.k + 3) / 4;  // k-steps, the last one possibly partial
  // k-steps of the pipelined loop: whole groups of D full k-steps.  The rest
  // (CHECK only) runs in the tail loop below.
  const int nkm = CHECK ? (g.k / (4 * D)) * D : nk;

  // Loads of the pipelined loop.  Past its end (prefetch of k-steps it will
  // not use) the address is clamped to stay inside the matrix.
  auto fetch = [&](int kl, z2* fa, z2* fb) {
    const int kidx = min(4 * kl + k4, 4 * nkm - 4 + k4);
#pragma unroll
    for (int i = 0; i < TM; ++i) fa[i] = pa[i][kidx * ka];
#pragma unroll
    for (int j = 0; j < TN; ++j) fb[j] = pb[j][kidx * kb_];
  };

  acc_t acc1[TM][TN], acc2[TM][TN], acc3[SPLIT == 3 ? TM : 1][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc1[i][j] = acc_t{0, 0, 0, 0};
      acc2[i][j] = acc_t{0, 0, 0, 0};
      if constexpr (SPLIT == 3) acc3[i][j] = acc_t{0, 0, 0, 0};
    }

  // One k-step of MFMAs on fragments fa[TM], fb[TN].
  auto step = [&](const z2* fa, const z2* fb) {
    if constexpr (!F64) {
      typedef MfmaTraits<T> MT;
      T ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        ar[i] = (T)fa[i].x;
        ai[i] = (T)(CA ? -fa[i].y : fa[i].y);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        br[j] = (T)fb[j].x;
        bi[j] = (T)(CB ? -fb[j].y : fb[j].y);
      }
      if constexpr (SPLIT == 3) {
        T as[TM], bs[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) as[i] = ar[i] + ai[i];
#pragma unroll
        for (int j = 0; j < TN; ++j) bs[j] = br[j] + bi[j];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc1[i][j] = MT::mma(ar[i], br[j], acc1[i][j]);
            acc2[i][j] = MT::mma(ai[i], bi[j], acc2[i][j]);
            acc3[i][j] = MT::mma(as[i], bs[j], acc3[i][j]);
          }
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc1[i][j] = MT::mma(ar[i], br[j], acc1[i][j]);
            acc2[i][j] = MT::mma(ar[i], bi[j], acc2[i][j]);
          }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            acc1[i][j] = MT::mma(-ai[i], bi[j], acc1[i][j]);
            acc2[i][j] = MT::mma(ai[i], br[j], acc2[i][j]);
          }
      }
    } else if constexpr (SPLIT == 3) {
      double as[TM], bs[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) as[i] = CA ? fa[i].x - fa[i].y : fa[i].x + fa[i].y;
#pragma unroll
      for (int j = 0; j < TN; ++j) bs[j] = CB ? fb[j].x - fb[j].y : fb[j].x + fb[j].y;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].x, acc1[i][j], 0, 0, 0);
          // P2 += (s_a im_a)(s_b im_b)
          acc2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].y, acc2[i][j], 0, 0,
                                                            (CA != CB) ? 1 : 0);
          acc3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[i], bs[j], acc3[i][j], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].x, acc1[i][j], 0, 0, 0);
          // Ci += ar (s_b im_b)
          acc2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].y, acc2[i][j], 0, 0,
                                                            CB ? 2 : 0);
        }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          // Cr += -(s_a im_a)(s_b im_b)
          acc1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].y, acc1[i][j], 0, 0,
                                                            (CA == CB) ? 1 : 0);
          // Ci += (s_a im_a) br
          acc2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].x, acc2[i][j], 0, 0,
                                                            CA ? 1 : 0);
        }
    }
  };

  if (nkm > 0) {
    z2 fa[D][TM], fb[D][TN];
#pragma unroll
    for (int s = 0; s < D - 1; ++s) fetch(s, fa[s], fb[s]);
    // Opaque copies of the prologue values: without them the compiler merges
    // a slot's prologue load and its in-loop load into a single load at the
    // point of use, which removes the prefetch distance.
#pragma unroll
    for (int s = 0; s < D - 1; ++s) {
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        asm volatile("" : "+v"(fa[s][i].x));
        asm volatile("" : "+v"(fa[s][i].y));
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        asm volatile("" : "+v"(fb[s][j].x));
        asm volatile("" : "+v"(fb[s][j].y));
      }
    }
    for (int kb = 0; kb < nkm; kb += D) {
#pragma unroll
      for (int s = 0; s < D; ++s) {
        fetch(kb + s + D - 1, fa[(s + D - 1) % D], fb[(s + D - 1) % D]);
        step(fa[s], fb[s]);
      }
    }
  }
  if constexpr (CHECK) {
    // Tail: the remaining k-steps, operands zeroed past k.
    for (int kl = nkm; kl < nk; ++kl) {
      const int kraw = 4 * kl + k4;
      const bool ok = kraw < g.k;
      const int kidx = min(kraw, g.k - 1);
      z2 fa[TM], fb[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const z2 v = pa[i][kidx * ka];
        fa[i] = ok ? v : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const z2 v = pb[j][kidx * kb_];
        fb[j] = ok ? v : make_double2(0.0, 0.0);
      }
      step(fa, fb);
    }
  }

  z2* C = P.C;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm * 16 * TM + i * 16 + (F64 ? k4 + 4 * q : 4 * k4 + q);
        const int col = n0 + wn * 16 * TN + j * 16 + r16;
        if (!CHECK || (row < g.m && col < g.n)) {
          double vr, vi;
          if constexpr (SPLIT == 3) {
            const double p1 = (double)acc1[i][j][q], p2 = (double)acc2[i][j][q];
            vr = p1 - p2;
            vi = (double)acc3[i][j][q] - p1 - p2;
          } else {
            vr = (double)acc1[i][j][q];
            vi = (double)acc2[i][j][q];
          }
          // explicit fma, as in zgemm_kernel
          double outr = fma(g.ar, vr, -(g.ai * vi)), outi = fma(g.ar, vi, g.ai * vr);
          if (!g.beta_zero) {
            const z2 c = beta_source(P, row, col);
            outr += fma(g.br, c.x, -(g.bi * c.y));
            outi += fma(g.br, c.y, g.bi * c.x);
          }
          C[(int64_t)row * P.ldc + col] = make_double2(outr, outi);
          if (P.Ct) P.Ct[(int64_t)col * P.ldct + row] = make_double2(-outr, outi);
        }
      }
}

template <typename T, int SPLIT, int BM, int BN, int WM, int WN>
void launch_shape(hipStream_t s, int64_t batch, int opA, int opB, GemmArgs g) {
  g.tilesM = (g.m + BM - 1) / BM;
  g.tilesN = (g.n + BN - 1) / BN;
  const int64_t blocks = batch * g.tilesM * g.tilesN;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "zgemm: grid too large");
  dim3 grid((unsigned)blocks), block(64 * WM * WN);
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);
  if (opA == QX_OP_N && opB == QX_OP_N)
    hipLaunchKernelGGL((zgemm_kernel<T, QX_OP_N, QX_OP_N, SPLIT, BM, BN, WM, WN>), grid, block, 0, s, g);
  else if (opA == QX_OP_N && opB == QX_OP_C)
    hipLaunchKernelGGL((zgemm_kernel<T, QX_OP_N, QX_OP_C, SPLIT, BM, BN, WM, WN>), grid, block, 0, s, g);
  else if (opA == QX_OP_C && opB == QX_OP_N)
    hipLaunchKernelGGL((zgemm_kernel<T, QX_OP_C, QX_OP_N, SPLIT, BM, BN, WM, WN>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((zgemm_kernel<T, QX_OP_C, QX_OP_C, SPLIT, BM, BN, WM, WN>), grid, block, 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::GEMM, 8.0 * g.m * g.n * (double)g.k * batch);
}

// Tile shape: QX_ZGEMM_TILE if set, else the calling handle's setting
// (qx_set_tuning, e.g. from an autotuner), else 128 x 128 when both
// dimensions fill it and the stack has at least one such tile per CU, else
// 64 x 64.
int gemm_tile(int m, int n, int64_t batch) {
  static const int forced = []() {
    const char* e = std::getenv("QX_ZGEMM_TILE");
    if (!e) return 0;
    const int v = std::atoi(e);
    return (v == 64 || v == 128) ? v : 0;
  }();
  if (forced) return forced;
  if (active_gemm_tile == 64 || active_gemm_tile == 128) return active_gemm_tile;
  const int64_t big = batch * ((m + 127) / 128) * ((n + 127) / 128);
  return (m >= 128 && n >= 128 && big >= 256) ? 128 : 64;
}

template <typename T, int SPLIT, int TM, int TN, int D>
void launch_reg(hipStream_t s, int64_t batch, int opA, int opB, GemmArgs g) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  g.tilesM = (g.m + BM - 1) / BM;
  g.tilesN = (g.n + BN - 1) / BN;
  const int64_t blocks = batch * g.tilesM * g.tilesN;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "zgemm: grid too large");
  const bool full = g.m % BM == 0 && g.n % BN == 0 && g.k % (4 * D) == 0;
  dim3 grid((unsigned)blocks), block(256);
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);
#define QX_REG_LAUNCH(OA, OB)                                                                   \
  if (full)                                                                                     \
    hipLaunchKernelGGL((zgemm_reg_kernel<T, OA, OB, SPLIT, TM, TN, D, false>), grid, block, 0, s, g); \
  else                                                                                          \
    hipLaunchKernelGGL((zgemm_reg_kernel<T, OA, OB, SPLIT, TM, TN, D, true>), grid, block, 0, s, g);
  if (opA == QX_OP_N && opB == QX_OP_N) {
    QX_REG_LAUNCH(QX_OP_N, QX_OP_N)
  } else if (opA == QX_OP_N && opB == QX_OP_C) {
    QX_REG_LAUNCH(QX_OP_N, QX_OP_C)
  } else if (opA == QX_OP_C && opB == QX_OP_N) {
    QX_REG_LAUNCH(QX_OP_C, QX_OP_N)
  } else {
    QX_REG_LAUNCH(QX_OP_C, QX_OP_C)
  }
#undef QX_REG_LAUNCH
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::GEMM, 8.0 * g.m * g.n * (double)g.k * batch);
}

// Shapes of the register-direct kernel, as (TM, TN, D): workgroup tile
// (32 TM) x (32 TN), D register slots.  3M keeps three accumulator sets and
// only fits 2 x 2 without spilling.
enum RegShape { REG_2x2x4 = 0, REG_2x4x2, REG_4x2x2, REG_NSHAPES };

int env_int(const char* name, int dflt) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : dflt;
}

// Register-direct shape for a complex128 launch, or -1 for the LDS-staged
// kernel.  Precedence: QX_ZGEMM_TILE (64 | 128: LDS shape), QX_ZGEMM_KERNEL=
// lds, QX_ZGEMM_REG (shape index), then the calling handle's gemm_tile
// (qx_set_tuning: 64 | 128 LDS, QX_GEMM_REG_* register-direct), then the
// automatic choice.
int reg_shape(int split, int opA, int opB, int m, int n) {
  static const int env_tile = env_int("QX_ZGEMM_TILE", 0);
  static const bool env_lds = [] {
    const char* e = std::getenv("QX_ZGEMM_KERNEL");
    return e && std::string(e) == "lds";
  }();
  static const int env_reg = env_int("QX_ZGEMM_REG", -1);
  if (env_tile == 64 || env_tile == 128 || env_lds) return -1;
  int shape;
  if (env_reg >= 0 && env_reg < REG_NSHAPES) {
    shape = env_reg;
  } else {
    switch (active_gemm_tile) {
      case QX_GEMM_LDS_64: case QX_GEMM_LDS_128: return -1;
      case QX_GEMM_REG_64x64: shape = REG_2x2x4; break;
      case QX_GEMM_REG_64x128: shape = REG_2x4x2; break;
      case QX_GEMM_REG_128x64: shape = REG_4x2x2; break;
      default: {
        // Automatic (measured at 32-256 matrices of 64-512): A op(B) with
        // op(B) = B^H reads both operands with a 64-byte stride pattern;
        // the 128 x 64 tile halves the fragment loads per MFMA there.  The
        // 3M split cannot use the larger tiles (three accumulator sets) and
        // is faster through LDS for that case.
        const bool nc = opA == QX_OP_N && opB == QX_OP_C;
        if (nc && split == 3) return -1;
        shape = (nc && m >= 256 && n >= 256) ? REG_4x2x2 : REG_2x2x4;
      }
    }
  }
  return split == 3 ? REG_2x2x4 : shape;
}

template <typename T, int SPLIT>
void launch_reg_shape(hipStream_t s, int64_t batch, int opA, int opB, const GemmArgs& g, int shape) {
  if constexpr (SPLIT == 4) {
    switch (shape) {
      case REG_2x4x2: launch_reg<T, SPLIT, 2, 4, 2>(s, batch, opA, opB, g); return;
      case REG_4x2x2: launch_reg<T, SPLIT, 4, 2, 2>(s, batch, opA, opB, g); return;
      default: break;
    }
  }
  launch_reg<T, SPLIT, 2, 2, 4>(s, batch, opA, opB, g);
}

// complex64 products on the register-direct kernel only with QX_ZGEMM_F32=reg
// (read per call).  Same bits as the LDS-staged kernel, but slower: each
// fragment is still 16 bytes of complex128 storage per lane and per wave
// while the f32 matrix core runs twice as fast, so the register-direct
// loads bind (32 x 256^3, A.B: 54.0 vs 45.7 us; mixed RGF at 256 energies:
// 405 vs 458 energies/s); the LDS kernel loads each element once per
// workgroup and keeps it as float.
bool reg_f32() {
  const char* e = std::getenv("QX_ZGEMM_F32");
  return e && std::string(e) == "reg";
}

template <typename T, int SPLIT>
void launch_typed(hipStream_t s, int64_t batch, int opA, int opB, const GemmArgs& g) {
  if (sizeof(T) == 8 || reg_f32()) {
    const int shape = reg_shape(SPLIT, opA, opB, g.m, g.n);
    if (shape >= 0) {
      launch_reg_shape<T, SPLIT>(s, batch, opA, opB, g, shape);
      return;
    }
  }
  // complex128 3M at 128 x 128 needs more than 256 registers per lane (it
  // would spill): that combination stays on the 64 x 64 shape.
  constexpr bool big_ok = !(sizeof(T) == 8 && SPLIT == 3);
  if constexpr (big_ok) {
    // the tile of one product alone, also in a paired launch
    if (gemm_tile(g.m, g.n, g.half) == 128) {
      launch_shape<T, SPLIT, 128, 128, 4, 2>(s, batch, opA, opB, g);
      return;
    }
  }
  launch_shape<T, SPLIT, 64, 64, 2, 2>(s, batch, opA, opB, g);
}

// Complex split of the GEMM: 4M for QX_PREC_C128, 3M for QX_PREC_C128_3M
// and QX_PREC_C64, unless QX_ZGEMM_SPLIT (3 or 4) is set.
int gemm_split(int precision) {
  static const int forced = []() {
    const char* e = std::getenv("QX_ZGEMM_SPLIT");
    if (!e) return 0;
    const int v = std::atoi(e);
    return (v == 3 || v == 4) ? v : 0;
  }();
  if (forced) return forced;
  return precision == QX_PREC_C128 ? 4 : 3;
}

}  // namespace

void zgemm(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
           Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& C,
           int precision) {
  zgemm_in(s, batch, m, n, k, opA, opB, alpha, A, B, beta, C, C, precision, Mat());
}

namespace {

void check_operands(int m, int n, const GemmOperands& o) {
  QX_REQUIRE(!o.C.zero(), "zgemm: output is a zero matrix");
  QX_REQUIRE(o.E.zero() || (m == n && !o.Cin.zero() && o.E.p != o.C.p && o.Cin.p != o.C.p),
             "zgemm: the D + E - E^H beta term needs m == n, a non-zero D and an output "
             "aliasing neither");
  QX_REQUIRE(o.Ct.zero() || (o.Ct.p != o.C.p && o.Ct.p != o.A.p && o.Ct.p != o.B.p &&
                             o.Ct.p != o.Cin.p),
             "zgemm: the transposed output aliases an operand");
}

// Operands of two products can share one launch when both or neither are
// present (strides and leading dimensions may differ).
bool same_layout(const Mat& a, const Mat& b) { return a.zero() == b.zero(); }

// An output of one product is no operand of the other.
bool independent(const GemmOperands& x, const GemmOperands& y) {
  for (const z2* w : {x.C.p, x.Ct.p}) {
    if (!w) continue;
    for (const z2* r : {y.A.p, y.B.p, y.Cin.p, y.E.p, y.C.p, y.Ct.p})
      if (r == w) return false;
  }
  return true;
}

// One launch of batch products (y, z null) or of two or three such groups
// of products (x, y[, z]: 2 or 3 batch matrices).
void zgemm_impl(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                Cplx alpha, Cplx beta, const GemmOperands& x, const GemmOperands* y,
                int precision, const GemmOperands* z = nullptr) {
  QX_REQUIRE(!z || y, "zgemm: a third group needs a second");
  const GemmOperands* gs[3] = {&x, y, z};
  const int ng = z ? 3 : (y ? 2 : 1);
  for (int a = 0; a < ng; ++a) check_operands(m, n, *gs[a]);
  QX_REQUIRE(opA == QX_OP_N || opA == QX_OP_C, "zgemm: bad op_a");
  QX_REQUIRE(opB == QX_OP_N || opB == QX_OP_C, "zgemm: bad op_b");
  QX_REQUIRE(precision == QX_PREC_C128 || precision == QX_PREC_C64 ||
                 precision == QX_PREC_C128_3M || oz2_moduli(precision) > 0 ||
                 ozbf_slices(precision) > 0,
             "zgemm: precision must be QX_PREC_C128, QX_PREC_C64, QX_PREC_C128_3M, "
             "QX_PREC_OZAKI2(8..20) or QX_PREC_OZAKI_BF16(2..7)");
  if (batch <= 0 || m <= 0 || n <= 0) return;
  if (const int nmod = oz2_moduli(precision)) {
    // one launch sequence per product, in order, on the calling thread's
    // residue workspace
    for (int a = 0; a < ng; ++a)
      zgemm_oz2(s, batch, m, n, k, opA, opB, alpha, beta, *gs[a], nmod, active_oz_ws.p,
                active_oz_ws.bytes);
    return;
  }
  if (const int sl = ozbf_slices(precision)) {
    // bf16 digit products; the D + E - E^H beta term is formed in C first
    // (C aliases neither D nor E) and -C^H written after
    for (int a = 0; a < ng; ++a) {
      const GemmOperands& o = *gs[a];
      const bool bz = (beta.re == 0.0 && beta.im == 0.0) || o.Cin.zero();
      Mat cin = o.Cin;
      if (!o.E.zero() && !bz) {
        ew_herm_add(s, batch, n, o.Cin, 1.0, o.E, o.C);
        cin = o.C;
      }
      if (o.A.zero() || o.B.zero() || k <= 0)
        ew_axpby(s, batch, m, n, bz ? Cplx{0, 0} : beta, bz ? Mat() : cin, Cplx{0, 0}, Mat(),
                 o.C);
      else
        zgemm_ozaki(s, batch, m, n, k, opA, opB, alpha, o.A, o.B, beta, bz ? Mat() : cin, o.C,
                    sl, active_oz_ws.p, active_oz_ws.bytes);
      if (!o.Ct.zero()) ew_neg_h(s, batch, m, n, o.C, o.Ct);
    }
    return;
  }
  if (ng > 1) {
    // otherwise x, y[, z] as separate launches in that order
    bool one_launch = !x.A.zero() && !x.B.zero() && k > 0;
    for (int a = 0; a < ng; ++a)
      for (int b = 0; b < ng; ++b)
        if (a != b) one_launch = one_launch && independent(*gs[a], *gs[b]);
    for (int a = 1; a < ng; ++a) {
      const GemmOperands& o = *gs[a];
      one_launch = one_launch && same_layout(x.A, o.A) && same_layout(x.B, o.B) &&
                   same_layout(x.Cin, o.Cin) && same_layout(x.C, o.C) &&
                   same_layout(x.Ct, o.Ct) && same_layout(x.E, o.E);
    }
    if (!one_launch) {
      for (int a = 0; a < ng; ++a)
        zgemm_impl(s, batch, m, n, k, opA, opB, alpha, beta, *gs[a], nullptr, precision);
      return;
    }
  }
  const bool beta_zero = (beta.re == 0.0 && beta.im == 0.0) || x.Cin.zero();
  if (x.A.zero() || x.B.zero() || k <= 0) {
    // op(A) op(B) == 0: C = beta Cin (Cin + E - E^H).
    if (!x.E.zero() && !beta_zero) {
      ew_herm_add(s, batch, n, x.Cin, 1.0, x.E, x.C);
      if (beta.re != 1.0 || beta.im != 0.0)
        ew_axpby(s, batch, m, n, beta, x.C, Cplx{0, 0}, Mat(), x.C);
    } else {
      ew_axpby(s, batch, m, n, beta_zero ? Cplx{0, 0} : beta, x.Cin, Cplx{0, 0}, Mat(), x.C);
    }
    if (!x.Ct.zero()) ew_neg_h(s, batch, m, n, x.C, x.Ct);
    return;
  }
  const GemmOperands& o2 = y ? *y : x;
  GemmArgs g;
  g.m = m;
  g.n = n;
  g.k = k;
  g.ar = alpha.re;
  g.ai = alpha.im;
  g.br = beta.re;
  g.bi = beta.im;
  g.A = x.A.p;
  g.sA = x.A.stride;
  g.lda = x.A.ld;
  g.B = x.B.p;
  g.sB = x.B.stride;
  g.ldb = x.B.ld;
  g.C = x.C.p;
  g.sC = x.C.stride;
  g.ldc = x.C.ld;
  const Mat& ci = x.Cin.zero() ? x.C : x.Cin;
  g.Cin = ci.p;
  g.sCin = ci.stride;
  g.ldcin = ci.ld;
  g.beta_zero = beta_zero ? 1 : 0;
  g.Ct = x.Ct.p;
  g.sCt = x.Ct.stride;
  g.ldct = x.Ct.ld;
  g.E = x.E.p;
  g.sE = x.E.stride;
  g.lde = x.E.ld;
  g.half = batch;
  g.A2 = o2.A.p;
  g.sA2 = o2.A.stride;
  g.lda2 = o2.A.ld;
  g.B2 = o2.B.p;
  g.sB2 = o2.B.stride;
  g.ldb2 = o2.B.ld;
  g.C2 = o2.C.p;
  g.sC2 = o2.C.stride;
  g.ldc2 = o2.C.ld;
  const Mat& ci2 = o2.Cin.zero() ? o2.C : o2.Cin;
  g.Cin2 = ci2.p;
  g.sCin2 = ci2.stride;
  g.ldcin2 = ci2.ld;
  g.Ct2 = o2.Ct.p;
  g.sCt2 = o2.Ct.stride;
  g.ldct2 = o2.Ct.ld;
  g.E2 = o2.E.p;
  g.sE2 = o2.E.stride;
  g.lde2 = o2.E.ld;
  const GemmOperands& o3 = z ? *z : x;
  g.A3 = o3.A.p;
  g.sA3 = o3.A.stride;
  g.lda3 = o3.A.ld;
  g.B3 = o3.B.p;
  g.sB3 = o3.B.stride;
  g.ldb3 = o3.B.ld;
  g.C3 = o3.C.p;
  g.sC3 = o3.C.stride;
  g.ldc3 = o3.C.ld;
  const Mat& ci3 = o3.Cin.zero() ? o3.C : o3.Cin;
  g.Cin3 = ci3.p;
  g.sCin3 = ci3.stride;
  g.ldcin3 = ci3.ld;
  g.Ct3 = o3.Ct.p;
  g.sCt3 = o3.Ct.stride;
  g.ldct3 = o3.Ct.ld;
  g.E3 = o3.E.p;
  g.sE3 = o3.E.stride;
  g.lde3 = o3.E.ld;
  const int64_t nb = ng * batch;
  const int split = gemm_split(precision);
  if (precision == QX_PREC_C64) {
    if (split == 3)
      launch_typed<float, 3>(s, nb, opA, opB, g);
    else
      launch_typed<float, 4>(s, nb, opA, opB, g);
  } else {
    if (split == 3)
      launch_typed<double, 3>(s, nb, opA, opB, g);
    else
      launch_typed<double, 4>(s, nb, opA, opB, g);
  }
}

}  // namespace

void zgemm_in(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
              Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& Cin,
              const Mat& C, int precision, const Mat& Ct) {
  zgemm_impl(s, batch, m, n, k, opA, opB, alpha, beta, GemmOperands{A, B, Cin, C, Ct, Mat()},
             nullptr, precision);
}

void zgemm_hin(hipStream_t s, int64_t batch, int n, int k, int opA, int opB, Cplx alpha,
               const Mat& A, const Mat& B, Cplx beta, const Mat& D, const Mat& E, const Mat& C,
               int precision) {
  zgemm_impl(s, batch, n, n, k, opA, opB, alpha, beta, GemmOperands{A, B, D, C, Mat(), E},
             nullptr, precision);
}

void zgemm_pair(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                Cplx alpha, Cplx beta, const GemmOperands& x, const GemmOperands& y,
                int precision) {
  zgemm_impl(s, batch, m, n, k, opA, opB, alpha, beta, x, &y, precision);
}

void zgemm_triple(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                  Cplx alpha, Cplx beta, const GemmOperands& x, const GemmOperands& y,
                  const GemmOperands& z, int precision) {
  zgemm_impl(s, batch, m, n, k, opA, opB, alpha, beta, x, &y, precision, &z);
}

}  // namespace qx
