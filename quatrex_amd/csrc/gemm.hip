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
};

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
  const z2* __restrict__ A = g.A + e * g.sA;
  const z2* __restrict__ B = g.B + e * g.sB;
  z2* C = g.C + e * g.sC;
  const z2* Ci = g.Cin + e * g.sCin;

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
        if (gm < g.m && gk < g.k) ra[r] = A[(int64_t)gm * g.lda + gk];
      } else {  // A stored k x m, op = conj transpose
        const int kk = idx / BM, mm = idx % BM;
        const int gm = m0 + mm, gk = k0 + kk;
        ra[r] = zero;
        if (gm < g.m && gk < g.k) ra[r] = A[(int64_t)gk * g.lda + gm];
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int idx = tid + NT * r;
      if (OPB == QX_OP_N) {  // B is k x n row-major
        const int kk = idx / BN, nn = idx % BN;
        const int gn = n0 + nn, gk = k0 + kk;
        rb[r] = zero;
        if (gn < g.n && gk < g.k) rb[r] = B[(int64_t)gk * g.ldb + gn];
      } else {  // B stored n x k, op = conj transpose
        const int nn = idx / BK, kk = idx % BK;
        const int gn = n0 + nn, gk = k0 + kk;
        rb[r] = zero;
        if (gn < g.n && gk < g.k) rb[r] = B[(int64_t)gn * g.ldb + gk];
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
          double outr = g.ar * vr - g.ai * vi, outi = g.ar * vi + g.ai * vr;
          z2* dst = C + (int64_t)row * g.ldc + col;
          if (!g.beta_zero) {
            const z2 c = Ci[(int64_t)row * g.ldcin + col];
            outr += g.br * c.x - g.bi * c.y;
            outi += g.br * c.y + g.bi * c.x;
          }
          *dst = make_double2(outr, outi);
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

template <typename T, int SPLIT>
void launch_typed(hipStream_t s, int64_t batch, int opA, int opB, const GemmArgs& g) {
  // complex128 3M at 128 x 128 needs more than 256 registers per lane (it
  // would spill): that combination stays on the 64 x 64 shape.
  constexpr bool big_ok = !(sizeof(T) == 8 && SPLIT == 3);
  if constexpr (big_ok) {
    if (gemm_tile(g.m, g.n, batch) == 128) {
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
  zgemm_in(s, batch, m, n, k, opA, opB, alpha, A, B, beta, C, C, precision);
}

void zgemm_in(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
              Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& Cin,
              const Mat& C, int precision) {
  QX_REQUIRE(!C.zero(), "zgemm: output is a zero matrix");
  QX_REQUIRE(opA == QX_OP_N || opA == QX_OP_C, "zgemm: bad op_a");
  QX_REQUIRE(opB == QX_OP_N || opB == QX_OP_C, "zgemm: bad op_b");
  QX_REQUIRE(precision == QX_PREC_C128 || precision == QX_PREC_C64 ||
                 precision == QX_PREC_C128_3M,
             "zgemm: precision must be QX_PREC_C128, QX_PREC_C64 or QX_PREC_C128_3M");
  if (batch <= 0 || m <= 0 || n <= 0) return;
  const bool beta_zero = (beta.re == 0.0 && beta.im == 0.0) || Cin.zero();
  if (A.zero() || B.zero() || k <= 0) {
    // op(A) op(B) == 0: C = beta Cin.
    ew_axpby(s, batch, m, n, beta_zero ? Cplx{0, 0} : beta, Cin, Cplx{0, 0}, Mat(), C);
    return;
  }
  GemmArgs g;
  g.m = m;
  g.n = n;
  g.k = k;
  g.ar = alpha.re;
  g.ai = alpha.im;
  g.br = beta.re;
  g.bi = beta.im;
  g.A = A.p;
  g.sA = A.stride;
  g.lda = A.ld;
  g.B = B.p;
  g.sB = B.stride;
  g.ldb = B.ld;
  g.C = C.p;
  g.sC = C.stride;
  g.ldc = C.ld;
  g.Cin = Cin.zero() ? C.p : Cin.p;
  g.sCin = Cin.zero() ? C.stride : Cin.stride;
  g.ldcin = Cin.zero() ? C.ld : Cin.ld;
  g.beta_zero = beta_zero ? 1 : 0;
  const int split = gemm_split(precision);
  if (precision == QX_PREC_C64) {
    if (split == 3)
      launch_typed<float, 3>(s, batch, opA, opB, g);
    else
      launch_typed<float, 4>(s, batch, opA, opB, g);
  } else {
    if (split == 3)
      launch_typed<double, 3>(s, batch, opA, opB, g);
    else
      launch_typed<double, 4>(s, batch, opA, opB, g);
  }
}

}  // namespace qx
