// Ozaki-style split complex GEMM on bf16 matrix cores (gfx950).
//
// C = alpha op(A) op(B) + beta Cin for complex128 data, computed from
// integer "digit" slices of the operands on v_mfma_f32_16x16x32_bf16.
//
// Splitting (ozaki_split_kernel): every row of op(A) (column of op(B)) is
// scaled by a power of two 2^e with max(|re|, |im|) < 2^e over the row, so
// that its entries t lie in (-1, 1), and written as s digits
//   t = sum_p d_p 2^(-7 p) + r,   d_p integer, |d_p| <= 127,
// by repeated t <- 2^7 t, d = trunc(t), t <- t - d.  A 7-bit integer is
// exact in bf16.  Real and imaginary parts are split separately with the
// row's common exponent.
//
// Products (ozaki_gemm_kernel): for digit pairs (p, q) with p + q <= s + 1
// the products A_p B_q are formed on the bf16 MFMA with fp32 accumulation.
// Every product of two digits is an integer below 2^14 in magnitude; one
// accumulator per t = p + q collects, over 64 values of k, at most
// 2 s * 64 * 127^2 < 2^24 in magnitude for s <= 8 (7 are built), so the fp32 sums are
// exact integers.  Every 64 values of k they are added, scaled by 2^(-7 t),
// into FP64 accumulators.  Each operand keeps a remainder below 2^(-7 s) of
// its line's largest entry and the pairs p + q <= s + 1 keep every product
// term of that order, so the relative error is a small multiple of
// 2^(-7 s): about 6e-6 for s = 3, 6e-10 for s = 5, 5e-14 for s = 7.
//
// Complex product by the 4M split: Cr = Ar Br - Ai Bi, Ci = Ar Bi + Ai Br,
// the minus sign applied by flipping the bf16 sign bits of the Bi digits.
//
// Layouts: digits of op(A) as [batch][p][re/im][row][k_pad] and of op(B) as
// [batch][p][re/im][col][k_pad], k contiguous and zero-padded to a multiple
// of 32, so one MFMA operand (8 consecutive k of one row or column) is one
// 16-byte load.  Exponents as int [batch][row] / [batch][col].
//
// bf16 MFMA 16x16x32 lane layout (checked on gfx950 hardware): lane l passes
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7; result
// element e of lane l is D[4 (l / 16) + e][l % 16].
#include <cstdint>

#include "qx_internal.h"

namespace qx {
namespace {

constexpr int kDigitBits = 7;
constexpr int kMaxSlices = 7;
constexpr int OBM = 64, OBN = 64, OBK = 32, ONT = 256;

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// One workgroup per line (row of op(A) or column of op(B)) of one matrix:
// element kk of line i is X[i][kk] (trans == 0) or X[kk][i] (trans == 1) of
// the stored matrix (leading dimension ld), conjugated when conj == 1.
__global__ __launch_bounds__(ONT) void ozaki_split_kernel(const z2* X, int64_t sX, int ld,
                                                         int trans, int conj, int lines, int k,
                                                         int k_pad, int slices,
                                                         unsigned short* digits,
                                                         int64_t s_digits, int* expo,
                                                         int64_t s_expo) {
  __shared__ double red[ONT / 64];
  const int64_t e = blockIdx.x / lines;
  const int line = (int)(blockIdx.x - e * lines);
  const z2* Xe = X + e * sX;
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int kk = tid; kk < k; kk += ONT) {
    const z2 v = trans ? Xe[(int64_t)kk * ld + line] : Xe[(int64_t)line * ld + kk];
    mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = red[0];
#pragma unroll
  for (int q = 1; q < ONT / 64; ++q) mx = fmax(mx, red[q]);
  // smallest e with mx < 2^e (e = 0 for an all-zero line)
  int ex = 0;
  if (mx > 0.0) {
    ex = ilogb(mx) + 1;
  }
  if (tid == 0) expo[e * s_expo + line] = ex;
  const int64_t plane = (int64_t)lines * k_pad;  // one (p, re/im) plane
  unsigned short* D = digits + e * s_digits + (int64_t)line * k_pad;
  for (int kk = tid; kk < k_pad; kk += ONT) {
    double tr = 0.0, ti = 0.0;
    if (kk < k) {
      const z2 v = trans ? Xe[(int64_t)kk * ld + line] : Xe[(int64_t)line * ld + kk];
      tr = ldexp(v.x, -ex);
      ti = ldexp(conj ? -v.y : v.y, -ex);
    }
    for (int p = 0; p < slices; ++p) {
      tr = ldexp(tr, kDigitBits);
      ti = ldexp(ti, kDigitBits);
      const double dr = trunc(tr), di = trunc(ti);
      tr -= dr;
      ti -= di;
      const __bf16 br = (__bf16)(float)dr, bi = (__bf16)(float)di;
      D[(2 * p) * plane + kk] = __builtin_bit_cast(unsigned short, br);
      D[(2 * p + 1) * plane + kk] = __builtin_bit_cast(unsigned short, bi);
    }
  }
}

struct OzArgs {
  int m, n, k_pad, slices;
  double ar, ai, br, bi;
  const unsigned short* Ad;  // [batch][p][2][m][k_pad]
  int64_t sAd;
  const unsigned short* Bd;  // [batch][p][2][n][k_pad]
  int64_t sBd;
  const int* ea;  // [batch][m]
  const int* eb;  // [batch][n]
  z2* C;
  int64_t sC;
  int ldc;
  const z2* Cin;
  int64_t sCin;
  int ldcin;
  int tilesM, tilesN;
  int beta_zero;
};

__device__ inline int oz_xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx, rem = nblocks % nx;
  const int x = bid % nx, j = bid / nx;
  return x * per + min(x, rem) + j;
}

// 64 x 64 output tile per workgroup of 4 waves (2 x 2, 32 x 32 each).
template <int S>
__global__ __launch_bounds__(ONT) void ozaki_gemm_kernel(OzArgs g) {
  // LDS: digits of the current 32-wide k slab, [p][re/im][row][8 x 16-byte
  // groups], row padded by one group against bank conflicts.
  constexpr int GP = OBK / 8 + 1;  // 16-byte groups per row (+1 pad)
  __shared__ us8 As[S][2][OBM][GP];
  __shared__ us8 Bs[S][2][OBN][GP];

  const int id = oz_xcd_remap(blockIdx.x, gridDim.x);
  const int tiles = g.tilesM * g.tilesN;
  const int64_t e = id / tiles;
  const int t = id - (int)(e * tiles);
  const int m0 = (t / g.tilesN) * OBM, n0 = (t % g.tilesN) * OBN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int64_t planeA = (int64_t)g.m * g.k_pad, planeB = (int64_t)g.n * g.k_pad;
  const us8* Ad = reinterpret_cast<const us8*>(g.Ad + e * g.sAd);
  const us8* Bd = reinterpret_cast<const us8*>(g.Bd + e * g.sBd);
  const int kg = g.k_pad / 8;  // 16-byte groups per line

  f4 acc_r[S][2][2], acc_i[S][2][2];  // index t - 2 = p + q - 2 (0-based p, q: p + q)
  double c64r[2][2][4], c64i[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) c64r[i][j][q] = c64i[i][j][q] = 0.0;

  const int nk = g.k_pad / OBK;
  for (int kt = 0; kt < nk; ++kt) {
    if ((kt & 1) == 0) {
#pragma unroll
      for (int s = 0; s < S; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc_r[s][i][j] = f4{0, 0, 0, 0};
            acc_i[s][i][j] = f4{0, 0, 0, 0};
          }
    }
    // global -> LDS: S * 2 planes * 64 lines * 4 groups for A and for B
    const us8 zero = us8{0, 0, 0, 0, 0, 0, 0, 0};
    for (int idx = tid; idx < S * 2 * OBM * 4; idx += ONT) {
      const int grp = idx & 3, line = (idx >> 2) & (OBM - 1), pc = idx >> 8;
      const int gm = m0 + line;
      us8 v = zero;
      if (gm < g.m) v = Ad[pc * (planeA / 8) + (int64_t)gm * kg + kt * 4 + grp];
      As[pc >> 1][pc & 1][line][grp] = v;
      const int gn = n0 + line;
      us8 u = zero;
      if (gn < g.n) u = Bd[pc * (planeB / 8) + (int64_t)gn * kg + kt * 4 + grp];
      Bs[pc >> 1][pc & 1][line][grp] = u;
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < S; ++p) {
      bf8 ar[2], ai[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wm * 32 + i * 16 + (lane & 15);
        ar[i] = __builtin_bit_cast(bf8, As[p][0][row][lane >> 4]);
        ai[i] = __builtin_bit_cast(bf8, As[p][1][row][lane >> 4]);
      }
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (p + q > S - 1) continue;  // keep p + q <= s + 1 (1-based)
        bf8 br[2], bi[2], bin[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int col = wn * 32 + j * 16 + (lane & 15);
          br[j] = __builtin_bit_cast(bf8, Bs[q][0][col][lane >> 4]);
          const us8 raw = Bs[q][1][col][lane >> 4];
          bi[j] = __builtin_bit_cast(bf8, raw);
          bin[j] = __builtin_bit_cast(bf8, raw ^ (unsigned short)0x8000u);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            f4 r = acc_r[p + q][i][j], im = acc_i[p + q][i][j];
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i], br[j], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[i], bin[j], r, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i], bi[j], im, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[i], br[j], im, 0, 0, 0);
            acc_r[p + q][i][j] = r;
            acc_i[p + q][i][j] = im;
          }
      }
    }
    __syncthreads();
    if ((kt & 1) == 1 || kt == nk - 1) {
      // exact integer partial sums -> FP64, scaled by 2^(-7 t), t = p + q + 2
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double sc = ldexp(1.0, -kDigitBits * (s + 2));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              c64r[i][j][q] += sc * (double)acc_r[s][i][j][q];
              c64i[i][j][q] += sc * (double)acc_i[s][i][j][q];
            }
      }
    }
  }

  const int* ea = g.ea + e * g.m;
  const int* eb = g.eb + e * g.n;
  z2* C = g.C + e * g.sC;
  const z2* Ci = g.Cin + e * g.sCin;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < g.m && col < g.n) {
          const int ex = ea[row] + eb[col];
          const double vr = ldexp(c64r[i][j][q], ex), vi = ldexp(c64i[i][j][q], ex);
          double outr = g.ar * vr - g.ai * vi, outi = g.ar * vi + g.ai * vr;
          if (!g.beta_zero) {
            const z2 c = Ci[(int64_t)row * g.ldcin + col];
            outr += g.br * c.x - g.bi * c.y;
            outi += g.br * c.y + g.bi * c.x;
          }
          C[(int64_t)row * g.ldc + col] = make_double2(outr, outi);
        }
      }
}

}  // namespace

size_t ozaki_workspace_bytes(int64_t batch, int m, int n, int k, int slices) {
  const int64_t k_pad = (k + OBK - 1) / OBK * OBK;
  const int64_t digits = batch * (int64_t)slices * 2 * (m + n) * k_pad;
  return (size_t)digits * 2 + (size_t)batch * (m + n) * sizeof(int) + 256;
}

void zgemm_ozaki(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB,
                 Cplx alpha, const Mat& A, const Mat& B, Cplx beta, const Mat& Cin,
                 const Mat& C, int slices, void* ws, size_t ws_bytes) {
  QX_REQUIRE(!C.zero() && !A.zero() && !B.zero(), "zgemm_ozaki: null operand");
  QX_REQUIRE(opA == QX_OP_N || opA == QX_OP_C, "zgemm_ozaki: bad op_a");
  QX_REQUIRE(opB == QX_OP_N || opB == QX_OP_C, "zgemm_ozaki: bad op_b");
  QX_REQUIRE(slices >= 2 && slices <= kMaxSlices, "zgemm_ozaki: slices must be in [2, 7]");
  QX_REQUIRE(k >= 1 && k <= 4096, "zgemm_ozaki: k must be in [1, 4096]");
  if (batch <= 0 || m <= 0 || n <= 0) return;
  QX_REQUIRE(ws_bytes >= ozaki_workspace_bytes(batch, m, n, k, slices),
             "zgemm_ozaki: workspace too small");
  const int k_pad = (k + OBK - 1) / OBK * OBK;
  unsigned short* dA = static_cast<unsigned short*>(ws);
  const int64_t sAd = (int64_t)slices * 2 * m * k_pad;
  const int64_t sBd = (int64_t)slices * 2 * n * k_pad;
  unsigned short* dB = dA + batch * sAd;
  int* ea = reinterpret_cast<int*>(dB + batch * sBd);
  int* eb = ea + batch * m;
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);
  // line i of op(A) = row i of A, or conj of column i of A for A^H;
  // line j of op(B) = column j of B, or conj of row j of B for B^H
  hipLaunchKernelGGL(ozaki_split_kernel, dim3((unsigned)(batch * m)), dim3(ONT), 0, s, A.p,
                     A.stride, A.ld, opA == QX_OP_C ? 1 : 0, opA == QX_OP_C ? 1 : 0, m, k,
                     k_pad, slices, dA, sAd, ea, (int64_t)m);
  hipLaunchKernelGGL(ozaki_split_kernel, dim3((unsigned)(batch * n)), dim3(ONT), 0, s, B.p,
                     B.stride, B.ld, opB == QX_OP_N ? 1 : 0, opB == QX_OP_C ? 1 : 0, n, k,
                     k_pad, slices, dB, sBd, eb, (int64_t)n);
  OzArgs g;
  g.m = m;
  g.n = n;
  g.k_pad = k_pad;
  g.slices = slices;
  g.ar = alpha.re;
  g.ai = alpha.im;
  g.br = beta.re;
  g.bi = beta.im;
  g.Ad = dA;
  g.sAd = sAd;
  g.Bd = dB;
  g.sBd = sBd;
  g.ea = ea;
  g.eb = eb;
  g.C = C.p;
  g.sC = C.stride;
  g.ldc = C.ld;
  g.Cin = Cin.zero() ? C.p : Cin.p;
  g.sCin = Cin.zero() ? C.stride : Cin.stride;
  g.ldcin = Cin.zero() ? C.ld : Cin.ld;
  g.tilesM = (m + OBM - 1) / OBM;
  g.tilesN = (n + OBN - 1) / OBN;
  g.beta_zero = ((beta.re == 0.0 && beta.im == 0.0) || Cin.zero()) ? 1 : 0;
  const int64_t blocks = batch * g.tilesM * g.tilesN;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "zgemm_ozaki: grid too large");
  switch (slices) {
#define QX_OZ_CASE(S)                                                                   \
  case S:                                                                               \
    hipLaunchKernelGGL(ozaki_gemm_kernel<S>, dim3((unsigned)blocks), dim3(ONT), 0, s, g); \
    break;
    QX_OZ_CASE(2)
    QX_OZ_CASE(3)
    QX_OZ_CASE(4)
    QX_OZ_CASE(5)
    QX_OZ_CASE(6)
    QX_OZ_CASE(7)
#undef QX_OZ_CASE
  }
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::GEMM, 8.0 * m * n * (double)k * batch);
}

}  // namespace qx
