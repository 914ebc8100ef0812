// Batched complex matrix inverse with partial pivoting (gfx950).
//
// Replaces the batched `linalg.inv` of the reference
// (src/qttools/kernels/linalg/inv.py:24-40, LAPACK/cuSOLVER getrf+getri).
//
// Algorithm: blocked in-place Gauss-Jordan elimination, one workgroup per
// matrix, the matrix itself resident in HBM / L2:
//   for each panel of NB = 16 columns
//     1. the n x NB panel is copied to LDS and reduced there with partial
//        pivoting (pivot = first row of maximal |re| + |im|, as LAPACK's
//        izamax); every step scales the pivot row and eliminates the pivot
//        column from all other rows;
//     2. the panel's row interchanges are applied to the other columns;
//     3. the other columns receive the rank-NB update
//          X[:, C] <- X[:, C] (pivot rows zeroed) + P * X[pivot rows, C]
//        on the f64 (or f32) matrix cores, 16 x 16 tiles per wave;
//     4. the reduced panel is written back.
//   Finally the column interchanges are undone in reverse order.
// The flop count equals getrf + getri (8 n^3 real flops for complex).
//
// Precision: PREC_C128 computes in double; PREC_C64 rounds the panel and the
// update operands to float (the update runs on f32 MFMA) while storing the
// result back in complex128.
#include "qx_internal.h"

namespace qx {
namespace {

constexpr int NB = 16;
constexpr int NT = 256;
constexpr int PST = NB + 1;  // LDS row stride of the panel (bank spread)
constexpr int XST = NB + 1;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct InvArgs {
  int n;
  const z2* A;
  int64_t sA;
  int lda;
  const z2* S;
  int64_t sS;
  int lds;
  z2* X;
  int64_t sX;
  int ldx;
  int* info;
};

template <typename T>
struct InvMma;
template <>
struct InvMma<double> {
  typedef d4 acc_t;
  static __device__ inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <>
struct InvMma<float> {
  typedef f4 acc_t;
  static __device__ inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

__device__ inline void cdiv_one(double ar, double ai, double& qr, double& qi) {
  // 1 / (ar + i ai) with scaling against overflow (Smith's method).
  if (fabs(ar) >= fabs(ai)) {
    const double r = ai / ar, d = ar + ai * r;
    qr = 1.0 / d;
    qi = -r / d;
  } else {
    const double r = ar / ai, d = ai + ar * r;
    qr = r / d;
    qi = -1.0 / d;
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void zinv_kernel(InvArgs g) {
  typedef InvMma<T> MM;
  typedef typename MM::acc_t acc_t;
  extern __shared__ double smem[];
  const int n = g.n;
  T* Pr = reinterpret_cast<T*>(smem);
  T* Pi = Pr + n * PST;
  T* Xr = Pi + n * PST;
  T* Xi = Xr + NB * XST;
  int* perm = reinterpret_cast<int*>(Xi + NB * XST);
  double* red_v = reinterpret_cast<double*>(perm + ((n + 1) & ~1));
  int* red_i = reinterpret_cast<int*>(red_v + NT / 64);

  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  z2* __restrict__ X = g.X + e * g.sX;
  const int ldx = g.ldx;

  // 0. X = A - S
  {
    const z2* A = g.A + e * g.sA;
    const z2* S = g.S ? g.S + e * g.sS : nullptr;
    const bool same = (A == X) && S == nullptr && g.lda == ldx;
    if (!same) {
      for (int idx = tid; idx < n * n; idx += NT) {
        const int r = idx / n, c = idx - r * n;
        z2 v = A[(int64_t)r * g.lda + c];
        if (S) {
          const z2 sv = S[(int64_t)r * g.lds + c];
          v.x -= sv.x;
          v.y -= sv.y;
        }
        X[(int64_t)r * ldx + c] = v;
      }
    }
  }
  __syncthreads();

  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nbw = min(NB, n - k0);
    // 1a. panel -> LDS
    for (int idx = tid; idx < n * NB; idx += NT) {
      const int r = idx / NB, c = idx - r * NB;
      z2 v = make_double2(0.0, 0.0);
      if (c < nbw) v = X[(int64_t)r * ldx + k0 + c];
      Pr[r * PST + c] = (T)v.x;
      Pi[r * PST + c] = (T)v.y;
    }
    __syncthreads();

    // 1b. unblocked Gauss-Jordan on the panel
    for (int kk = 0; kk < nbw; ++kk) {
      const int k = k0 + kk;
      double best = -1.0;
      int bi = n;
      for (int r = k + tid; r < n; r += NT) {
        const double v = fabs((double)Pr[r * PST + kk]) + fabs((double)Pi[r * PST + kk]);
        if (v > best) {
          best = v;
          bi = r;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) {
          best = ov;
          bi = oi;
        }
      }
      if (lane == 0) {
        red_v[w] = best;
        red_i[w] = bi;
      }
      __syncthreads();
      best = red_v[0];
      bi = red_i[0];
#pragma unroll
      for (int q = 1; q < NT / 64; ++q) {
        const double ov = red_v[q];
        const int oi = red_i[q];
        if (ov > best || (ov == best && oi < bi)) {
          best = ov;
          bi = oi;
        }
      }
      const int p = (bi < n) ? bi : k;
      if (tid == 0) {
        perm[k] = p;
        if (!(best > 0.0)) g.info[e] = 1;  // plain vector store, idempotent
      }
      if (p != k && tid < NB) {
        const T tr = Pr[k * PST + tid], ti = Pi[k * PST + tid];
        Pr[k * PST + tid] = Pr[p * PST + tid];
        Pi[k * PST + tid] = Pi[p * PST + tid];
        Pr[p * PST + tid] = tr;
        Pi[p * PST + tid] = ti;
      }
      __syncthreads();

      // pivot row, scaled, in registers of every thread
      double rkr[NB], rki[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        rkr[c] = (double)Pr[k * PST + c];
        rki[c] = (double)Pi[k * PST + c];
      }
      double ivr, ivi;
      cdiv_one((double)Pr[k * PST + kk], (double)Pi[k * PST + kk], ivr, ivi);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const double a = rkr[c], b = rki[c];
        rkr[c] = a * ivr - b * ivi;
        rki[c] = a * ivi + b * ivr;
      }
      // in-place trick: the pivot column of the pivot row holds 1/pivot
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c == kk) {
          rkr[c] = ivr;
          rki[c] = ivi;
        }
      __syncthreads();
      for (int r = tid; r < n; r += NT) {
        if (r == k) {
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            Pr[r * PST + c] = (T)rkr[c];
            Pi[r * PST + c] = (T)rki[c];
          }
        } else {
          const double fr = (double)Pr[r * PST + kk], fi = (double)Pi[r * PST + kk];
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            double xr = (c == kk) ? 0.0 : (double)Pr[r * PST + c];
            double xi = (c == kk) ? 0.0 : (double)Pi[r * PST + c];
            xr -= fr * rkr[c] - fi * rki[c];
            xi -= fr * rki[c] + fi * rkr[c];
            Pr[r * PST + c] = (T)xr;
            Pi[r * PST + c] = (T)xi;
          }
        }
      }
      __syncthreads();
    }

    // 2. row interchanges of this panel on the other columns
    for (int c = tid; c < n; c += NT) {
      if (c >= k0 && c < k0 + nbw) continue;
      for (int kk = 0; kk < nbw; ++kk) {
        const int k = k0 + kk, p = perm[k];
        if (p != k) {
          const z2 a = X[(int64_t)k * ldx + c];
          X[(int64_t)k * ldx + c] = X[(int64_t)p * ldx + c];
          X[(int64_t)p * ldx + c] = a;
        }
      }
    }
    __syncthreads();

    // 3. rank-NB update of the other columns, NB columns at a time
    const int nrt = (n + 15) / 16;
    for (int cc = 0; cc < n; cc += NB) {
      if (cc == k0) continue;
      for (int idx = tid; idx < NB * NB; idx += NT) {
        const int r = idx / NB, c = idx - r * NB;
        z2 v = make_double2(0.0, 0.0);
        if (r < nbw && cc + c < n) v = X[(int64_t)(k0 + r) * ldx + cc + c];
        Xr[r * XST + c] = (T)v.x;
        Xi[r * XST + c] = (T)v.y;
      }
      __syncthreads();
      const int col = cc + (lane & 15);
      for (int rt = w; rt < nrt; rt += NT / 64) {
        acc_t aR, aI;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = rt * 16 + MM::row(lane, q);
          z2 v = make_double2(0.0, 0.0);
          if (row < n && col < n && !(row >= k0 && row < k0 + nbw))
            v = X[(int64_t)row * ldx + col];
          aR[q] = (T)v.x;
          aI[q] = (T)v.y;
        }
        const int arow = rt * 16 + (lane & 15);
#pragma unroll
        for (int kq = 0; kq < NB / 4; ++kq) {
          const int kc = kq * 4 + (lane >> 4);
          const T ar = (arow < n) ? Pr[arow * PST + kc] : (T)0;
          const T ai = (arow < n) ? Pi[arow * PST + kc] : (T)0;
          const T br = Xr[kc * XST + (lane & 15)];
          const T bi2 = Xi[kc * XST + (lane & 15)];
          aR = MM::mma(ar, br, aR);
          aI = MM::mma(ar, bi2, aI);
          aR = MM::mma(-ai, bi2, aR);
          aI = MM::mma(ai, br, aI);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = rt * 16 + MM::row(lane, q);
          if (row < n && col < n)
            X[(int64_t)row * ldx + col] = make_double2((double)aR[q], (double)aI[q]);
        }
      }
      __syncthreads();
    }

    // 4. reduced panel -> X
    for (int idx = tid; idx < n * NB; idx += NT) {
      const int r = idx / NB, c = idx - r * NB;
      if (c < nbw) X[(int64_t)r * ldx + k0 + c] = make_double2((double)Pr[r * PST + c], (double)Pi[r * PST + c]);
    }
    __syncthreads();
  }

  // undo the column interchanges in reverse order
  for (int r = tid; r < n; r += NT) {
    z2* xr = X + (int64_t)r * ldx;
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        const z2 a = xr[k];
        xr[k] = xr[p];
        xr[p] = a;
      }
    }
  }
}

size_t smem_bytes(int n, int elem) {
  return (size_t)(2 * n * PST + 2 * NB * XST) * elem + (size_t)((n + 1) & ~1) * 4 +
         (NT / 64) * (8 + 4);
}

}  // namespace

void zinv(hipStream_t s, int64_t batch, int n, const Mat& A, const Mat& S,
          const Mat& X, int* info, int precision) {
  QX_REQUIRE(n >= 1 && n <= kInvMaxN, "inv: block size must be in [1, 512]");
  QX_REQUIRE(!A.zero() && !X.zero(), "inv: null operand");
  if (batch <= 0) return;
  QX_REQUIRE(batch < (int64_t(1) << 31), "inv: batch too large");
  InvArgs g{n, A.p, A.stride, A.ld, S.p, S.stride, S.ld, X.p, X.stride, X.ld, info};
  const bool f32 = precision == QX_PREC_C64;
  const size_t sm = smem_bytes(n, f32 ? 4 : 8);
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);
  if (f32) {
    static bool attr = false;
    if (!attr) {
      QX_HIP_CHECK(hipFuncSetAttribute((const void*)zinv_kernel<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      attr = true;
    }
    hipLaunchKernelGGL(zinv_kernel<float>, dim3((unsigned)batch), dim3(NT), sm, s, g);
  } else {
    static bool attr = false;
    if (!attr) {
      QX_HIP_CHECK(hipFuncSetAttribute((const void*)zinv_kernel<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      attr = true;
    }
    hipLaunchKernelGGL(zinv_kernel<double>, dim3((unsigned)batch), dim3(NT), sm, s, g);
  }
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::INV, 8.0 * (double)n * n * n * batch);
}

}  // namespace qx
