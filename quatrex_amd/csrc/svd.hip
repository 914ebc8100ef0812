// Batched singular value decomposition of tall complex matrices by one-sided
// Jacobi rotations (Hestenes), for the projection step of Beyn's contour
// method (the SVD of the probed moment P0, src/qttools/nevp/beyn.py:96-145).
//
// A = U diag(S) V^H for each d x m matrix A (d >= m), S descending.  One
// workgroup per matrix; the columns of G = A V are kept column-major in the
// workspace (each column contiguous), V likewise.  A sweep visits every
// column pair once in m' - 1 rounds of the round-robin tournament ordering
// (m' = m rounded up to even): the m' / 2 pairs of a round are disjoint and
// run concurrently, one wave per pair.  For a pair (p, q) with
// alpha = |g_p|^2, beta = |g_q|^2, gamma = g_p^H g_q = |gamma| e^(i phi), the
// rotation
//   g_p <- c g_p - s e^(-i phi) g_q,   g_q <- s e^(i phi) g_p + c g_q
// (and the same on the columns of V) with t = s / c the smaller root of
// t^2 + 2 zeta t - 1 = 0, zeta = (beta - alpha) / (2 |gamma|), makes the two
// columns orthogonal.  Pairs already orthogonal to |gamma| <= tol sqrt(alpha
// beta) are skipped; the iteration stops after a sweep without a rotation or
// after max_sweeps.  Then S_j = |g_j|, U_j = g_j / S_j, sorted by S.
#include <cmath>

#include "qx_internal.h"

namespace qx {
namespace {

constexpr int kSvdThreads = 1024;
constexpr int kSvdWaves = kSvdThreads / 64;

struct SvdArgs {
  int d, m, max_sweeps;
  double tol;
  const z2* A;  // (d, m) row-major, leading dimension lda
  int64_t sA;
  int lda;
  z2* G;  // workspace: m columns of d per matrix
  z2* Vw;  // workspace: m columns of m per matrix
  z2* U;  // (d, m) row-major
  int64_t sU;
  double* S;  // (m)
  int64_t sS;
  z2* V;  // (m, m) row-major
  int64_t sV;
  int* sweeps;  // per matrix, nullable
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Player at position k of round r of the round-robin ordering of mp players
// (mp even): position 0 is fixed, the others rotate by one per round.
__device__ inline int rr_player(int k, int r, int mp) {
  if (k == 0) return 0;
  return 1 + (k - 1 + r) % (mp - 1);
}

__global__ __launch_bounds__(kSvdThreads) void jacobi_svd_kernel(SvdArgs g) {
  __shared__ int rotated;
  __shared__ int order[512];
  const int64_t e = blockIdx.x;
  const int d = g.d, m = g.m;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  z2* G = g.G + e * (int64_t)m * d;
  z2* Vw = g.Vw + e * (int64_t)m * m;
  const z2* A = g.A + e * g.sA;
  for (int idx = tid; idx < m * d; idx += kSvdThreads) {
    const int j = idx / d, r = idx - j * d;
    G[idx] = A[(int64_t)r * g.lda + j];
  }
  for (int idx = tid; idx < m * m; idx += kSvdThreads) {
    const int j = idx / m, i = idx - j * m;
    Vw[idx] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  const int mp = m + (m & 1);
  int sweep = 0;
  for (; sweep < g.max_sweeps; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < mp - 1; ++round) {
      for (int k = w; k < mp / 2; k += kSvdWaves) {
        int p = rr_player(k, round, mp), q = rr_player(mp - 1 - k, round, mp);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q >= m) continue;  // the idle player of an odd m
        z2* gp = G + (int64_t)p * d;
        z2* gq = G + (int64_t)q * d;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int r = lane; r < d; r += 64) {
          const z2 x = gp[r], y = gq[r];
          al = fma(x.x, x.x, fma(x.y, x.y, al));
          be = fma(y.x, y.x, fma(y.y, y.y, be));
          gr = fma(x.x, y.x, fma(x.y, y.y, gr));   // Re conj(x) y
          gi = fma(x.x, y.y, fma(-x.y, y.x, gi));  // Im conj(x) y
        }
        al = wave_sum(al);
        be = wave_sum(be);
        gr = wave_sum(gr);
        gi = wave_sum(gi);
        const double ga = sqrt(gr * gr + gi * gi);
        if (!(ga > g.tol * sqrt(al * be))) continue;  // uniform over the wave
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        const double er = gr / ga, ei = gi / ga;  // e^(i phi)
        // g_p' = c g_p - s e^(-i phi) g_q,  g_q' = s e^(i phi) g_p + c g_q
        for (int r = lane; r < d; r += 64) {
          const z2 x = gp[r], y = gq[r];
          const double yr = er * y.x + ei * y.y, yi = er * y.y - ei * y.x;  // e^(-i phi) y
          const double xr = er * x.x - ei * x.y, xi = er * x.y + ei * x.x;  // e^(i phi) x
          gp[r] = make_double2(c * x.x - s * yr, c * x.y - s * yi);
          gq[r] = make_double2(s * xr + c * y.x, s * xi + c * y.y);
        }
        z2* vp = Vw + (int64_t)p * m;
        z2* vq = Vw + (int64_t)q * m;
        for (int r = lane; r < m; r += 64) {
          const z2 x = vp[r], y = vq[r];
          const double yr = er * y.x + ei * y.y, yi = er * y.y - ei * y.x;
          const double xr = er * x.x - ei * x.y, xi = er * x.y + ei * x.x;
          vp[r] = make_double2(c * x.x - s * yr, c * x.y - s * yi);
          vq[r] = make_double2(s * xr + c * y.x, s * xi + c * y.y);
        }
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  if (g.sweeps && tid == 0) g.sweeps[e] = sweep;
  // singular values (column norms), then the descending order by rank
  double* S = g.S + e * g.sS;
  for (int j = w; j < m; j += kSvdWaves) {
    const z2* gj = G + (int64_t)j * d;
    double nrm = 0.0;
    for (int r = lane; r < d; r += 64) nrm = fma(gj[r].x, gj[r].x, fma(gj[r].y, gj[r].y, nrm));
    nrm = wave_sum(nrm);
    if (lane == 0) S[j] = sqrt(nrm);
  }
  __syncthreads();
  for (int j = tid; j < m; j += kSvdThreads) {
    const double sj = S[j];
    int rank = 0;
    for (int k = 0; k < m; ++k) {
      const double sk = S[k];
      rank += (sk > sj || (sk == sj && k < j)) ? 1 : 0;
    }
    order[rank] = j;
  }
  __syncthreads();
  // U and V columns in rank order; S rewritten in rank order below
  z2* U = g.U + e * g.sU;
  z2* V = g.V + e * g.sV;
  for (int idx = tid; idx < d * m; idx += kSvdThreads) {
    const int r = idx / m, c = idx - r * m;
    const int j = order[c];
    const double sj = S[j];
    const z2 x = G[(int64_t)j * d + r];
    U[(int64_t)r * m + c] = sj > 0.0 ? make_double2(x.x / sj, x.y / sj) : make_double2(0.0, 0.0);
  }
  for (int idx = tid; idx < m * m; idx += kSvdThreads) {
    const int r = idx / m, c = idx - r * m;
    V[(int64_t)r * m + c] = Vw[(int64_t)order[c] * m + r];
  }
  __syncthreads();
  double mine = 0.0;
  if (tid < m) mine = S[order[tid]];
  __syncthreads();
  if (tid < m) S[tid] = mine;
}

}  // namespace

size_t svd_workspace_bytes(int64_t n, int d, int m) {
  return (size_t)(n * ((int64_t)m * d + (int64_t)m * m)) * sizeof(z2);
}

void svd_tall(hipStream_t s, int64_t n, int d, int m, const Mat& A, const Mat& U, double* S,
              int64_t sS, const Mat& V, int max_sweeps, double tol, void* ws, size_t ws_bytes,
              int* sweeps) {
  QX_REQUIRE(d >= 1 && m >= 1 && m <= d && m <= 512, "svd: need 1 <= cols <= rows, cols <= 512");
  QX_REQUIRE(!A.zero() && !U.zero() && S && !V.zero(), "svd: null operand");
  QX_REQUIRE(max_sweeps >= 1 && tol > 0.0, "svd: bad iteration limits");
  if (n <= 0) return;
  QX_REQUIRE(n < (int64_t(1) << 31), "svd: batch too large");
  QX_REQUIRE(ws_bytes >= svd_workspace_bytes(n, d, m), "svd: workspace too small");
  SvdArgs g;
  g.d = d;
  g.m = m;
  g.max_sweeps = max_sweeps;
  g.tol = tol;
  g.A = A.p;
  g.sA = A.stride;
  g.lda = A.ld;
  g.G = static_cast<z2*>(ws);
  g.Vw = g.G + n * (int64_t)m * d;
  g.U = U.p;
  g.sU = U.stride;
  g.S = S;
  g.sS = sS;
  g.V = V.p;
  g.sV = V.stride;
  g.sweeps = sweeps;
  QX_REQUIRE(U.ld == m && V.ld == m, "svd: U and V must be contiguous (leading dimension = cols)");
  hipLaunchKernelGGL(jacobi_svd_kernel, dim3((unsigned)n), dim3(kSvdThreads), 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
}

}  // namespace qx
