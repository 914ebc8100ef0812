// Elementwise and reduction kernels around the GEMM / inverse workhorses.
//
// The conjugate-transpose forms (X - X^H, T^H - T, -X^H) read the transposed
// operand through a 32 x 32 LDS tile so that both the plain and the
// transposed reads are coalesced.  Reductions (traces, norms) use one
// workgroup per matrix and a fixed summation order, so results do not
// depend on how many matrices share a launch.
#include "qx_internal.h"

namespace qx {
namespace {

constexpr int TS = 32;    // tile edge
constexpr int NT = 256;   // threads per workgroup

struct AxpbyArgs {
  int m, n;
  double ar, ai, br, bi;
  const z2* A;
  int64_t sA;
  int lda;
  const z2* B;
  int64_t sB;
  int ldb;
  z2* Y;
  int64_t sY;
  int ldy;
};

__global__ __launch_bounds__(NT) void axpby_kernel(AxpbyArgs g, int64_t total) {
  const int64_t mn = (int64_t)g.m * g.n;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t e = i / mn;
    const int64_t rc = i - e * mn;
    const int r = (int)(rc / g.n), c = (int)(rc - (int64_t)r * g.n);
    double yr = 0.0, yi = 0.0;
    if (g.A) {
      const z2 a = g.A[e * g.sA + (int64_t)r * g.lda + c];
      yr += g.ar * a.x - g.ai * a.y;
      yi += g.ar * a.y + g.ai * a.x;
    }
    if (g.B) {
      const z2 b = g.B[e * g.sB + (int64_t)r * g.ldb + c];
      yr += g.br * b.x - g.bi * b.y;
      yi += g.br * b.y + g.bi * b.x;
    }
    g.Y[e * g.sY + (int64_t)r * g.ldy + c] = make_double2(yr, yi);
  }
}

__global__ __launch_bounds__(NT) void ident_sub_kernel(z2* Y, int64_t sY, int ldy, int n,
                                                       int64_t total) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t e = i / nn;
    const int64_t rc = i - e * nn;
    const int r = (int)(rc / n), c = (int)(rc - (int64_t)r * n);
    z2* y = Y + e * sY + (int64_t)r * ldy + c;
    const z2 v = *y;
    *y = make_double2((r == c ? 1.0 : 0.0) - v.x, -v.y);
  }
}

// Y = cy*Y + ca*A + cb*conj(B^T) + cc*D; Y, A, D are m x n, B is n x m.  Real coefs.
struct HermArgs {
  int m, n;
  double cy, ca, cb;
  const z2* A;
  int64_t sA;
  int lda;
  const z2* B;
  int64_t sB;
  int ldb;
  z2* Y;
  int64_t sY;
  int ldy;
  int tm, tn;
  double cc;
  const z2* D;
  int64_t sD;
  int ldd;
  // Two groups of matrices in one launch: matrices e >= half use the second
  // A, B and Y (index e - half); D is single-group only.  One group sets
  // half = batch and the second set equal to the first.
  int64_t half;
  const z2* A2;
  int64_t sA2;
  int lda2;
  const z2* B2;
  int64_t sB2;
  int ldb2;
  z2* Y2;
  int64_t sY2;
  int ldy2;
};

// Workgroup ids are dealt round-robin over the 8 XCDs; this maps them so
// that each XCD receives one contiguous range of logical ids (the tiles of
// one matrix together).
__device__ inline int tile_xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx, rem = nblocks % nx;
  const int x = bid % nx, j = bid / nx;
  return x * per + min(x, rem) + j;
}

// Tiles (r, c) and (c, r) read each other's data: with the tiles of a matrix
// on one XCD the second read comes from its L2.
__global__ __launch_bounds__(NT) void herm_kernel(HermArgs g) {
  __shared__ z2 tile[TS][TS + 1];
  const int tiles = g.tm * g.tn;
  const int id = tile_xcd_remap(blockIdx.x, gridDim.x);
  const int64_t eg = id / tiles;
  const int t = id - (int)(eg * tiles);
  const bool second = eg >= g.half;
  const int64_t e = second ? eg - g.half : eg;
  const z2* A = second ? g.A2 + e * g.sA2 : (g.A ? g.A + e * g.sA : nullptr);
  const z2* B = second ? g.B2 + e * g.sB2 : (g.B ? g.B + e * g.sB : nullptr);
  z2* Y = second ? g.Y2 + e * g.sY2 : g.Y + e * g.sY;
  const int lda = second ? g.lda2 : g.lda, ldb = second ? g.ldb2 : g.ldb,
            ldy = second ? g.ldy2 : g.ldy;
  const int r0 = (t / g.tn) * TS, c0 = (t % g.tn) * TS;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  if (g.B) {
    // tile[j][i] = B[c0 + j][r0 + i]
    for (int j = ty; j < TS; j += 8) {
      const int br = c0 + j, bc = r0 + tx;
      z2 v = make_double2(0.0, 0.0);
      if (br < g.n && bc < g.m) v = B[(int64_t)br * ldb + bc];
      tile[j][tx] = v;
    }
  }
  __syncthreads();
  for (int i = ty; i < TS; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r >= g.m || c >= g.n) continue;
    double yr = 0.0, yi = 0.0;
    z2* y = Y + (int64_t)r * ldy + c;
    if (g.cy != 0.0) {
      const z2 v = *y;
      yr = g.cy * v.x;
      yi = g.cy * v.y;
    }
    if (g.A) {
      const z2 a = A[(int64_t)r * lda + c];
      yr += g.ca * a.x;
      yi += g.ca * a.y;
    }
    if (g.B) {
      const z2 b = tile[tx][i];  // B[c][r]
      yr += g.cb * b.x;
      yi -= g.cb * b.y;
    }
    if (g.D) {
      const z2 d = g.D[e * g.sD + (int64_t)r * g.ldd + c];
      yr += g.cc * d.x;
      yi += g.cc * d.y;
    }
    *y = make_double2(yr, yi);
  }
}

// One launch over batch matrices, or over two groups of batch matrices when
// A2 / B2 / Y2 are given (the second group's operands; D must then be zero).
void herm(hipStream_t s, int64_t batch, int m, int n, double cy, double ca,
          const Mat& A, double cb, const Mat& B, const Mat& Y, double cc = 0.0,
          const Mat& D = Mat(), const Mat* A2 = nullptr, const Mat* B2 = nullptr,
          const Mat* Y2 = nullptr) {
  if (batch <= 0 || m <= 0 || n <= 0) return;
  const bool two = Y2 != nullptr;
  QX_REQUIRE(!two || (D.zero() && A2 && B2 && A2->zero() == A.zero() &&
                      B2->zero() == B.zero() && !Y2->zero()),
             "herm: a two-group launch needs matching operands and no D term");
  const Mat& a2 = two ? *A2 : A;
  const Mat& b2 = two ? *B2 : B;
  const Mat& y2 = two ? *Y2 : Y;
  HermArgs g{m, n, cy, ca, cb, A.p, A.stride, A.ld, B.p, B.stride, B.ld,
             Y.p, Y.stride, Y.ld, (m + TS - 1) / TS, (n + TS - 1) / TS,
             cc, D.p, D.stride, D.ld, batch,
             a2.p, a2.stride, a2.ld, b2.p, b2.stride, b2.ld, y2.p, y2.stride, y2.ld};
  const int64_t blocks = (two ? 2 : 1) * batch * g.tm * g.tn;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "herm: grid too large");
  hipLaunchKernelGGL(herm_kernel, dim3((unsigned)blocks), dim3(NT), 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
}

// Contact self-energies from Sigma^R (one 32 x 32 tile per workgroup, the
// transposed tile through LDS): D = S - S^H, L = -f D, G = (1 - f) D with
// one occupation f per matrix (i f Gamma and i (f - 1) Gamma for
// Gamma = i (S - S^H)).
__global__ __launch_bounds__(NT) void contact_sigma_kernel(const z2* S, int64_t sS, int n,
                                                           const double* f, z2* L, int64_t sL,
                                                           z2* G, int64_t sG, int tn) {
  __shared__ z2 tile[TS][TS + 1];
  const int tiles = tn * tn;
  const int id = tile_xcd_remap(blockIdx.x, gridDim.x);  // as herm_kernel
  const int64_t e = id / tiles;
  const int t = id - (int)(e * tiles);
  const int r0 = (t / tn) * TS, c0 = (t % tn) * TS;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const z2* Se = S + e * sS;
  for (int j = ty; j < TS; j += 8) {
    const int sr = c0 + j, sc = r0 + tx;
    z2 v = make_double2(0.0, 0.0);
    if (sr < n && sc < n) v = Se[(int64_t)sr * n + sc];
    tile[j][tx] = v;  // S[c0 + j][r0 + tx]
  }
  __syncthreads();
  const double fe = f[e];
  for (int i = ty; i < TS; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r >= n || c >= n) continue;
    const z2 a = Se[(int64_t)r * n + c];
    const z2 b = tile[tx][i];  // S[c][r]
    const double dr = a.x - b.x, di = a.y + b.y;
    L[e * sL + (int64_t)r * n + c] = make_double2(-fe * dr, -fe * di);
    G[e * sG + (int64_t)r * n + c] = make_double2((1.0 - fe) * dr, (1.0 - fe) * di);
  }
}

__global__ __launch_bounds__(NT) void diag_imag_kernel(const z2* X, int64_t sX, int ldx, int n,
                                                       double* out, int64_t so, int64_t total,
                                                       double scale) {
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t e = i / n;
    const int r = (int)(i - e * n);
    out[e * so + r] = scale * X[e * sX + (int64_t)r * ldx + r].y;
  }
}

// System matrix on the nnz pattern: A[e, k] = z_e S_k + T_k - Sig[e, k].
__global__ __launch_bounds__(NT) void assemble_kernel(const z2* z, const z2* S, const z2* T,
                                                      const z2* Sig, int64_t sSig, z2* A,
                                                      int64_t sA, int64_t nnz, int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t e = i / nnz;
    const int64_t k = i - e * nnz;
    double vr = 0.0, vi = 0.0;
    if (S) {
      const z2 ze = z[e], sk = S[k];
      vr = ze.x * sk.x - ze.y * sk.y;
      vi = ze.x * sk.y + ze.y * sk.x;
    }
    if (T) {
      const z2 tk = T[k];
      vr += tk.x;
      vi += tk.y;
    }
    if (Sig) {
      const z2 sg = Sig[e * sSig + k];
      vr -= sg.x;
      vi -= sg.y;
    }
    A[e * sA + k] = make_double2(vr, vi);
  }
}

// Block reduction of NT doubles in a fixed order.
__device__ inline double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) s += red[q];
  __syncthreads();
  return s;
}

struct TraceArgs {
  int n;
  const z2 *X, *Y, *Z, *W;
  int64_t sX, sY, sZ, sW;
  int lx, ly, lz, lw;
  z2* out;
  int64_t so;
};

// tr(X Y - Z W) = sum_ab X[a][b] Y[b][a] - Z[a][b] W[b][a]
__global__ __launch_bounds__(NT) void trace_kernel(TraceArgs g) {
  __shared__ double red[NT / 64];
  const int64_t e = blockIdx.x;
  const int n = g.n;
  double sr = 0.0, si = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int a = idx / n, b = idx - a * n;
    if (g.X) {
      const z2 x = g.X[e * g.sX + (int64_t)a * g.lx + b];
      const z2 y = g.Y[e * g.sY + (int64_t)b * g.ly + a];
      sr += x.x * y.x - x.y * y.y;
      si += x.x * y.y + x.y * y.x;
    }
    if (g.Z) {
      const z2 z = g.Z[e * g.sZ + (int64_t)a * g.lz + b];
      const z2 w = g.W[e * g.sW + (int64_t)b * g.lw + a];
      sr -= z.x * w.x - z.y * w.y;
      si -= z.x * w.y + z.y * w.x;
    }
  }
  sr = block_sum(sr, red);
  si = block_sum(si, red);
  if (threadIdx.x == 0) g.out[e * g.so] = make_double2(sr, si);
}

__global__ __launch_bounds__(NT) void absnorm_kernel(int n, const z2* A, int64_t sA, int lda,
                                                     const z2* B, int64_t sB, int ldb,
                                                     double* out) {
  __shared__ double red[NT / 64];
  const int64_t e = blockIdx.x;
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int r = idx / n, c = idx - r * n;
    const z2 a = A[e * sA + (int64_t)r * lda + c];
    const z2 b = B[e * sB + (int64_t)r * ldb + c];
    const double v = hypot(a.x, a.y) + hypot(b.x, b.y);
    acc += v * v;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) out[e] = sqrt(acc);
}

// Newton-Schulz check of one complex64 block inverse per workgroup (see
// ew_ns_check).  Every thread reads the flag after the block sum, so the
// zeroing pass is uniform over the workgroup.
__global__ __launch_bounds__(NT) void ns_check_kernel(int n, z2* R, int64_t sR, int ldr,
                                                      double tol, int* flagged, int* info) {
  __shared__ double red[NT / 64];
  __shared__ int flag;
  const int64_t e = blockIdx.x;
  z2* Re = R + e * sR;
  double acc = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += NT) {
    const int r = idx / n, c = idx - r * n;
    const z2 v = Re[(int64_t)r * ldr + c];
    acc += v.x * v.x + v.y * v.y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) {
    // NaN and Inf fail the comparison
    flag = sqrt(acc) < tol ? 0 : 1;
    flagged[e] = flag;
    if (flag) info[e] |= 2;
  }
  __syncthreads();
  if (flag)
    for (int idx = threadIdx.x; idx < n * n; idx += NT) {
      const int r = idx / n, c = idx - r * n;
      Re[(int64_t)r * ldr + c] = make_double2(0.0, 0.0);
    }
}

struct ContourArgs {
  int nq, b;
  const z2 *aji, *aii, *aij;
  int64_t s0, s1, s2;
  int l0, l1, l2;
  const z2* z;
  z2* T;
  int64_t sT;
  int lT;
};

__device__ inline z2 cmul(z2 a, z2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// T[e*nq + q] = aji[e] / z_q + aii[e] + aij[e] z_q  (operator.py:57-110)
__global__ __launch_bounds__(NT) void contour_kernel(ContourArgs g, int64_t total) {
  const int64_t bb = (int64_t)g.b * g.b;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t eq = i / bb;
    const int64_t rc = i - eq * bb;
    const int r = (int)(rc / g.b), c = (int)(rc - (int64_t)r * g.b);
    const int64_t e = eq / g.nq;
    const int q = (int)(eq - e * g.nq);
    const z2 z = g.z[q];
    const double den = z.x * z.x + z.y * z.y;
    const z2 zi = make_double2(z.x / den, -z.y / den);
    const z2 a0 = g.aji[e * g.s0 + (int64_t)r * g.l0 + c];
    const z2 a1 = g.aii[e * g.s1 + (int64_t)r * g.l1 + c];
    const z2 a2 = g.aij[e * g.s2 + (int64_t)r * g.l2 + c];
    const z2 t0 = cmul(a0, zi), t2 = cmul(a2, z);
    g.T[eq * g.sT + (int64_t)r * g.lT + c] =
        make_double2(t0.x + a1.x + t2.x, t0.y + a1.y + t2.y);
  }
}

struct AccArgs {
  int nq, b;
  const z2* Ti;
  int64_t sT;
  int lT;
  const z2 *w0, *w1;
  z2 *q0, *q1;
  int64_t sq;
  int lq;
};

__global__ __launch_bounds__(NT) void contour_acc_kernel(AccArgs g, int64_t total) {
  const int64_t bb = (int64_t)g.b * g.b;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * NT) {
    const int64_t e = i / bb;
    const int64_t rc = i - e * bb;
    const int r = (int)(rc / g.b), c = (int)(rc - (int64_t)r * g.b);
    z2 s0 = make_double2(0, 0), s1 = make_double2(0, 0);
    for (int q = 0; q < g.nq; ++q) {
      const z2 t = g.Ti[(e * g.nq + q) * g.sT + (int64_t)r * g.lT + c];
      const z2 a = cmul(g.w0[q], t), b = cmul(g.w1[q], t);
      s0.x += a.x;
      s0.y += a.y;
      s1.x += b.x;
      s1.y += b.y;
    }
    z2* p0 = g.q0 + e * g.sq + (int64_t)r * g.lq + c;
    z2* p1 = g.q1 + e * g.sq + (int64_t)r * g.lq + c;
    const z2 o0 = *p0, o1 = *p1;
    *p0 = make_double2(o0.x + s0.x, o0.y + s0.y);
    *p1 = make_double2(o1.x + s1.x, o1.y + s1.y);
  }
}

unsigned grid_for(int64_t total) {
  int64_t b = (total + NT - 1) / NT;
  if (b > 65536 * 4) b = 65536 * 4;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

void ew_axpby(hipStream_t s, int64_t batch, int m, int n, Cplx a, const Mat& A,
              Cplx b, const Mat& B, const Mat& Y) {
  if (batch <= 0 || m <= 0 || n <= 0) return;
  QX_REQUIRE(!Y.zero(), "axpby: null output");
  const bool use_a = !A.zero() && (a.re != 0.0 || a.im != 0.0);
  const bool use_b = !B.zero() && (b.re != 0.0 || b.im != 0.0);
  AxpbyArgs g{m, n, a.re, a.im, b.re, b.im,
              use_a ? A.p : nullptr, A.stride, A.ld,
              use_b ? B.p : nullptr, B.stride, B.ld,
              Y.p, Y.stride, Y.ld};
  const int64_t total = batch * m * (int64_t)n;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(total)), dim3(NT), 0, s, g, total);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_contact_sigma(hipStream_t s, int64_t batch, int n, const Mat& S, const double* f,
                      const Mat& L, const Mat& G) {
  if (batch <= 0 || n <= 0) return;
  QX_REQUIRE(!S.zero() && !L.zero() && !G.zero() && f, "contact_sigma: null argument");
  QX_REQUIRE(S.ld == n && L.ld == n && G.ld == n, "contact_sigma: matrices must be dense n x n");
  const int tn = (n + TS - 1) / TS;
  const int64_t blocks = batch * tn * tn;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "contact_sigma: grid too large");
  hipLaunchKernelGGL(contact_sigma_kernel, dim3((unsigned)blocks), dim3(NT), 0, s, S.p, S.stride,
                     n, f, L.p, L.stride, G.p, G.stride, tn);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_diag_imag(hipStream_t s, int64_t batch, int n, const Mat& X, double* out,
                  int64_t out_stride, double scale) {
  if (batch <= 0 || n <= 0) return;
  QX_REQUIRE(!X.zero() && out, "diag_imag: null argument");
  const int64_t total = batch * n;
  hipLaunchKernelGGL(diag_imag_kernel, dim3(grid_for(total)), dim3(NT), 0, s, X.p, X.stride, X.ld,
                     n, out, out_stride, total, scale);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_assemble(hipStream_t s, int64_t batch, int64_t nnz, const z2* z, const z2* S,
                 const z2* T, const z2* Sig, int64_t sSig, z2* A, int64_t sA) {
  if (batch <= 0 || nnz <= 0) return;
  QX_REQUIRE(A && (z || !S), "assemble: null argument");
  const int64_t total = batch * nnz;
  hipLaunchKernelGGL(assemble_kernel, dim3(grid_for(total)), dim3(NT), 0, s, z, S, T, Sig, sSig, A,
                     sA, nnz, total);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_ident_sub(hipStream_t s, int64_t batch, int n, const Mat& Y) {
  if (batch <= 0 || n <= 0) return;
  QX_REQUIRE(!Y.zero(), "ident_sub: null matrix");
  const int64_t total = batch * n * (int64_t)n;
  hipLaunchKernelGGL(ident_sub_kernel, dim3(grid_for(total)), dim3(NT), 0, s, Y.p, Y.stride, Y.ld,
                     n, total);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_antiherm(hipStream_t s, int64_t batch, int n, double scale,
                 const Mat& X, const Mat& Y) {
  QX_REQUIRE(X.p != Y.p, "antiherm: in-place use is not supported");
  herm(s, batch, n, n, 0.0, scale, X, -scale, X, Y);
}

void ew_antiherm_pair(hipStream_t s, int64_t batch, int n, double scale, const Mat& X0,
                      const Mat& Y0, const Mat& X1, const Mat& Y1) {
  QX_REQUIRE(X0.p != Y0.p && X1.p != Y1.p, "antiherm: in-place use is not supported");
  if (Y0.p == Y1.p || Y0.p == X1.p || Y1.p == X0.p) {
    // one group's output is the other's operand: in order
    ew_antiherm(s, batch, n, scale, X0, Y0);
    ew_antiherm(s, batch, n, scale, X1, Y1);
    return;
  }
  herm(s, batch, n, n, 0.0, scale, X0, -scale, X0, Y0, 0.0, Mat(), &X1, &X1, &Y1);
}

void ew_herm_add(hipStream_t s, int64_t batch, int n, const Mat& D, double scale,
                 const Mat& X, const Mat& Y) {
  QX_REQUIRE(X.p != Y.p, "herm_add: Y must not alias X");
  herm(s, batch, n, n, 0.0, scale, X, -scale, X, Y, 1.0, D);
}

void ew_add_hdiff(hipStream_t s, int64_t batch, int n, const Mat& T,
                  const Mat& Y) {
  QX_REQUIRE(T.p != Y.p, "add_hdiff: in-place use is not supported");
  herm(s, batch, n, n, 1.0, -1.0, T, 1.0, T, Y);
}

void ew_neg_h(hipStream_t s, int64_t batch, int m, int n, const Mat& X,
              const Mat& Y) {
  // X is m x n, Y = -X^H is n x m.
  QX_REQUIRE(X.p != Y.p, "neg_h: in-place use is not supported");
  herm(s, batch, n, m, 0.0, 0.0, Mat(), -1.0, X, Y);
}

void trace_diff(hipStream_t s, int64_t batch, int n, const Mat& X, const Mat& Y,
                const Mat& Z, const Mat& W, z2* out, int64_t out_stride) {
  if (batch <= 0) return;
  TraceArgs g{n, X.p, Y.p, Z.p, W.p, X.stride, Y.stride, Z.stride, W.stride,
              X.ld, Y.ld, Z.ld, W.ld, out, out_stride};
  if (X.zero() || Y.zero()) g.X = nullptr;
  if (Z.zero() || W.zero()) g.Z = nullptr;
  hipLaunchKernelGGL(trace_kernel, dim3((unsigned)batch), dim3(NT), 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
}

void absnorm2(hipStream_t s, int64_t batch, int n, const Mat& A, const Mat& B,
              double* out) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(absnorm_kernel, dim3((unsigned)batch), dim3(NT), 0, s, n, A.p,
                     A.stride, A.ld, B.p, B.stride, B.ld, out);
  QX_HIP_CHECK(hipGetLastError());
}

void ew_ns_check(hipStream_t s, int64_t batch, int n, const Mat& R, double tol, int* flagged,
                 int* info) {
  if (batch <= 0) return;
  QX_REQUIRE(batch < (int64_t(1) << 31), "ns_check: batch too large");
  hipLaunchKernelGGL(ns_check_kernel, dim3((unsigned)batch), dim3(NT), 0, s, n, R.p, R.stride,
                     R.ld, tol, flagged, info);
  QX_HIP_CHECK(hipGetLastError());
}

void contour_operator(hipStream_t s, int64_t batch, int nq, int b,
                      const Mat& aji, const Mat& aii, const Mat& aij,
                      const z2* z, const Mat& T) {
  ContourArgs g{nq, b, aji.p, aii.p, aij.p, aji.stride, aii.stride, aij.stride,
                aji.ld, aii.ld, aij.ld, z, T.p, T.stride, T.ld};
  const int64_t total = batch * nq * (int64_t)b * b;
  if (total <= 0) return;
  hipLaunchKernelGGL(contour_kernel, dim3(grid_for(total)), dim3(NT), 0, s, g, total);
  QX_HIP_CHECK(hipGetLastError());
}

void contour_accumulate(hipStream_t s, int64_t batch, int nq, int b,
                        const Mat& Tinv, const z2* w0, const z2* w1,
                        const Mat& q0, const Mat& q1) {
  AccArgs g{nq, b, Tinv.p, Tinv.stride, Tinv.ld, w0, w1, q0.p, q1.p, q0.stride, q0.ld};
  QX_REQUIRE(q0.stride == q1.stride && q0.ld == q1.ld, "contour: q0/q1 layouts differ");
  const int64_t total = batch * (int64_t)b * b;
  if (total <= 0) return;
  hipLaunchKernelGGL(contour_acc_kernel, dim3(grid_for(total)), dim3(NT), 0, s, g, total);
  QX_HIP_CHECK(hipGetLastError());
}

}  // namespace qx
