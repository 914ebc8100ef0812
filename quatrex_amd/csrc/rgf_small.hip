// Fused selected solve for small blocks (gfx950): one workgroup per energy.
//
// Same algorithm, outputs and operation order as rgf_selected_solve in
// rgf.hip (reference src/qttools/greens_function_solver/rgf.py:164-414), for
// stacks whose blocks are all at most kSB = 32 wide and at most kSN = 16 long.
// At such sizes every block product is a few microseconds of matrix-core
// work, so the general path (one launch per product over the whole stack)
// is bound by launch latency; here the whole sweep of one energy runs in one
// workgroup of two teams of four waves (one wave of each team per SIMD),
// with every intermediate block in a per-energy workspace that stays in L2.
//
// - Block products (sgemm): the four waves of a team each own a 16 x 16
//   quarter of the output and load their MFMA fragments straight from
//   memory, all k-steps issued before the first v_mfma_f64_16x16x4_f64.
//   4M or 3M split, summed over k in the same order as the batched kernels.
// - The two teams run independent products of one step side by side: the
//   lesser and greater chains of a layer, pairs of retarded products
//   (rgf_small_kernel describes the schedule).  The sweep is a chain of
//   dependent steps, each about a microsecond of one SIMD's matrix-core
//   work behind a memory round trip and a barrier, so fewer steps make it
//   shorter.
// - Block inverse (sinv_wg): the four waves of team 0, eight lanes per row
//   (four columns each), Gauss-Jordan with partial pivoting and one barrier
//   per column: the pivot is a DPP wave maximum of 32-bit keys (float
//   magnitude with the row index in the low bits, as in inv.hip) over each
//   wave's rows, then the largest of the four published keys; the candidate
//   rows and row k go through double-buffered LDS; the column interchanges
//   are undone at the end.  (sinv_wave is the one-wave variant, kept for
//   QX_INV_SMALL=1.)
// - Elementwise steps (folds, X - X^H parts, traces) as team loops whose
//   loads are all issued before the first use.
// - A barrier closes every step, which makes its global writes visible to
//   the other waves of the workgroup.
#include <cstdlib>

#include "qx_internal.h"

namespace qx {
namespace {

// (the workspace bound of rgf_workspace_bytes in rgf.hip follows these)
constexpr int kSB = 32;   // largest block size
constexpr int kSN = 16;   // largest block count
constexpr int kThreads = 256;       // a team: four waves, one per SIMD
constexpr int kSolveThreads = 512;  // the fused kernels: two teams

typedef double d4 __attribute__((ext_vector_type(4)));

struct SLay {
  int32_t diag[kSN], up[kSN], lo[kSN];  // element offsets, -1: not stored
};

struct SmallArgs {
  int N, flags, NS, split;
  int bs[kSN];
  int orb[kSN];  // first orbital of each block (density row offsets)
  int64_t n_orb;
  int64_t sA, sSL, sSG, sXL, sXG, sXR;  // stack strides (elements)
  const z2 *A, *SL, *SG;
  z2 *XL, *XG, *XR;
  SLay lA, lSL, lSG, lXL, lXG, lXR;
  const z2* obc[3][kSN];  // retarded, lesser, greater boundary blocks or null
  z2* ws;                 // per energy NS slots of kSB * kSB
  int* info;
  z2* current;   // (S, N - 1) or null
  z2* corner;    // (S, b_0, b_{N-1}) or null
  double* density;  // (S, n_orb) or null
};

struct M {
  z2* p;
  int ld;
};

__device__ inline M blk(const z2* base, int64_t stride, int64_t e, int32_t off, int ld) {
  if (!base || off < 0) return M{nullptr, ld};
  return M{const_cast<z2*>(base) + e * stride + off, ld};
}

// Wave maximum of pivot keys (DPP), as in inv.hip.
__device__ inline unsigned s_wave_max(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xf, 0xf, false));
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
  const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
  const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
  return max(max(a, b), max(c, d));
}

// Value of lane QQ of each pair of lanes (DPP quad_perm [0,0,2,2] / [1,1,3,3]).
template <int QQ>
__device__ inline double s_pair_bcast(double x) {
  constexpr int sel = QQ == 0 ? 0xA0 : 0xF5;
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), sel, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), sel, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// LDS of one wave's inverse.
struct WaveShared {
  z2 prow[kSB], krow[kSB];
  int perm[kSB], idx[kSB];
  z2 tile[kSB][kSB + 1];
};

struct WgInvShared {
  z2 cand[2][kThreads / 64][kSB];
  z2 krow[2][kSB];
  unsigned keys[2][kThreads / 64];
  int perm[kSB], idx[kSB];
  z2 tile[kSB][kSB + 1];
};

struct Shared {
  WgInvShared winv;
  double red[2][kSolveThreads / 64];
};

// C = alpha op(A) op(B) + beta Cin, m, n, k <= 32, by one team of four waves
// (tt: the thread's index in the team), each wave owning a 16 x 16 quarter
// of C; a null A or B is zero.  Every load (the fragments of all k-steps and
// the Cin values) is issued before the first v_mfma_f64_16x16x4_f64, from
// clamped addresses, with no branch between the loads.  A lane whose row of
// A (column of B) lies outside the block feeds only output rows (columns)
// that are not stored, so its value is left as loaded; only a partial last
// k-step (k not a multiple of 4) has its fragments beyond k zeroed.  No
// barrier.
__device__ __forceinline__ void sgemm(int tt, int split, int m, int n, int k, int opa, int opb,
                                      Cplx al, M A, M B, Cplx be, M Cin, M C) {
  const int w = tt >> 6, lane = tt & 63;
  const int r0 = (w >> 1) * 16, c0 = (w & 1) * 16;
  const int r16 = lane & 15, k4 = lane >> 4;
  if (r0 >= m || c0 >= n) return;
  const bool use_b = Cin.p && (be.re != 0.0 || be.im != 0.0);
  z2 cin[4];
  if (use_b) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r0 + k4 + 4 * q, col = c0 + r16;
      const bool ok = row < m && col < n;
      cin[q] = Cin.p[ok ? row * Cin.ld + col : 0];
    }
  }
  d4 a1 = {0, 0, 0, 0}, a2 = {0, 0, 0, 0}, a3 = {0, 0, 0, 0};
  if (A.p && B.p) {
    z2 fa[kSB / 4], fb[kSB / 4];
#pragma unroll
    for (int s = 0; s < kSB / 4; ++s) {
      const int kk = 4 * s + k4;
      const bool oka = kk < k && r0 + r16 < m, okb = kk < k && c0 + r16 < n;
      const int ia = opa == QX_OP_N ? (r0 + r16) * A.ld + kk : kk * A.ld + r0 + r16;
      const int ib = opb == QX_OP_N ? kk * B.ld + c0 + r16 : (c0 + r16) * B.ld + kk;
      fa[s] = A.p[oka ? ia : 0];
      fb[s] = B.p[okb ? ib : 0];
    }
#pragma unroll
    for (int s = 0; s < kSB / 4; ++s) {
      if (opa != QX_OP_N) fa[s].y = -fa[s].y;
      if (opb != QX_OP_N) fb[s].y = -fb[s].y;
    }
    if (k & 3) {
      const int s = k >> 2;  // the partial k-step
#pragma unroll
      for (int u = 0; u < kSB / 4; ++u) {
        if (u == s && 4 * u + k4 >= k) {
          fa[u] = make_double2(0.0, 0.0);
          fb[u] = make_double2(0.0, 0.0);
        }
      }
    }
#pragma unroll
    for (int s = 0; s < kSB / 4; ++s) {
      if (4 * s < k) {
        if (split == 3) {
          a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].x, fb[s].x, a1, 0, 0, 0);
          a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].y, fb[s].y, a2, 0, 0, 0);
          a3 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].x + fa[s].y, fb[s].x + fb[s].y, a3,
                                                    0, 0, 0);
        } else {
          a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].x, fb[s].x, a1, 0, 0, 0);
          a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].x, fb[s].y, a2, 0, 0, 0);
          a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].y, fb[s].y, a1, 0, 0, 1);
          a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s].y, fb[s].x, a2, 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = r0 + k4 + 4 * q, col = c0 + r16;
    if (row < m && col < n) {
      double vr, vi;
      if (split == 3) {
        vr = a1[q] - a2[q];
        vi = a3[q] - a1[q] - a2[q];
      } else {
        vr = a1[q];
        vi = a2[q];
      }
      double outr = al.re * vr - al.im * vi, outi = al.re * vi + al.im * vr;
      if (use_b) {
        const z2 c = cin[q];
        outr += be.re * c.x - be.im * c.y;
        outi += be.re * c.y + be.im * c.x;
      }
      C.p[row * C.ld + col] = make_double2(outr, outi);
    }
  }
}

// The elementwise steps below run over the threads tt = 0 .. nt - 1 of a
// team (nt a multiple of 64); each thread loads the four elements it owns
// (element t + u nt, u = 0..3, addresses clamped into the block) before it
// uses any of them.  A null operand reads from Y and is not used.  No barrier.

// Y = a A + b B (null operands are zero).
__device__ __forceinline__ void saxpby(int tt, int nt, int m, int n, Cplx a, M A, Cplx b, M B,
                                       M Y) {
  const int mn = m * n;
  const z2* pa = A.p ? A.p : Y.p;
  const z2* pb = B.p ? B.p : Y.p;
  const int la = A.p ? A.ld : Y.ld, lb = B.p ? B.ld : Y.ld;
  for (int t0 = tt; t0 < mn; t0 += 4 * nt) {
    z2 va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = min(t0 + u * nt, mn - 1), r = t / n, c = t - r * n;
      va[u] = pa[r * la + c];
      vb[u] = pb[r * lb + c];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = t0 + u * nt;
      if (t < mn) {
        const int r = t / n, c = t - r * n;
        double yr = 0.0, yi = 0.0;
        if (A.p) {
          yr += a.re * va[u].x - a.im * va[u].y;
          yi += a.re * va[u].y + a.im * va[u].x;
        }
        if (B.p) {
          yr += b.re * vb[u].x - b.im * vb[u].y;
          yi += b.re * vb[u].y + b.im * vb[u].x;
        }
        Y.p[r * Y.ld + c] = make_double2(yr, yi);
      }
    }
  }
}

// Y = cy Y + ca A + cb B^H + cc D (real coefficients; Y, m x n, aliases none
// of A, B, D; B is n x m).
__device__ __forceinline__ void sherm(int tt, int nt, int m, int n, double cy, double ca, M A,
                                      double cb, M B, M Y, double cc = 0.0,
                                      M D = M{nullptr, 0}) {
  const int mn = m * n;
  const bool uy = cy != 0.0, ua = ca != 0.0 && A.p, ub = cb != 0.0 && B.p, ud = cc != 0.0 && D.p;
  const z2* pa = ua ? A.p : Y.p;
  const z2* pb = ub ? B.p : Y.p;
  const z2* pd = ud ? D.p : Y.p;
  const int la = ua ? A.ld : Y.ld, ld = ud ? D.ld : Y.ld;
  for (int t0 = tt; t0 < mn; t0 += 4 * nt) {
    z2 vy[4], va[4], vb[4], vd[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = min(t0 + u * nt, mn - 1), r = t / n, c = t - r * n;
      vy[u] = Y.p[r * Y.ld + c];
      va[u] = pa[r * la + c];
      vb[u] = pb[ub ? c * B.ld + r : r * Y.ld + c];
      vd[u] = pd[r * ld + c];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = t0 + u * nt;
      if (t < mn) {
        const int r = t / n, c = t - r * n;
        double yr = 0.0, yi = 0.0;
        if (uy) {
          yr = cy * vy[u].x;
          yi = cy * vy[u].y;
        }
        if (ua) {
          yr += ca * va[u].x;
          yi += ca * va[u].y;
        }
        if (ub) {
          yr += cb * vb[u].x;
          yi -= cb * vb[u].y;
        }
        if (ud) {
          yr += cc * vd[u].x;
          yi += cc * vd[u].y;
        }
        Y.p[r * Y.ld + c] = make_double2(yr, yi);
      }
    }
  }
}

__device__ __forceinline__ void sdiag_imag(int tt, int nt, int n, M X, double* out) {
  for (int t = tt; t < n; t += nt) out[t] = X.p[t * X.ld + t].y;
}

// out = Tr(X Y) - Tr(Z W), n x n, by the whole workgroup of nt threads
// (X, Y, Z, W all stored); two barriers.
__device__ __forceinline__ void strace_diff(int nt, int n, M X, M Y, M Z, M W, z2* out,
                                            Shared& sh) {
  double sr = 0.0, si = 0.0;
  for (int t0 = threadIdx.x; t0 < n * n; t0 += 2 * nt) {
    z2 x[2], y[2], z[2], w[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = min(t0 + u * nt, n * n - 1), a = t / n, b = t - a * n;
      x[u] = X.p[a * X.ld + b];
      y[u] = Y.p[b * Y.ld + a];
      z[u] = Z.p[a * Z.ld + b];
      w[u] = W.p[b * W.ld + a];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (t0 + u * nt < n * n) {
        sr += x[u].x * y[u].x - x[u].y * y[u].y;
        si += x[u].x * y[u].y + x[u].y * y[u].x;
        sr -= z[u].x * w[u].x - z[u].y * w[u].y;
        si -= z[u].x * w[u].y + z[u].y * w[u].x;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh.red[0][threadIdx.x >> 6] = sr;
    sh.red[1][threadIdx.x >> 6] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, ti = 0.0;
    for (int u = 0; u < nt / 64; ++u) {
      tr += sh.red[0][u];
      ti += sh.red[1][u];
    }
    *out = make_double2(tr, ti);
  }
  __syncthreads();
}

// X = inv(A - S) for one n x n matrix (n <= 32; S may be null; X may
// alias A), computed by one whole wave with no workgroup barrier.
__device__ __forceinline__ void sinv_wave(int n, const z2* A, int lda, const z2* S, int lds, z2* X, int ldx,
                          int* info, WaveShared& sh) {
  {
    const int lane = threadIdx.x & 63, r = lane >> 1, h = lane & 1;
    double xr[16], xi[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int col = 16 * h + c;
      z2 v = make_double2(0.0, 0.0);
      if (r < n && col < n) {
        v = A[r * lda + col];
        if (S) {
          const z2 u = S[r * lds + col];
          v.x -= u.x;
          v.y -= u.y;
        }
      }
      xr[c] = v.x;
      xi[c] = v.y;
    }
#pragma unroll
    for (int k = 0; k < kSB; ++k) {
      if (k < n) {
        const int hk = k >> 4, ck = k & 15;
        unsigned key = 0;
        if (h == hk && r >= k && r < n) key = pivot_key(xr[ck], xi[ck], r);
        const unsigned kmax = s_wave_max(key);
        int p = 511 - (int)(kmax & 0x1FFu);
        if (p < k || p >= n) p = k;
        if (lane == 0) {
          sh.perm[k] = p;
          if ((kmax >> 9) == 0u) *info = 1;  // no nonzero pivot: singular
        }
        // publish the pivot row and row k (one wave: LDS keeps program order)
        if (r == p) {
#pragma unroll
          for (int c = 0; c < 16; ++c) sh.prow[16 * h + c] = make_double2(xr[c], xi[c]);
        }
        if (r == k) {
#pragma unroll
          for (int c = 0; c < 16; ++c) sh.krow[16 * h + c] = make_double2(xr[c], xi[c]);
        }
        __builtin_amdgcn_wave_barrier();
        const z2 x = sh.prow[k];
        double vr, vi;
        complex_recip(x.x, x.y, vr, vi);
        double pr[16], pi[16];  // this lane's half of the scaled pivot row
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          const int col = 16 * h + c;
          if (col == k) {
            pr[c] = vr;
            pi[c] = vi;
          } else {
            const z2 v = sh.prow[col];
            pr[c] = v.x * vr - v.y * vi;
            pi[c] = v.x * vi + v.y * vr;
          }
        }
        if (r == k) {
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            xr[c] = pr[c];
            xi[c] = pi[c];
          }
        } else if (r == p) {
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            const z2 v = sh.krow[16 * h + c];
            xr[c] = v.x;
            xi[c] = v.y;
          }
        }
        // multiplier: the row's (swapped-in) value in column k
        const double fr = hk == 0 ? s_pair_bcast<0>(xr[ck]) : s_pair_bcast<1>(xr[ck]);
        const double fi = hk == 0 ? s_pair_bcast<0>(xi[ck]) : s_pair_bcast<1>(xi[ck]);
        if (r != k && r < n) {
#pragma unroll
          for (int c = 0; c < 16; ++c) {
            const int col = 16 * h + c;
            if (col == k) {
              xr[c] = -(fr * pr[c] - fi * pi[c]);
              xi[c] = -(fr * pi[c] + fi * pr[c]);
            } else {
              xr[c] -= fr * pr[c] - fi * pi[c];
              xi[c] -= fr * pi[c] + fi * pr[c];
            }
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    // undo the column interchanges: reverse order of the pivots
    if (lane == 0) {
      for (int c = 0; c < n; ++c) sh.idx[c] = c;
      for (int k = n - 1; k >= 0; --k) {
        const int pq = sh.perm[k];
        const int t = sh.idx[k];
        sh.idx[k] = sh.idx[pq];
        sh.idx[pq] = t;
      }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int col = 16 * h + c;
      if (r < n && col < n) sh.tile[r][col] = make_double2(xr[c], xi[c]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int col = 16 * h + c;
      if (r < n && col < n) X[r * ldx + col] = sh.tile[r][sh.idx[col]];
    }
  }
}

// X = inv(A - S) for one n x n matrix (n <= 32) by the first four waves of
// the workgroup (eight lanes per row, four columns per lane; further waves
// hold no row): Gauss-Jordan with partial pivoting and one barrier per
// column.  Before the barrier each wave publishes its best candidate row
// (DPP maximum of the pivot keys over its eight rows) and the lanes of row k
// publish row k; after it every thread takes the largest of the four keys,
// reads the pivot row and computes the reciprocal itself (one wave per SIMD,
// so nothing is serialised).  The buffers alternate between columns: no
// wave can write a buffer before every wave has passed the barrier that
// follows its last read.


__device__ __forceinline__ void sinv_wg(int n, const z2* A, int lda, const z2* S, int lds,
                                        z2* X, int ldx, int* info, WgInvShared& sh) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = tid >> 3, q = tid & 7;  // row, quarter-of-eight within the row
  double xr[4], xi[4];
  {
    // all loads in flight together (clamped addresses, no branch between them)
    z2 va[4], vs[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = 4 * q + c;
      const bool ok = r < n && col < n;
      va[c] = A[ok ? r * lda + col : 0];
      vs[c] = S ? S[ok ? r * lds + col : 0] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bool ok = r < n && 4 * q + c < n;
      xr[c] = ok ? va[c].x - vs[c].x : 0.0;
      xi[c] = ok ? va[c].y - vs[c].y : 0.0;
    }
  }
  // Groups of four columns run as a loop, the four columns of a group are
  // unrolled (register indices stay constants): the fully unrolled 32
  // columns are ~64 KB of code per call site, beyond the instruction cache.
#pragma unroll 1
  for (int kq = 0; kq < kSB / 4; ++kq) {
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
      const int k = 4 * kq + kc, b = kc & 1;
      if (k < n) {
        unsigned key = 0;
        if (q == kq && r >= k && r < n) key = pivot_key(xr[kc], xi[kc], r);
        const unsigned wkey = s_wave_max(key);
        if (wkey != 0u && r == 511 - (int)(wkey & 0x1FFu)) {
#pragma unroll
          for (int c = 0; c < 4; ++c) sh.cand[b][w][4 * q + c] = make_double2(xr[c], xi[c]);
        }
        if (lane == 0 && w < kThreads / 64) sh.keys[b][w] = wkey;  // waves 4..7 hold no row
        if (r == k) {
#pragma unroll
          for (int c = 0; c < 4; ++c) sh.krow[b][4 * q + c] = make_double2(xr[c], xi[c]);
        }
        __syncthreads();
        unsigned kmax = sh.keys[b][0];
        int ws = 0;
#pragma unroll
        for (int u = 1; u < kThreads / 64; ++u) {
          const unsigned v = sh.keys[b][u];
          if (v > kmax) {
            kmax = v;
            ws = u;
          }
        }
        int p = 511 - (int)(kmax & 0x1FFu);
        if (p < k || p >= n) p = k;
        if (tid == 0) {
          sh.perm[k] = p;
          if ((kmax >> 9) == 0u) *info = 1;  // no nonzero pivot: singular
        }
        const z2 x = sh.cand[b][ws][k];
        double vr, vi;
        complex_recip(x.x, x.y, vr, vi);
        double pr[4], pi[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int col = 4 * q + c;
          if (col == k) {
            pr[c] = vr;
            pi[c] = vi;
          } else {
            const z2 v = sh.cand[b][ws][col];
            pr[c] = v.x * vr - v.y * vi;
            pi[c] = v.x * vi + v.y * vr;
          }
        }
        if (r == k) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            xr[c] = pr[c];
            xi[c] = pi[c];
          }
        } else if (r == p) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const z2 v = sh.krow[b][4 * q + c];
            xr[c] = v.x;
            xi[c] = v.y;
          }
        }
        // multiplier: the row's (swapped-in) value in column k, from the lane
        // of this row that holds column k
        const int src = (lane & ~7) + kq;
        const double fr = __shfl(xr[kc], src, 64), fi = __shfl(xi[kc], src, 64);
        if (r < n && r != k) {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int col = 4 * q + c;
            if (col == k) {
              xr[c] = -(fr * pr[c] - fi * pi[c]);
              xi[c] = -(fr * pi[c] + fi * pr[c]);
            } else {
              xr[c] -= fr * pr[c] - fi * pi[c];
              xi[c] -= fr * pi[c] + fi * pr[c];
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int col = 4 * q + c;
    if (r < n && col < n) sh.tile[r][col] = make_double2(xr[c], xi[c]);
  }
  __syncthreads();  // tile, and the pivot record written by thread 0
  // Undo the column interchanges (the identity after the swaps (k, perm k)
  // for k = n - 1 down to 0): position c is traced forward through the
  // swaps for k = 0 .. n - 1, one position per thread.
  if (tid < n) {
    int pos = tid;
#pragma unroll 8
    for (int k = 0; k < n; ++k) {
      const int p = sh.perm[k];
      pos = (pos == k) ? p : ((pos == p) ? k : pos);
    }
    sh.idx[tid] = pos;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int col = 4 * q + c;
    if (r < n && col < n) X[r * ldx + col] = sh.tile[r][sh.idx[col]];
  }
  __syncthreads();
}



// In-place inverse of the block X by the workgroup.
__device__ __forceinline__ void sinv(int n, M X, int* info, Shared& sh) {
  sinv_wg(n, X.p, X.ld, nullptr, 0, X.p, X.ld, info, sh.winv);
}

// Batched X_e = inv(A_e - S_e), n <= 32: one matrix per workgroup.
__global__ __launch_bounds__(kThreads) void inv_small_wg_kernel(int n, const z2* A, int64_t sA,
                                                                int lda, const z2* S, int64_t sS,
                                                                int lds, z2* X, int64_t sX,
                                                                int ldx, int* info,
                                                                const int* mask) {
  __shared__ WgInvShared sh;
  const int64_t e = blockIdx.x;
  if (mask && !mask[e]) return;  // uniform over the workgroup
  sinv_wg(n, A + e * sA, lda, S ? S + e * sS : nullptr, lds, X + e * sX, ldx, info + e, sh);
}

// Batched X_e = inv(A_e - S_e), n <= 32: one matrix per wave, four per
// workgroup (the whole-wave condition e < batch is uniform per wave).
__global__ __launch_bounds__(kThreads) void inv_small_kernel(int64_t batch, int n, const z2* A,
                                                             int64_t sA, int lda, const z2* S,
                                                             int64_t sS, int lds, z2* X,
                                                             int64_t sX, int ldx, int* info,
                                                             const int* mask) {
  __shared__ WaveShared sh[kThreads / 64];
  const int w = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * (kThreads / 64) + w;
  if (e < batch && !(mask && !mask[e]))
    sinv_wave(n, A + e * sA, lda, S ? S + e * sS : nullptr, lds, X + e * sX, ldx, info + e,
              sh[w]);
}

const Cplx kOneS{1.0, 0.0}, kMinusS{-1.0, 0.0}, kZeroS{0.0, 0.0};

// SP: the complex product split (4 = 4M, 3 = 3M), a template parameter so
// that each inlined block product carries only its own MFMA sequence.
//
// Two teams of four waves share the workgroup.  After the retarded block of
// a layer is known, the lesser (q = 1, team 0) and greater (q = 2, team 1)
// chains have the same shape and run side by side; the retarded products
// that do not depend on each other run one per team (b1 | b2, b3 | b4), and
// the products of the lesser and greater chains that do not need the new
// retarded block (a_ji x^≶_ii, then its product with a_ji^H) run on team 1
// while team 0 forms the retarded block.  Every step ends in one workgroup
// barrier, reached by both teams: a team with nothing to do in a step only
// waits.  Within a step a team may run several products and elementwise
// updates that do not read each other's results.
template <int SP>
__global__ __launch_bounds__(kSolveThreads) void rgf_small_kernel(SmallArgs g) {
  __shared__ Shared sh;
  const int64_t e = blockIdx.x;
  const int N = g.N;
  constexpr int sp = SP;
  const bool want_r = (g.flags & QX_SOLVE_RETARDED) != 0;
  const bool want_r_off = (g.flags & QX_SOLVE_RETARDED_OFFDIAG) != 0;
  const bool want_cur = (g.flags & QX_SOLVE_CURRENT) != 0 && g.current;
  const bool skip_lower[2] = {(g.flags & QX_SOLVE_SKIP_LOWER_L) != 0,
                              (g.flags & QX_SOLVE_SKIP_LOWER_G) != 0};
  int* info = g.info + e;
  z2* W = g.ws + e * (int64_t)g.NS * kSB * kSB;
  const int team = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 8));
  const int tt = threadIdx.x & (kThreads - 1);  // thread index in the team
  const int tq = 1 + team;                     // the team's chain: 1 lesser, 2 greater
  auto sync = [] { __syncthreads(); };
  auto slot = [&](int s, int ld) { return M{W + (int64_t)s * kSB * kSB, ld}; };
  auto bs = [&](int i) { return g.bs[i]; };
  // slots: forward x^R, x^<, x^> (0..3N-1), full G^R, G^<, G^> from the
  // backward sweep (3N..6N-1), then temporaries (per chain where both
  // chains need one at the same time)
  auto F = [&](int q, int i) { return q * N + i; };  // q: 0 r, 1 l, 2 g
  auto full = [&](int q, int i) { return i == N - 1 ? q * N + i : (3 + q) * N + i; };
  const int R0 = 6 * N, T1 = R0;
  auto T3 = [&](int q) { return R0 + q; };
  auto T4 = [&](int q) { return R0 + 2 + q; };
  auto T5 = [&](int q) { return R0 + 4 + q; };
  auto T6 = [&](int q) { return R0 + 6 + q; };
  auto TB = [&](int q) { return R0 + 8 + q; };
  auto C4 = [&](int q) { return R0 + 10 + q; };
  auto CT = [&](int q) { return R0 + 12 + q; };
  const int B1 = R0 + 15, B2 = R0 + 16, B3 = R0 + 17, B4 = R0 + 18, C1 = R0 + 19,
            GC0 = R0 + 20, GC1 = R0 + 21, GCT = R0 + 22;  // 6 N + 23 slots
  auto Ad = [&](int i) { return blk(g.A, g.sA, e, g.lA.diag[i], bs(i)); };
  auto Au = [&](int i) { return blk(g.A, g.sA, e, g.lA.up[i], bs(i + 1)); };   // (i, i+1)
  auto Al = [&](int i) { return blk(g.A, g.sA, e, g.lA.lo[i], bs(i)); };       // (i+1, i)
  auto Sd = [&](int q, int i) {
    return q == 1 ? blk(g.SL, g.sSL, e, g.lSL.diag[i], bs(i)) : blk(g.SG, g.sSG, e, g.lSG.diag[i], bs(i));
  };
  auto Su = [&](int q, int i) {
    return q == 1 ? blk(g.SL, g.sSL, e, g.lSL.up[i], bs(i + 1))
                  : blk(g.SG, g.sSG, e, g.lSG.up[i], bs(i + 1));
  };
  auto Od = [&](int q, int i) {
    return q == 1 ? blk(g.XL, g.sXL, e, g.lXL.diag[i], bs(i)) : blk(g.XG, g.sXG, e, g.lXG.diag[i], bs(i));
  };
  auto Ou = [&](int q, int i) {
    return q == 1 ? blk(g.XL, g.sXL, e, g.lXL.up[i], bs(i + 1)) : blk(g.XG, g.sXG, e, g.lXG.up[i], bs(i + 1));
  };
  auto Ol = [&](int q, int i) {
    return q == 1 ? blk(g.XL, g.sXL, e, g.lXL.lo[i], bs(i)) : blk(g.XG, g.sXG, e, g.lXG.lo[i], bs(i));
  };
  auto obc = [&](int q, int i) {
    const z2* p = g.obc[q][i];
    return M{p ? const_cast<z2*>(p) + e * (int64_t)bs(i) * bs(i) : nullptr, bs(i)};
  };
  const M none{nullptr, 0};
  const int b0 = bs(0), bl = bs(N - 1);
  int gc_cur = GC0, gc_nxt = GC1;

  // ---- block 0 (rgf.py:207-231) ----
  {
    const M X0 = slot(F(0, 0), b0);
    saxpby(threadIdx.x, kSolveThreads, b0, b0, kOneS, Ad(0), kMinusS, obc(0, 0), X0);
    sync();
    sinv(b0, X0, info, sh);
    const M t3 = slot(T3(tq), b0), t6 = slot(T6(tq), b0);
    saxpby(tt, kThreads, b0, b0, kOneS, Sd(tq, 0), kOneS, obc(tq, 0), t3);
    if (g.corner && team == 0) saxpby(tt, kThreads, b0, b0, kOneS, X0, kZeroS, none, slot(gc_cur, kSB));
    sync();
    sgemm(tt, sp, b0, b0, b0, QX_OP_N, QX_OP_N, kOneS, X0, t3, kZeroS, none, t6);
    sync();
    sgemm(tt, sp, b0, b0, b0, QX_OP_N, QX_OP_C, kOneS, t6, X0, kZeroS, none, slot(F(tq, 0), b0));
    sync();
  }

  // ---- forward sweep (rgf.py:234-295) ----
  for (int i = 0; i + 1 < N; ++i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const M Xri = slot(F(0, i), bi), Xrj = slot(F(0, j), bj), t1 = slot(T1, bi);
    const bool ob0 = obc(0, j).p != nullptr;
    // team 0: a_ji x^R_ii, then x^R_jj's argument a_jj - a_ji x^R_ii a_ij;
    // team 1 meanwhile: a_ji x^<_ii, then a_ji x^>_ii (and the boundary folds)
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
      if (team == 0) {
        if (s == 0) {
          sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kOneS, Al(i), Xri, kZeroS, none, t1);
          if (ob0) saxpby(tt, kThreads, bj, bj, kOneS, Ad(j), kMinusS, obc(0, j), Xrj);
        } else {
          sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_N, kMinusS, t1, Au(i), kOneS,
                ob0 ? Xrj : Ad(j), Xrj);
        }
      } else {
        const int q = 1 + s;
        sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kOneS, Al(i), slot(F(q, i), bi), kZeroS, none,
              slot(T4(q), bi));
        if (obc(q, j).p)
          saxpby(tt, kThreads, bj, bj, kOneS, Sd(q, j), kOneS, obc(q, j), slot(T3(q), bj));
      }
      sync();
    }
    // each chain: t3 = a_ji x_ii a_ji^H + s_jj (+ boundary); team 0 also the
    // corner's g_c a_ij
    const M t3 = slot(T3(tq), bj), t5 = slot(T5(tq), bj), t6 = slot(T6(tq), bj);
    sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_C, kOneS, slot(T4(tq), bi), Al(i), kOneS,
          obc(tq, j).p ? t3 : Sd(tq, j), t3);
    if (g.corner && team == 0)
      sgemm(tt, sp, b0, bj, bi, QX_OP_N, QX_OP_N, kOneS, slot(gc_cur, kSB), Au(i), kZeroS, none,
            slot(GCT, kSB));
    sync();
    if (Su(1, i).p || Su(2, i).p) {
      const M sij = Su(tq, i);
      if (sij.p) sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_N, kOneS, t1, sij, kZeroS, none, t5);
      sync();
      if (sij.p) sherm(tt, kThreads, bj, bj, 1.0, -1.0, t5, 1.0, t5, t3);  // t3 += t5^H - t5
      sync();
    }
    sinv(bj, Xrj, info, sh);
    sgemm(tt, sp, bj, bj, bj, QX_OP_N, QX_OP_N, kOneS, Xrj, t3, kZeroS, none, t6);
    sync();
    sgemm(tt, sp, bj, bj, bj, QX_OP_N, QX_OP_C, kOneS, t6, Xrj, kZeroS, none, slot(F(tq, j), bj));
    if (g.corner && team == 1)
      sgemm(tt, sp, b0, bj, bj, QX_OP_N, QX_OP_N, kMinusS, slot(GCT, kSB), Xrj, kZeroS, none,
            slot(gc_nxt, kSB));
    sync();
    if (g.corner) {
      const int t = gc_cur;
      gc_cur = gc_nxt;
      gc_nxt = t;
    }
  }

  // ---- last diagonal blocks (rgf.py:297-305), the corner ----
  {
    const int bn = bs(N - 1);
    const M Xq = slot(F(tq, N - 1), bn), od = Od(tq, N - 1);
    if (g.density && team == 0) sdiag_imag(tt, kThreads, bn, Xq, g.density + e * g.n_orb + g.orb[N - 1]);
    if (od.p) sherm(tt, kThreads, bn, bn, 0.0, 0.5, Xq, -0.5, Xq, od);
    if (want_r && team == 0) {
      const M xr = blk(g.XR, g.sXR, e, g.lXR.diag[N - 1], bn);
      if (xr.p) saxpby(tt, kThreads, bn, bn, kOneS, slot(F(0, N - 1), bn), kZeroS, none, xr);
    }
    if (g.corner && team == 1)
      saxpby(tt, kThreads, b0, bl, kOneS, slot(gc_cur, kSB), kZeroS, none,
             M{g.corner + e * (int64_t)b0 * bl, bl});
    sync();
  }

  // ---- backward sweep (rgf.py:308-404) ----
  for (int i = N - 2; i >= 0; --i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const M Xri = slot(F(0, i), bi), Xrj = slot(full(0, j), bj);
    const M Aij = Au(i), Aji = Al(i);
    const M b1 = slot(B1, bj), b2 = slot(B2, bi), b3 = slot(B3, bj), b4 = slot(B4, bi);
    if (team == 0)
      sgemm(tt, sp, bi, bj, bi, QX_OP_N, QX_OP_N, kOneS, Xri, Aij, kZeroS, none, b1);  // xr_ii a_ij
    else
      sgemm(tt, sp, bj, bi, bj, QX_OP_N, QX_OP_N, kOneS, Xrj, Aji, kZeroS, none, b2);  // xr_jj a_ji
    sync();
    if (team == 0)
      sgemm(tt, sp, bi, bj, bj, QX_OP_N, QX_OP_N, kOneS, b1, Xrj, kZeroS, none, b3);  // b1 xr_jj
    else
      sgemm(tt, sp, bi, bi, bj, QX_OP_N, QX_OP_N, kOneS, b1, b2, kZeroS, none, b4);  // b1 b2
    sync();

    const bool any_s = Su(1, i).p || Su(2, i).p;
    const M sij = Su(tq, i);
    if (want_cur) {
      // layer current from the forward blocks of layer i and the full blocks
      // of layer j (rgf.py:381-400): tl = a_ji x_ii a_ji^H (+ the Σ_ij terms)
      const M c1 = slot(C1, bi), t4 = slot(T4(tq), bi), tl = slot(CT(tq), bj),
              c4 = slot(C4(tq), bj);
      sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kOneS, Aji, slot(F(tq, i), bi), kZeroS, none, t4);
      if (team == 0)
        sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kOneS, Aji, Xri, kZeroS, none, c1);  // a_ji xr_ii
      sync();
      sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_C, kOneS, t4, Aji, kZeroS, none, tl);
      sync();
      if (any_s) {
        if (sij.p) sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_N, kOneS, c1, sij, kZeroS, none, c4);
        sync();
        if (sij.p) sherm(tt, kThreads, bj, bj, 1.0, -1.0, c4, 1.0, c4, tl);
        sync();
      }
      strace_diff(kSolveThreads, bj, slot(CT(2), bj), slot(full(1, j), bj), slot(full(2, j), bj),
                  slot(CT(1), bj), g.current + e * (int64_t)(N - 1) + i, sh);
    }

    // each chain (q = tq): G_ii, G_ij and the outputs
    const int q = tq;
    const M Xi = slot(F(q, i), bi), Xj = slot(full(q, j), bj);
    const M t1 = slot(TB(q), bi), t5 = slot(T5(q), bj), P = slot(T6(q), bj);
    const M Z = slot(full(q, i), bi), Zr = slot(full(0, i), bi);
    // temp_1x = B4 x_ii; team 0 also G^R_ii = xr_ii + B4 xr_ii
    sgemm(tt, sp, bi, bi, bi, QX_OP_N, QX_OP_N, kOneS, b4, Xi, kZeroS, none, t1);
    if (team == 0) sgemm(tt, sp, bi, bi, bi, QX_OP_N, QX_OP_N, kOneS, b4, Xri, kOneS, Xri, Zr);
    sync();
    if (any_s) {  // temp_1x -= xr_ii s_ij B3^H
      if (sij.p) sgemm(tt, sp, bi, bj, bi, QX_OP_N, QX_OP_N, kOneS, Xri, sij, kZeroS, none, t5);
      sync();
      if (sij.p) sgemm(tt, sp, bi, bi, bj, QX_OP_N, QX_OP_C, kMinusS, t5, b3, kOneS, t1, t1);
      sync();
    }
    // x_ii' = x_ii + (temp_1x - temp_1x^H); P = -temp_2x = -B1 x_jj
    sherm(tt, kThreads, bi, bi, 0.0, 1.0, t1, -1.0, t1, Z, 1.0, Xi);
    sgemm(tt, sp, bi, bj, bj, QX_OP_N, QX_OP_N, kMinusS, b1, Xj, kZeroS, none, P);
    sync();
    // x_ii' += temp_2x B1^H;  x_ij = -temp_2x - x_ii B2^H (+ xr_ii s_ij xr_jj^H)
    M Oij = Ou(q, i);
    if (!Oij.p) Oij = slot(T4(q), bj);
    sgemm(tt, sp, bi, bi, bj, QX_OP_N, QX_OP_C, kMinusS, P, b1, kOneS, Z, Z);
    sgemm(tt, sp, bi, bj, bi, QX_OP_N, QX_OP_C, kMinusS, Xi, b2, kOneS, P, Oij);
    sync();
    if (any_s) {
      if (sij.p) sgemm(tt, sp, bi, bj, bj, QX_OP_N, QX_OP_C, kOneS, t5, Xrj, kOneS, Oij, Oij);
      sync();
    }
    // outputs: x_ji = -x_ij^H, the diagonal block, the density; team 0 the
    // retarded diagonal block, team 1 its off-diagonal blocks on request
    const M low = Ol(q, i), od = Od(q, i);
    if (!skip_lower[q - 1] && low.p) sherm(tt, kThreads, bj, bi, 0.0, 0.0, none, -1.0, Oij, low);
    if (od.p) sherm(tt, kThreads, bi, bi, 0.0, 0.5, Z, -0.5, Z, od);
    if (g.density && q == 1) sdiag_imag(tt, kThreads, bi, Z, g.density + e * g.n_orb + g.orb[i]);
    if (want_r && team == 0) {
      const M xr = blk(g.XR, g.sXR, e, g.lXR.diag[i], bi);
      if (xr.p) saxpby(tt, kThreads, bi, bi, kOneS, Zr, kZeroS, none, xr);
    }
    if (want_r && want_r_off && team == 1) {
      const M up = blk(g.XR, g.sXR, e, g.lXR.up[i], bj), lo = blk(g.XR, g.sXR, e, g.lXR.lo[i], bi);
      if (up.p) saxpby(tt, kThreads, bi, bj, kMinusS, b3, kZeroS, none, up);
      if (lo.p) sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kMinusS, b2, Xri, kZeroS, none, lo);
    }
    sync();
  }
}

// Fused selected inversion (rgf.py:25-113 restated in rgf_selected_inv of
// rgf.hip), one workgroup per energy.
struct SmallInvArgs {
  int N, NS, split;
  int bs[kSN];
  int64_t sA, sX;
  const z2* A;
  z2* X;
  SLay lA, lX;
  const z2* obc[kSN];
  z2* ws;
  int* info;
};

// Two teams as in rgf_small_kernel: the forward sweep is one chain (team 1
// folds the boundary block while team 0 forms a_ji x_ii); in the backward
// sweep x_ji and x_ij run one per team.
template <int SP>
__global__ __launch_bounds__(kSolveThreads) void rgf_small_inv_kernel(SmallInvArgs g) {
  __shared__ Shared sh;
  const int64_t e = blockIdx.x;
  const int N = g.N;
  constexpr int sp = SP;
  int* info = g.info + e;
  z2* W = g.ws + e * (int64_t)g.NS * kSB * kSB;
  const int team = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 8));
  const int tt = threadIdx.x & (kThreads - 1);
  auto sync = [] { __syncthreads(); };
  auto slot = [&](int s, int ld) { return M{W + (int64_t)s * kSB * kSB, ld}; };
  auto bs = [&](int i) { return g.bs[i]; };
  // slots: forward x_ii (0..N-1), full X_ii from the backward sweep
  // (N..2N-1; block N-1 keeps its forward slot), temporaries
  auto F = [&](int i) { return i; };
  auto full = [&](int i) { return i == N - 1 ? i : N + i; };
  const int T1 = 2 * N, T2 = T1 + 1, T3 = T1 + 2;
  const M none{nullptr, 0};
  auto Ad = [&](int i) { return blk(g.A, g.sA, e, g.lA.diag[i], bs(i)); };
  auto Au = [&](int i) { return blk(g.A, g.sA, e, g.lA.up[i], bs(i + 1)); };
  auto Al = [&](int i) { return blk(g.A, g.sA, e, g.lA.lo[i], bs(i)); };
  auto obc = [&](int i) {
    const z2* p = g.obc[i];
    return M{p ? const_cast<z2*>(p) + e * (int64_t)bs(i) * bs(i) : nullptr, bs(i)};
  };

  // forward sweep (rgf.py:70-92)
  saxpby(threadIdx.x, kSolveThreads, bs(0), bs(0), kOneS, Ad(0), kMinusS, obc(0), slot(F(0), bs(0)));
  sync();
  sinv(bs(0), slot(F(0), bs(0)), info, sh);
  for (int i = 0; i + 1 < N; ++i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const M t1 = slot(T1, bi), Xj = slot(F(j), bj);
    const bool ob = obc(j).p != nullptr;
    if (team == 0)
      sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kOneS, Al(i), slot(F(i), bi), kZeroS, none, t1);
    else if (ob)
      saxpby(tt, kThreads, bj, bj, kOneS, Ad(j), kMinusS, obc(j), Xj);
    sync();
    if (team == 0)
      sgemm(tt, sp, bj, bj, bi, QX_OP_N, QX_OP_N, kMinusS, t1, Au(i), kOneS, ob ? Xj : Ad(j), Xj);
    sync();
    sinv(bj, Xj, info, sh);
  }
  {
    const M xd = blk(g.X, g.sX, e, g.lX.diag[N - 1], bs(N - 1));
    if (xd.p)
      saxpby(threadIdx.x, kSolveThreads, bs(N - 1), bs(N - 1), kOneS, slot(F(N - 1), bs(N - 1)),
             kZeroS, none, xd);
  }
  // backward sweep (rgf.py:97-110)
  for (int i = N - 2; i >= 0; --i) {
    const int j = i + 1, bi = bs(i), bj = bs(j);
    const M Xii = slot(F(i), bi), Xjj = slot(full(j), bj);
    const M t1 = slot(T1, bi), t2 = slot(T2, bj);
    M Xji = blk(g.X, g.sX, e, g.lX.lo[i], bi);
    if (!Xji.p) Xji = slot(T3, bi);
    const M up = blk(g.X, g.sX, e, g.lX.up[i], bj);
    // x_ji = -(x_jj a_ji) x_ii on team 0, x_ij = -(x_ii a_ij) x_jj on team 1
    if (team == 0)
      sgemm(tt, sp, bj, bi, bj, QX_OP_N, QX_OP_N, kOneS, Xjj, Al(i), kZeroS, none, t1);
    else
      sgemm(tt, sp, bi, bj, bi, QX_OP_N, QX_OP_N, kOneS, Xii, Au(i), kZeroS, none, t2);
    sync();
    if (team == 0)
      sgemm(tt, sp, bj, bi, bi, QX_OP_N, QX_OP_N, kMinusS, t1, Xii, kZeroS, none, Xji);
    else if (up.p)
      sgemm(tt, sp, bi, bj, bj, QX_OP_N, QX_OP_N, kMinusS, t2, Xjj, kZeroS, none, up);
    sync();
    // x_ii <- x_ii - x_ii a_ij x_ji
    const M Z = slot(full(i), bi);
    if (team == 0)
      sgemm(tt, sp, bi, bi, bj, QX_OP_N, QX_OP_N, kMinusS, t2, Xji, kOneS, Xii, Z);
    sync();
    const M xd = blk(g.X, g.sX, e, g.lX.diag[i], bi);
    if (xd.p) saxpby(threadIdx.x, kSolveThreads, bi, bi, kOneS, Z, kZeroS, none, xd);
  }
}

SLay make_lay(const qx_btd_layout* l, int N) {
  SLay s;
  for (int i = 0; i < kSN; ++i) s.diag[i] = s.up[i] = s.lo[i] = -1;
  if (!l) return s;
  for (int i = 0; i < N; ++i) s.diag[i] = (int32_t)l->diag[i];
  for (int i = 0; i + 1 < N; ++i) {
    s.up[i] = (int32_t)l->upper[i];
    s.lo[i] = (int32_t)l->lower[i];
  }
  return s;
}

bool fits32(const qx_btd_layout* l, int N) {
  if (!l) return true;
  auto ok = [](int64_t v) { return v >= -1 && v < (int64_t(1) << 31); };
  for (int i = 0; i < N; ++i)
    if (!ok(l->diag[i])) return false;
  for (int i = 0; i + 1 < N; ++i)
    if (!ok(l->upper[i]) || !ok(l->lower[i])) return false;
  return true;
}

}  // namespace

// Batched inverse X = inv(A - S) of blocks n <= 32 in one launch (complex128
// arithmetic), one matrix per workgroup: measured at n = 32 it beats both the
// split panel kernels and one matrix per wave from 16 to 4096 matrices (37 vs
// 49 / 69 us at 256, 207 vs 623 / 296 us at 4096).  QX_INV_SMALL=0 sends every
// size to the split panel kernels, 1 to the per-wave kernel.
bool inv_small(hipStream_t st, int64_t batch, int n, const Mat& A, const Mat& S, const Mat& X,
               int* info, const int* mask) {
  const char* env = std::getenv("QX_INV_SMALL");
  const int mode = env ? std::atoi(env) : -1;
  if (mode == 0) return false;
  if (n < 1 || n > kSB) return false;
  if (batch <= 0) return true;
  QX_REQUIRE(batch < (int64_t(1) << 31), "inv: batch too large");
  const bool per_wave = mode == 1;
  if (per_wave) {
    const int64_t blocks = (batch + kThreads / 64 - 1) / (kThreads / 64);
    hipLaunchKernelGGL(inv_small_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, st, batch,
                       n, A.p, A.stride, A.ld, S.p, S.stride, S.ld, X.p, X.stride, X.ld, info,
                       mask);
  } else {
    hipLaunchKernelGGL(inv_small_wg_kernel, dim3((unsigned)batch), dim3(kThreads), 0, st, n,
                       A.p, A.stride, A.ld, S.p, S.stride, S.ld, X.p, X.stride, X.ld, info, mask);
  }
  QX_HIP_CHECK(hipGetLastError());
  return true;
}

// Fused path of rgf_selected_solve for small blocks; returns false (and does
// nothing) when the call is not eligible: complex128 arithmetic (4M or 3M),
// at most 16 blocks of at most 32, offsets below 2^31, and QX_RGF_FUSED not 0.
bool rgf_small_solve(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                     const qx_btd_layout* lsl, const void* sl, const qx_btd_layout* lsg,
                     const void* sg, const void* const* obc_r, const void* const* obc_l,
                     const void* const* obc_g, const qx_btd_layout* lxl, void* xl,
                     const qx_btd_layout* lxg, void* xg, const qx_btd_layout* lxr, void* xr,
                     void* current, void* corner, double* density, double* ldos,
                     uint32_t flags, int prec) {
  const char* env = std::getenv("QX_RGF_FUSED");
  if (env && std::atoi(env) == 0) return false;
  if (ldos) return false;  // the per-product path writes the LDOS
  if (prec != QX_PREC_C128 && prec != QX_PREC_C128_3M) return false;
  const int N = la->n_blocks;
  if (N < 1 || N > kSN) return false;
  for (int i = 0; i < N; ++i)
    if (la->block_sizes[i] < 1 || la->block_sizes[i] > kSB) return false;
  const qx_btd_layout* ls[6] = {la, lsl, lsg, lxl, lxg, lxr};
  for (auto* l : ls)
    if (!fits32(l, N)) return false;
  const int64_t S = la->n_stack;
  if (S <= 0) return true;
  QX_REQUIRE(S < (int64_t(1) << 31), "selected_solve: stack too large");
  SmallArgs g;
  g.N = N;
  g.flags = (int)flags;
  g.NS = 6 * N + 23;
  g.split = prec == QX_PREC_C128_3M ? 3 : 4;
  int off = 0;
  for (int i = 0; i < kSN; ++i) {
    g.bs[i] = i < N ? la->block_sizes[i] : 0;
    g.orb[i] = off;
    if (i < N) off += la->block_sizes[i];
  }
  g.n_orb = off;
  g.sA = la->stack_stride;
  g.sSL = lsl ? lsl->stack_stride : 0;
  g.sSG = lsg ? lsg->stack_stride : 0;
  g.sXL = lxl->stack_stride;
  g.sXG = lxg->stack_stride;
  g.sXR = lxr ? lxr->stack_stride : 0;
  g.A = static_cast<const z2*>(a);
  g.SL = static_cast<const z2*>(sl);
  g.SG = static_cast<const z2*>(sg);
  g.XL = static_cast<z2*>(xl);
  g.XG = static_cast<z2*>(xg);
  g.XR = static_cast<z2*>(xr);
  g.lA = make_lay(la, N);
  g.lSL = make_lay(sl ? lsl : nullptr, N);
  g.lSG = make_lay(sg ? lsg : nullptr, N);
  g.lXL = make_lay(lxl, N);
  g.lXG = make_lay(lxg, N);
  g.lXR = make_lay(xr ? lxr : nullptr, N);
  const void* const* ob[3] = {obc_r, obc_l, obc_g};
  for (int q = 0; q < 3; ++q)
    for (int i = 0; i < kSN; ++i)
      g.obc[q][i] = (ob[q] && i < N) ? static_cast<const z2*>(ob[q][i]) : nullptr;
  g.ws = static_cast<z2*>(h->ensure_ws((size_t)S * g.NS * kSB * kSB * sizeof(z2)));
  g.info = h->ensure_info(S);
  g.current = static_cast<z2*>(current);
  g.corner = static_cast<z2*>(corner);
  g.density = density;
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(st, &ev0);
  if (g.split == 3)
    hipLaunchKernelGGL(rgf_small_kernel<3>, dim3((unsigned)S), dim3(kSolveThreads), 0, st, g);
  else
    hipLaunchKernelGGL(rgf_small_kernel<4>, dim3((unsigned)S), dim3(kSolveThreads), 0, st, g);
  QX_HIP_CHECK(hipGetLastError());
  if (timer) {
    // matrix flops of the products and inverses the kernel runs when no
    // optional term is present (no current, corner, Σ off-diagonal blocks or
    // off-diagonal G^R): 8 m n k per product, 8 n^3 per inverse; all counted
    // as GEMM work of this one launch
    auto cube = [](double m, double n, double k) { return 8.0 * m * n * k; };
    const double b0 = la->block_sizes[0];
    double fl = 4.0 * cube(b0, b0, b0) + cube(b0, b0, b0);
    for (int i = 0; i + 1 < N; ++i) {
      const double bi = la->block_sizes[i], bj = la->block_sizes[i + 1];
      fl += 3.0 * cube(bj, bi, bi) + 3.0 * cube(bj, bj, bi) + 4.0 * cube(bj, bj, bj)  // forward
            + cube(bj, bj, bj)                                                          // inverse
            + cube(bi, bj, bi) + cube(bj, bi, bj) + cube(bi, bj, bj) + cube(bi, bi, bj)  // b1..b4
            + 3.0 * cube(bi, bi, bi) + 2.0 * cube(bi, bj, bj) + 2.0 * cube(bi, bi, bj) +
            2.0 * cube(bi, bj, bi);
    }
    timer->end(st, ev0, KernelTimer::GEMM, fl * (double)S);
  }
  return true;
}

// Fused path of rgf_selected_inv for small blocks (same eligibility as the
// fused solve); returns false, having done nothing, when not eligible.
bool rgf_small_inv(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                   const void* const* obc_r, const qx_btd_layout* lx, void* x, int prec) {
  const char* env = std::getenv("QX_RGF_FUSED");
  if (env && std::atoi(env) == 0) return false;
  if (prec != QX_PREC_C128 && prec != QX_PREC_C128_3M) return false;
  const int N = la->n_blocks;
  if (N < 1 || N > kSN) return false;
  for (int i = 0; i < N; ++i)
    if (la->block_sizes[i] < 1 || la->block_sizes[i] > kSB) return false;
  if (!fits32(la, N) || !fits32(lx, N)) return false;
  const int64_t S = la->n_stack;
  if (S <= 0) return true;
  QX_REQUIRE(S < (int64_t(1) << 31), "selected_inv: stack too large");
  SmallInvArgs g;
  g.N = N;
  g.NS = 2 * N + 3;
  g.split = prec == QX_PREC_C128_3M ? 3 : 4;
  for (int i = 0; i < kSN; ++i) {
    g.bs[i] = i < N ? la->block_sizes[i] : 0;
    g.obc[i] = (obc_r && i < N) ? static_cast<const z2*>(obc_r[i]) : nullptr;
  }
  g.sA = la->stack_stride;
  g.sX = lx->stack_stride;
  g.A = static_cast<const z2*>(a);
  g.X = static_cast<z2*>(x);
  g.lA = make_lay(la, N);
  g.lX = make_lay(lx, N);
  g.ws = static_cast<z2*>(h->ensure_ws((size_t)S * g.NS * kSB * kSB * sizeof(z2)));
  g.info = h->ensure_info(S);
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(st, &ev0);
  if (g.split == 3)
    hipLaunchKernelGGL(rgf_small_inv_kernel<3>, dim3((unsigned)S), dim3(kSolveThreads), 0, st, g);
  else
    hipLaunchKernelGGL(rgf_small_inv_kernel<4>, dim3((unsigned)S), dim3(kSolveThreads), 0, st, g);
  QX_HIP_CHECK(hipGetLastError());
  if (timer) {
    // matrix flops of its products and inverses (8 m n k, 8 n^3), counted as
    // GEMM work of this one launch
    auto cube = [](double m, double n, double k) { return 8.0 * m * n * k; };
    const double b0 = la->block_sizes[0];
    double fl = cube(b0, b0, b0);
    for (int i = 0; i + 1 < N; ++i) {
      const double bi = la->block_sizes[i], bj = la->block_sizes[i + 1];
      fl += cube(bj, bi, bi) + cube(bj, bj, bi) + cube(bj, bj, bj)  // forward + inverse
            + cube(bj, bi, bj) + cube(bj, bi, bi) + cube(bi, bj, bi) + cube(bi, bj, bj) +
            cube(bi, bi, bj);                                          // backward
    }
    timer->end(st, ev0, KernelTimer::GEMM, fl * (double)S);
  }
  return true;
}

}  // namespace qx
