// Batched complex matrix inverse with partial pivoting (gfx950).
//
// Replaces the batched `linalg.inv` of the reference
// (src/qttools/kernels/linalg/inv.py:24-40, LAPACK/cuSOLVER getrf+getri).
//
// Algorithm: blocked in-place Gauss-Jordan elimination split into kernels so
// that the O(n^3) part runs on every CU of the chip while the sequential
// pivoting chain stays short:
//   for each panel of NB = 32 columns (QX_INV_NB=64: 64 for n <= 256)
//     panel kernel  (one workgroup per matrix): by default inv_panel_cg, in
//        which each wave holds a group of the panel's columns for all rows
//        and the rows never move (one barrier per column; see the kernel);
//        the earlier row layouts stay selectable (QX_INV_PANEL=4: four
//        threads per row for n <= 256, inv_panel4; QX_INV_PANEL=1: one or two
//        rows per thread, 16-column panels above n = 256, inv_panel).  The
//        pivot of a column is the first row of maximal |re| + |im| among the
//        unreduced rows, compared through a 32-bit key (float magnitude with
//        9 low mantissa bits replaced by the row index).  The reduced panel,
//        the pivot rows and the composed row gather sig are written out.
//     update kernel (64 x 64 tiles of every matrix, MFMA; the rank-NB sum
//        staged through LDS in 16- or 32-deep k-chunks): the other columns
//        receive the row interchanges and the rank-NB update in one pass,
//          Xn[r, C] = (r in pivot rows ? 0 : Xc[sig r, C]) + P[r, :] Xc[sig K, C],
//        reading buffer Xc and writing buffer Xn (ping-pong), so the gather
//        is race free across workgroups.
//   Below 256 matrices per call the panel of step p + 1 runs beside the
//   update of step p (look-ahead, inv_step): only the update of the column
//   tile that holds the next panel precedes it.
//   A final gather undoes the column interchanges (reverse order of the
//   pivots) into the output.  The flop count equals getrf + getri
//   (8 n^3 real flops per complex matrix).
//
// Precision: PREC_C128 computes in double; PREC_C64 holds the panel in
// float registers and runs the update on f32 MFMA; storage stays complex128.
//
// MFMA 16x16x4 layouts as in gemm.hip: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; result element e of lane l is D[l/16 + 4e][l%16] (f64) or
// D[4(l/16) + e][l%16] (f32).
#include <cstdlib>

#include "qx_internal.h"

namespace qx {
namespace {

constexpr int NT = 256;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

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

__device__ inline unsigned wave_max(unsigned v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Wave-wide unsigned max without LDS: DPP butterflies inside each quad and
// each row of 16 lanes, then the four row maxima read into scalars.  Every
// lane of the wave must be active.
__device__ inline unsigned wave_max_dpp(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xf, 0xf, false));   // quad xor 1
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xf, 0xf, false));   // quad xor 2
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xf, 0xf, false));  // row_ror 4
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xf, 0xf, false));  // row_ror 8
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
  const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
  const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
  return max(max(a, b), max(c, d));
}

// Value of lane QQ of each quad, to all four lanes of the quad (DPP).
template <int QQ>
__device__ inline double quad_bcast(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), QQ * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), QQ * 0x55, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double quad_bcast(double x, int qq) {
  switch (qq) {
    case 0: return quad_bcast<0>(x);
    case 1: return quad_bcast<1>(x);
    case 2: return quad_bcast<2>(x);
    default: return quad_bcast<3>(x);
  }
}

struct PanelArgs {
  int n, k0, nbw;
  const z2* Xc;
  int64_t sXc;
  int ldc;
  z2* Xn;
  int64_t sXn;
  int ldn;
  int* ints;  // per matrix: two row-gather buffers sig[2][n], then idx[n]
  int64_t sI;
  int* info;
  const int* mask;  // nullable: matrices with mask[e] == 0 are left alone
  int sig_sel;  // the panel's row gather goes to ints + n (0) or ints (1)
};

// Row and column gathers of a finished panel.  Each position is traced
// through the panel's transpositions S_kk = (k0 + kk <-> piv[kk]) in
// registers, so every thread resolves its own positions (the swaps applied
// one after another to an array would be a serial chain of dependent LDS
// round trips on one thread).
//   sig[u]: row gather of this panel, the identity after the swaps
//           S_0, ..., S_{nbw-1} in that order: S_0(S_1(... S_{nbw-1}(u)));
//   idx[u]: column gather of the whole inverse, the identity after the
//           swaps of all panels in reverse order, i.e. u pushed forward
//           through every swap of the factorisation: each panel extends the
//           previous panel's idx by its own swaps, so the last one holds the
//           complete gather and no panel traces more than NB steps.
template <int NTH>
__device__ __forceinline__ void panel_gathers(const int* piv, int n, int k0, int nbw, int tid,
                                              int* sig, int* idx) {
  for (int u = tid; u < n; u += NTH) {
    int pc = (k0 == 0) ? u : idx[u];
    int pr = u;
#pragma unroll 8
    for (int kk = 0; kk < nbw; ++kk) {
      const int k = k0 + kk, p = piv[kk];
      pc = (pc == k) ? p : ((pc == p) ? k : pc);
      const int k2 = k0 + nbw - 1 - kk, p2 = piv[nbw - 1 - kk];
      pr = (pr == k2) ? p2 : ((pr == p2) ? k2 : pr);
    }
    sig[u] = pr;
    idx[u] = pc;
  }
}

// Each thread keeps its row(s) of the panel in registers, rotated so that
// the column being eliminated is always element 0: at step kk the row holds
// columns (kk, kk+1, ..., NB-1, 0, ..., kk-1).  All register indices are then
// compile-time constants; a rotation by one element closes every step.
template <typename T, int NB, int RPT>
__global__ __launch_bounds__(NT) void inv_panel(PanelArgs g) {
  __shared__ T prow_r[NB], prow_i[NB], krow_r[NB], krow_i[NB];
  __shared__ unsigned red[NT / 64];
  __shared__ int perm_l[NB];
  const int64_t e = blockIdx.x;
  if (g.mask && !g.mask[e]) return;  // uniform over the workgroup
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = g.n, k0 = g.k0, nbw = g.nbw;
  const z2* Xc = g.Xc + e * g.sXc;
  z2* Xn = g.Xn + e * g.sXn;
  int* sig = g.ints + e * g.sI + (g.sig_sel ? 0 : n);
  int* idx = g.ints + e * g.sI + 2 * n;

  T rr[RPT][NB], ri[RPT][NB];
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int r = tid + NT * j;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      z2 v = make_double2(0.0, 0.0);
      if (r < n && c < nbw) v = Xc[(int64_t)r * g.ldc + k0 + c];
      rr[j][c] = (T)v.x;
      ri[j][c] = (T)v.y;
    }
  }
  // pivot candidates of the first column
  unsigned key = 0;
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int r = tid + NT * j;
    if (r >= k0 && r < n) {
      const unsigned kj = pivot_key((double)rr[j][0], (double)ri[j][0], r);
      key = kj > key ? kj : key;
    }
  }
  key = wave_max(key);
  if (lane == 0) red[w] = key;
  __syncthreads();

  for (int kk = 0; kk < nbw; ++kk) {
    const int k = k0 + kk;
    unsigned kmax = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) kmax = red[q] > kmax ? red[q] : kmax;
    int p = 511 - (int)(kmax & 0x1FFu);
    if (p < k || p >= n) p = k;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = tid + NT * j;
      if (r == p) {
        double vr, vi;
        complex_recip((double)rr[j][0], (double)ri[j][0], vr, vi);
        prow_r[0] = (T)vr;
        prow_i[0] = (T)vi;
#pragma unroll
        for (int c = 1; c < NB; ++c) {
          const double a = (double)rr[j][c], b = (double)ri[j][c];
          prow_r[c] = (T)(a * vr - b * vi);
          prow_i[c] = (T)(a * vi + b * vr);
        }
      }
      if (r == k && p != k) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          krow_r[c] = rr[j][c];
          krow_i[c] = ri[j][c];
        }
      }
    }
    if (tid == 0) {
      perm_l[kk] = p;
      if ((kmax >> 9) == 0u) g.info[e] = 1;  // no nonzero pivot: singular
    }
    __syncthreads();

    key = 0;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = tid + NT * j;
      if (r < n) {
        if (r == k) {
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            rr[j][c] = prow_r[c];
            ri[j][c] = prow_i[c];
          }
        } else {
          if (r == p) {
#pragma unroll
            for (int c = 0; c < NB; ++c) {
              rr[j][c] = krow_r[c];
              ri[j][c] = krow_i[c];
            }
          }
          const T fr = rr[j][0], fi = ri[j][0];
          {
            const T br = prow_r[0], bi = prow_i[0];
            rr[j][0] = -(fr * br - fi * bi);
            ri[j][0] = -(fr * bi + fi * br);
          }
#pragma unroll
          for (int c = 1; c < NB; ++c) {
            const T br = prow_r[c], bi = prow_i[c];
            rr[j][c] -= fr * br - fi * bi;
            ri[j][c] -= fr * bi + fi * br;
          }
          if (kk + 1 < nbw && r > k) {
            const unsigned kj = pivot_key((double)rr[j][1], (double)ri[j][1], r);
            key = kj > key ? kj : key;
          }
        }
      }
      // rotate left by one: the next column moves to element 0
      const T t_r = rr[j][0], t_i = ri[j][0];
#pragma unroll
      for (int c = 0; c + 1 < NB; ++c) {
        rr[j][c] = rr[j][c + 1];
        ri[j][c] = ri[j][c + 1];
      }
      rr[j][NB - 1] = t_r;
      ri[j][NB - 1] = t_i;
    }
    if (kk + 1 < nbw) {
      key = wave_max(key);
      if (lane == 0) red[w] = key;
    }
    __syncthreads();
  }

  // reduced panel -> Xn (element c now holds panel column (c + nbw) mod NB)
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int r = tid + NT * j;
    if (r < n) {
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        int col = c + nbw;
        col -= (col >= NB) ? NB : 0;
        if (col < nbw)
          Xn[(int64_t)r * g.ldn + k0 + col] = make_double2((double)rr[j][c], (double)ri[j][c]);
      }
    }
  }
  // gathers (this panel's pivots are in perm_l since the column loop's last
  // barrier)
  panel_gathers<NT>(perm_l, n, k0, nbw, tid, sig, idx);
}

// Panel kernel for n <= 256 with four threads per row (1024 threads): thread
// t holds columns 8 q .. 8 q + 7 (q = t % 4) of row t / 4 of the 32-column
// panel, so each elimination step costs a thread 8 complex updates instead
// of 32, and the four threads of a row are neighbouring lanes of one wave
// (the active column's value reaches them by a DPP quad broadcast).  The
// eight columns of a quarter are unrolled, so every register index is a
// compile-time constant and nothing is rotated.  Pivot choice, outputs and
// the row / column gathers are those of inv_panel.
constexpr int NT4 = 1024;

template <typename T>
__global__ __launch_bounds__(NT4) void inv_panel4(PanelArgs g) {
  constexpr int NB = 32, CQ = 8;
  __shared__ T prow_r[NB], prow_i[NB], krow_r[NB], krow_i[NB];
  __shared__ unsigned red[NT4 / 64];
  __shared__ int perm_l[NB];
  const int64_t e = blockIdx.x;
  if (g.mask && !g.mask[e]) return;  // uniform over the workgroup
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = tid >> 2, q = tid & 3;
  const int n = g.n, k0 = g.k0, nbw = g.nbw;
  const z2* Xc = g.Xc + e * g.sXc;
  z2* Xn = g.Xn + e * g.sXn;
  int* sig = g.ints + e * g.sI + (g.sig_sel ? 0 : n);
  int* idx = g.ints + e * g.sI + 2 * n;

  T rr[CQ], ri[CQ];
#pragma unroll
  for (int c = 0; c < CQ; ++c) {
    z2 v = make_double2(0.0, 0.0);
    const int col = q * CQ + c;
    if (r < n && col < nbw) v = Xc[(int64_t)r * g.ldc + k0 + col];
    rr[c] = (T)v.x;
    ri[c] = (T)v.y;
  }
  // pivot candidates of the first panel column (quarter 0, register 0)
  unsigned key = 0;
  if (q == 0 && r >= k0 && r < n) key = pivot_key((double)rr[0], (double)ri[0], r);
  key = wave_max_dpp(key);
  if (lane == 0) red[w] = key;
  __syncthreads();

  // The quarters run as a loop and only the eight columns of a quarter are
  // unrolled: the fully unrolled 32 columns are ~78 KB of code, more than
  // the instruction cache, and every column then waits on instruction
  // fetches from L2.
#pragma unroll 1
  for (int qq = 0; qq < 4; ++qq) {
#pragma unroll
    for (int c = 0; c < CQ; ++c) {
      const int kk = qq * CQ + c;
      if (kk < nbw) {
        const int k = k0 + kk;
        unsigned kmax = red[0];
#pragma unroll
        for (int u = 1; u < NT4 / 64; ++u) kmax = red[u] > kmax ? red[u] : kmax;
        int p = 511 - (int)(kmax & 0x1FFu);
        if (p < k || p >= n) p = k;
        // pivot row: 1 / x_pk to all four threads of the row, scaled row to LDS
        const double pr = quad_bcast((double)rr[c], qq);
        const double pi = quad_bcast((double)ri[c], qq);
        if (r == p) {
          double vr, vi;
          complex_recip(pr, pi, vr, vi);
#pragma unroll
          for (int cc = 0; cc < CQ; ++cc) {
            if (q == qq && cc == c) {
              prow_r[kk] = (T)vr;
              prow_i[kk] = (T)vi;
            } else {
              const double a = (double)rr[cc], b = (double)ri[cc];
              prow_r[q * CQ + cc] = (T)(a * vr - b * vi);
              prow_i[q * CQ + cc] = (T)(a * vi + b * vr);
            }
          }
        }
        if (r == k && p != k) {
#pragma unroll
          for (int cc = 0; cc < CQ; ++cc) {
            krow_r[q * CQ + cc] = rr[cc];
            krow_i[q * CQ + cc] = ri[cc];
          }
        }
        if (tid == 0) {
          perm_l[kk] = p;
          if ((kmax >> 9) == 0u) g.info[e] = 1;  // no nonzero pivot: singular
        }
        __syncthreads();

        if (r < n) {
          if (r == k) {
#pragma unroll
            for (int cc = 0; cc < CQ; ++cc) {
              rr[cc] = prow_r[q * CQ + cc];
              ri[cc] = prow_i[q * CQ + cc];
            }
          } else if (r == p) {
#pragma unroll
            for (int cc = 0; cc < CQ; ++cc) {
              rr[cc] = krow_r[q * CQ + cc];
              ri[cc] = krow_i[q * CQ + cc];
            }
          }
        }
        // multiplier of this row: its (swapped-in) value in the pivot column
        const T fr = (T)quad_bcast((double)rr[c], qq);
        const T fi = (T)quad_bcast((double)ri[c], qq);
        if (r < n && r != k) {
#pragma unroll
          for (int cc = 0; cc < CQ; ++cc) {
            const T br = prow_r[q * CQ + cc], bi = prow_i[q * CQ + cc];
            if (q == qq && cc == c) {
              rr[cc] = -(fr * br - fi * bi);
              ri[cc] = -(fr * bi + fi * br);
            } else {
              rr[cc] -= fr * br - fi * bi;
              ri[cc] -= fr * bi + fi * br;
            }
          }
        }
        // pivot candidates of the next column
        key = 0;
        const int kn = kk + 1;
        const int cn = (c + 1) % CQ;  // register of column kn (constant once unrolled)
        if (kn < nbw && r > k && r < n && q == qq + (c + 1) / CQ)
          key = pivot_key((double)rr[cn], (double)ri[cn], r);
        key = wave_max_dpp(key);
        if (lane == 0) red[w] = key;
        __syncthreads();
      }
    }
  }

  if (r < n) {
#pragma unroll
    for (int c = 0; c < CQ; ++c) {
      const int col = q * CQ + c;
      if (col < nbw)
        Xn[(int64_t)r * g.ldn + k0 + col] = make_double2((double)rr[c], (double)ri[c]);
    }
  }
  // gathers (this panel's pivots are in perm_l since the column loop's last
  // barrier)
  panel_gathers<NT4>(perm_l, n, k0, nbw, tid, sig, idx);
}

// Wave-wide unsigned max with DPP only: quad and row butterflies give every
// lane its row's maximum, row_bcast:15 / row_bcast:31 carry the row maxima
// up to lane 63, which is read once.  Every lane of the wave must be active.
__device__ inline unsigned wave_max_bcast(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false));   // quad xor 1
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false));   // quad xor 2
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xf, 0xf, false));  // row_ror 4
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false));  // row_ror 8
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Panel kernel with column groups (n <= 64 RPL, NB-column panels): wave w of
// the NB / CW waves holds columns CW w .. CW w + CW - 1 of the panel for all
// rows, lane l the rows l + 64 j (j < RPL).  Rows stay in their registers;
// the row interchanges of the factorisation are tracked as positions (pos_j:
// where row l + 64 j currently stands), so the candidates of a column are the
// rows whose position is not below it, and every row is written to its
// final position at the end.
//
// The scaling of a pivot row by 1 / x_pk is deferred to the end of the panel:
// the pivot row keeps its values (its multiplier is taken as zero, its entry
// in the active column becomes 1), and every later elimination step acts on
// it linearly, so scaling it once at the end gives the same row.  Nothing
// then has to be written into the register slot of a runtime row.
//
// Per column, with one barrier:
//   - the wave owning the column has published the column (all rows), the
//     pivot key (wave max of the candidates' keys) and the pivot value to
//     LDS, double-buffered by column parity so no second barrier guards the
//     reuse;
//   - after the barrier every wave reads the pivot key and value and the
//     multipliers of its rows, and takes its part of row p (and the row's
//     position) from the lane that holds it (v_readlane, p being uniform);
//     1 / x_pk, the scaled pivot row and the elimination follow;
//   - the owner of the next column computes and publishes its pivot data.
// Pivot choice (largest |re| + |im| through the 32-bit key, ties to the lower
// row), arithmetic and outputs are those of the other panel kernels up to
// rounding; ties are broken by the row's index at the start of the panel.
// Value of v in lane l of the wave (l uniform).
__device__ inline double lane_read(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline float lane_read(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <typename T, int RPL, int CW, int NB>
struct PanelCgShared {
  T col_r[2][64 * RPL], col_i[2][64 * RPL];
  T piv_r[2], piv_i[2];
  unsigned key_b[2];
  double rcp_r[NB], rcp_i[NB];
  int perm_l[NB];
};

// The panel of matrix e, by a workgroup of 64 NB / CW threads.
template <typename T, int RPL, int CW, int NB>
__device__ __forceinline__ void panel_cg(const PanelArgs& g, int64_t e,
                                         PanelCgShared<T, RPL, CW, NB>& sh) {
  constexpr int NW = NB / CW, NTH = 64 * NW;
  auto& col_r = sh.col_r;
  auto& col_i = sh.col_i;
  auto& piv_r = sh.piv_r;
  auto& piv_i = sh.piv_i;
  auto& key_b = sh.key_b;
  auto& rcp_r = sh.rcp_r;
  auto& rcp_i = sh.rcp_i;
  auto& perm_l = sh.perm_l;
  if (g.mask && !g.mask[e]) return;  // uniform over the workgroup
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = g.n, k0 = g.k0, nbw = g.nbw;
  const z2* Xc = g.Xc + e * g.sXc;
  z2* Xn = g.Xn + e * g.sXn;
  int* sig = g.ints + e * g.sI + (g.sig_sel ? 0 : n);
  int* idx = g.ints + e * g.sI + 2 * n;

  T xr[RPL][CW], xi[RPL][CW];
  int pos[RPL];  // where row lane + 64 j currently stands
#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    const int r = lane + 64 * j;
    pos[j] = r;
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      // unconditional loads from a clamped address, so that all of them are
      // in flight together (a load under a branch is waited for in the branch)
      const int cc = w * CW + c;
      const bool ok = r < n && cc < nbw;
      const z2 v = Xc[ok ? (int64_t)r * g.ldc + k0 + cc : 0];
      xr[j][c] = ok ? (T)v.x : (T)0;
      xi[j][c] = ok ? (T)v.y : (T)0;
    }
  }
  // Pivot data of panel column cn (register cn of the owner wave) into
  // buffer b: the column, the key of the candidates (rows at position > kb)
  // and the pivot value; the pivot row's entry of the published column (its
  // multiplier in the elimination) is then zeroed, as the deferred scaling
  // requires.
  auto publish = [&](int cn, int b, int kb) {
    unsigned key = 0;
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      const int r = lane + 64 * j;
      col_r[b][r] = xr[j][cn];
      col_i[b][r] = xi[j][cn];
      if (pos[j] > kb && r < n) {
        const unsigned kj = pivot_key((double)xr[j][cn], (double)xi[j][cn], r);
        key = kj > key ? kj : key;
      }
    }
    key = wave_max_bcast(key);
    const int pn = min(511 - (int)(key & 0x1FFu), n - 1);
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      if (lane + 64 * j == pn) {
        piv_r[b] = xr[j][cn];
        piv_i[b] = xi[j][cn];
        col_r[b][pn] = (T)0;
        col_i[b][pn] = (T)0;
      }
    }
    if (lane == 0) key_b[b] = key;
  };
  if (w == 0) publish(0, 0, k0 - 1);
  __syncthreads();

  // The column groups run as a loop and only the CW columns of a group are
  // unrolled (register indices stay compile-time constants).
#pragma unroll 1
  for (int qa = 0; qa < NW; ++qa) {
#pragma unroll
    for (int ca = 0; ca < CW; ++ca) {
      const int kk = qa * CW + ca;
      if (kk < nbw) {
        const int k = k0 + kk;
        const int buf = ca & 1;  // kk & 1 (CW is even)
        const unsigned kmax = __builtin_amdgcn_readfirstlane(key_b[buf]);
        // a candidate always exists (n - k rows), so p < n; min() keeps every
        // index inside the panel whatever the key holds
        const int p = min(511 - (int)(kmax & 0x1FFu), n - 1);
        const double pvr = (double)piv_r[buf], pvi = (double)piv_i[buf];
        T fr[RPL], fi[RPL];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          fr[j] = col_r[buf][lane + 64 * j];  // zero for the pivot row
          fi[j] = col_i[buf][lane + 64 * j];
        }
        // this wave's part of row p and its position, read from the lane
        // that holds them (p is uniform: register row p >> 6 of lane p & 63)
        const int pl = p & 63, pj = p >> 6;
        T ar[CW], ai[CW];
        int q = 0;
#pragma unroll
        for (int c = 0; c < CW; ++c) ar[c] = ai[c] = (T)0;
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          if (j == pj) {  // uniform: a scalar branch
#pragma unroll
            for (int c = 0; c < CW; ++c) {
              ar[c] = lane_read(xr[j][c], pl);
              ai[c] = lane_read(xi[j][c], pl);
            }
            q = __builtin_amdgcn_readlane(pos[j], pl);
          }
        }
        double vr, vi;
        complex_recip(pvr, pvi, vr, vi);
        T sr[CW], si[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          const double a = (double)ar[c], b = (double)ai[c];
          sr[c] = (T)(a * vr - b * vi);
          si[c] = (T)(a * vi + b * vr);
        }
        if (w == qa) {
          // active column: -f / x_pk for the other rows, 1 (scaled at the
          // end) for the pivot row
          sr[ca] = (T)vr;
          si[ca] = (T)vi;
#pragma unroll
          for (int j = 0; j < RPL; ++j) {
            xr[j][ca] = (lane + 64 * j == p) ? (T)1 : (T)0;
            xi[j][ca] = (T)0;
          }
        }
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
#pragma unroll
          for (int c = 0; c < CW; ++c) {
            xr[j][c] = fma(-fr[j], sr[c], fma(fi[j], si[c], xr[j][c]));
            xi[j][c] = fma(-fr[j], si[c], fma(-fi[j], sr[c], xi[j][c]));
          }
        }
        // interchange k <-> q: row p moves to position k, the row at k to q
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          const int r = lane + 64 * j;
          pos[j] = (r == p) ? k : ((pos[j] == k) ? q : pos[j]);
        }
        if (tid == 0) {
          perm_l[kk] = q;
          rcp_r[kk] = vr;
          rcp_i[kk] = vi;
          if ((kmax >> 9) == 0u) g.info[e] = 1;  // no nonzero pivot: singular
        }
        if (kk + 1 < nbw) {
          if (w == qa + (ca + 1) / CW) publish((ca + 1) % CW, buf ^ 1, k);
          __syncthreads();
        }
      }
    }
  }
  __syncthreads();  // perm_l and the pivot reciprocals complete

#pragma unroll
  for (int j = 0; j < RPL; ++j) {
    const int r = lane + 64 * j;
    if (r < n) {
      // the pivot of panel column kk ends at position k0 + kk, and those
      // positions hold exactly this panel's pivots
      const int st = pos[j] - k0;
      const bool pivot = st >= 0 && st < nbw;
      double ar = 1.0, ai = 0.0;
      if (pivot) {
        ar = rcp_r[st];
        ai = rcp_i[st];
      }
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const int cc = w * CW + c;
        if (cc < nbw) {
          const double a = (double)xr[j][c], b = (double)xi[j][c];
          double yr = a, yi = b;
          if (pivot) {
            yr = (double)(T)(a * ar - b * ai);
            yi = (double)(T)(a * ai + b * ar);
          }
          Xn[(int64_t)pos[j] * g.ldn + k0 + cc] = make_double2(yr, yi);
        }
      }
    }
  }
  panel_gathers<NTH>(perm_l, n, k0, nbw, tid, sig, idx);
}

template <typename T, int RPL, int CW, int NB>
__global__ __launch_bounds__(64 * (NB / CW)) void inv_panel_cg(PanelArgs g) {
  __shared__ PanelCgShared<T, RPL, CW, NB> sh;
  panel_cg<T, RPL, CW, NB>(g, blockIdx.x, sh);
}

struct UpdArgs {
  int n, k0, nbw, tiles;  // tiles per matrix: row tiles x column tiles handled
  const z2* Xc;
  int64_t sXc;
  int ldc;
  z2* Xn;
  int64_t sXn;
  int ldn;
  const int* ints;
  int64_t sI;
  const int* mask;  // nullable, as PanelArgs::mask
  int sig_sel;      // the panel's row gather: ints + n (0) or ints (1)
  // Column tiles handled: cfirst + ci for ci < ncols (0: all), stepping over
  // tile cskip1 - 1 (cskip1 = 0: none).
  int cfirst, ncols, cskip1;
};

// Workgroup ids are dealt round-robin over the 8 XCDs; this maps them so
// that each XCD receives one contiguous range of logical ids.
__device__ inline int update_xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx, rem = nblocks % nx;
  const int x = bid % nx, j = bid / nx;
  return x * per + min(x, rem) + j;
}

// Matrix, first row and first column of update tile id.
struct UpdTile {
  int64_t e;
  int r0, c0;
};

__device__ inline UpdTile update_tile_of(const UpdArgs& g, int id) {
  const int tn = (g.n + 63) / 64;
  const int nct = g.ncols ? g.ncols : tn;
  UpdTile u;
  u.e = id / g.tiles;
  const int t = id - (int)(u.e * g.tiles);
  const int ci = g.cfirst + t % nct;
  u.r0 = (t / nct) * 64;
  u.c0 = (ci + (g.cskip1 > 0 && ci >= g.cskip1 - 1 ? 1 : 0)) * 64;
  return u;
}

// A column tile that lies inside the panel, or a matrix left alone, has
// nothing to do.
__device__ inline bool update_live(const UpdArgs& g, const UpdTile& u) {
  if (u.c0 >= g.k0 && u.c0 + 64 <= g.k0 + g.nbw) return false;
  return !(g.mask && !g.mask[u.e]);
}

template <typename T, int NB, int KC>
struct UpdShared {
  static constexpr int KW = NB < KC ? NB : KC;  // staged k per chunk
  T As[2][KW][64 + 1];
  T Bs[2][KW][64 + 1];
  int sig_r[64], sig_k[NB];
};

// Xn[r, c] = (r in K ? 0 : Xc[sig r, c]) + sum_k Xn[r, k0+k] Xc[sig(k0+k), c]
// for the columns c outside the panel, on one 64 x 64 tile by 256 threads
// (tid: the thread's index among them), 4 waves of 32 x 32.  The rank-NB sum
// runs over k-chunks of KC staged through LDS one after another (KC = 32;
// NB = 64: two chunks), so a wider panel costs no more LDS per workgroup.
// KC = 16 with a register target of four waves per SIMD halves the LDS and
// registers of a workgroup so that four instead of two share a CU: more
// loads in flight for large batches, where the update streams every matrix
// from HBM.  The k order of the sum is the same.  A tile that is not live
// loads only from its matrix's first element and stores nothing, and passes
// the same workgroup barriers as a live one.
template <typename T, int NB, int KC>
__device__ __forceinline__ void update_tile(const UpdArgs& g, const UpdTile& u, bool live,
                                            int tid, UpdShared<T, NB, KC>& sh) {
  typedef InvMma<T> MM;
  typedef typename MM::acc_t acc_t;
  static_assert(NB % KC == 0 || NB < KC, "panel width: a multiple of the chunk, or below it");
  constexpr int KW = UpdShared<T, NB, KC>::KW;
  auto& As = sh.As;
  auto& Bs = sh.Bs;
  auto& sig_r = sh.sig_r;
  auto& sig_k = sh.sig_k;
  const int64_t e = u.e;
  const int r0 = u.r0, c0 = u.c0;
  const int n = g.n, k0 = g.k0, nbw = g.nbw;
  const z2* Xc = g.Xc + e * g.sXc;
  z2* Xn = g.Xn + e * g.sXn;
  const int* sig = g.ints + e * g.sI + (g.sig_sel ? 0 : n);
  const int lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;

  if (tid < 64) sig_r[tid] = (r0 + tid < n) ? sig[r0 + tid] : 0;
  if (tid < NB) sig_k[tid] = (tid < nbw) ? sig[k0 + tid] : 0;
  __syncthreads();
  // C init: gathered rows, pivot rows zeroed.  Issued before the operand
  // staging below so that both sets of loads are in flight together.
  acc_t aR[2][2], aI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rl = wm * 32 + i * 16 + MM::row(lane, q);
        const int r = r0 + rl;
        const int c = c0 + wn * 32 + j * 16 + (lane & 15);
        // unconditional loads from clamped addresses: all in flight together
        const bool ok = live && r < n && c < n && !(r >= k0 && r < k0 + nbw);
        const z2 v = Xc[ok ? (int64_t)sig_r[rl] * g.ldc + c : 0];
        aR[i][j][q] = ok ? (T)v.x : (T)0;
        aI[i][j][q] = ok ? (T)v.y : (T)0;
      }
  // per chunk: A = reduced panel rows of this tile, B = gathered pivot rows;
  // all loads first (unconditional, clamped addresses), then the LDS writes
  constexpr int NS = 64 * KW / NT;
#pragma unroll
  for (int kc = 0; kc < NB; kc += KW) {
    if (kc > 0) __syncthreads();  // the previous chunk's LDS reads are done
    z2 va[NS], vb[NS];
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * NT;
      const int r = q / KW, k = q - r * KW;
      const bool oka = live && r0 + r < n && kc + k < nbw;
      va[it] = Xn[oka ? (int64_t)(r0 + r) * g.ldn + k0 + kc + k : 0];
      if (!oka) va[it] = make_double2(0.0, 0.0);
      const int k2 = q >> 6, c = q & 63;
      const bool okb = live && kc + k2 < nbw && c0 + c < n;
      vb[it] = Xc[okb ? (int64_t)sig_k[kc + k2] * g.ldc + c0 + c : 0];
      if (!okb) vb[it] = make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int it = 0; it < NS; ++it) {
      const int q = tid + it * NT;
      const int r = q / KW, k = q - r * KW;
      As[0][k][r] = (T)va[it].x;
      As[1][k][r] = (T)va[it].y;
      const int k2 = q >> 6, c = q & 63;
      Bs[0][k2][c] = (T)vb[it].x;
      Bs[1][k2][c] = (T)vb[it].y;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < KW / 4; ++kq) {
      const int kr = kq * 4 + (lane >> 4);
      T a_r[2], a_i[2], na_i[2], b_r[2], b_i[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a_r[i] = As[0][kr][wm * 32 + i * 16 + (lane & 15)];
        a_i[i] = As[1][kr][wm * 32 + i * 16 + (lane & 15)];
        na_i[i] = -a_i[i];
        b_r[i] = Bs[0][kr][wn * 32 + i * 16 + (lane & 15)];
        b_i[i] = Bs[1][kr][wn * 32 + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          aR[i][j] = MM::mma(a_r[i], b_r[j], aR[i][j]);
          aI[i][j] = MM::mma(a_r[i], b_i[j], aI[i][j]);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          aR[i][j] = MM::mma(na_i[i], b_i[j], aR[i][j]);
          aI[i][j] = MM::mma(a_i[i], b_r[j], aI[i][j]);
        }
    }
  }
  if (!live) return;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + wm * 32 + i * 16 + MM::row(lane, q);
        const int c = c0 + wn * 32 + j * 16 + (lane & 15);
        if (r < n && c < n && !(c >= k0 && c < k0 + nbw))
          Xn[(int64_t)r * g.ldn + c] = make_double2((double)aR[i][j][q], (double)aI[i][j][q]);
      }
}

// The update as its own launch: one tile per workgroup, the tiles of one
// matrix on one XCD, so that the panel columns and pivot rows every tile
// reads again come from that XCD's L2.  WPE: register target (waves per SIMD).
template <typename T, int NB, int KC, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
void inv_update(UpdArgs g) {
  __shared__ UpdShared<T, NB, KC> sh;
  const UpdTile u = update_tile_of(g, update_xcd_remap(blockIdx.x, gridDim.x));
  if (!update_live(g, u)) return;  // uniform over the workgroup
  update_tile<T, NB, KC>(g, u, true, threadIdx.x, sh);
}

// Look-ahead step of the 32-column factorisation: the panel of step p + 1 on
// workgroups [0, np) (np: the batch rounded up to a multiple of 8, so that
// the update's workgroup ids below keep their XCD), and the update of step p
// on the others, two tiles per workgroup of 512 threads (one per half).
// The update passed here covers every column tile except the one holding
// panel p + 1's columns, which an update launch before this one completed,
// so the two parts touch disjoint data: the panel reads its columns of the
// buffer the update writes and writes its columns of the buffer the update
// reads, and the row gathers of consecutive panels live in different sig
// buffers.
template <typename T, int RPL>
__global__ __launch_bounds__(512) void inv_step(PanelArgs pa, UpdArgs ua, int batch, int np,
                                                int64_t total) {
  __shared__ union {
    PanelCgShared<T, RPL, 4, 32> p;
    UpdShared<T, 32, 32> u[2];
  } sh;
  if ((int)blockIdx.x < np) {
    if ((int)blockIdx.x < batch) panel_cg<T, RPL, 4, 32>(pa, blockIdx.x, sh.p);
    return;
  }
  const int half = threadIdx.x >> 8;
  const int64_t id =
      2 * (int64_t)update_xcd_remap(blockIdx.x - np, gridDim.x - np) + half;
  const bool in = id < total;
  const UpdTile u = update_tile_of(ua, in ? (int)id : 0);
  update_tile<T, 32, 32>(ua, u, in && update_live(ua, u), threadIdx.x & (NT - 1), sh.u[half]);
}

// X[r, c] = Y[r, idx c]; X may not alias Y.
__global__ __launch_bounds__(NT) void inv_gather(int n, const z2* Y, int64_t sY, int ldy, z2* X,
                                                 int64_t sX, int ldx, const int* ints, int64_t sI,
                                                 int64_t total, const int* mask) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t q = blockIdx.x * (int64_t)NT + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * NT) {
    const int64_t e = q / nn;
    if (mask && !mask[e]) continue;
    const int64_t rc = q - e * nn;
    const int r = (int)(rc / n), c = (int)(rc - (int64_t)r * n);
    const int* idx = ints + e * sI + 2 * n;
    X[e * sX + (int64_t)r * ldx + c] = Y[e * sY + (int64_t)r * ldy + idx[c]];
  }
}

// Y = A - S (S nullable) for the matrices with mask[e] != 0.
__global__ __launch_bounds__(NT) void inv_masked_start(int n, const z2* A, int64_t sA, int lda,
                                                       const z2* S, int64_t sS, int lds, z2* Y,
                                                       int64_t sY, int ldy, int64_t total,
                                                       const int* mask) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t q = blockIdx.x * (int64_t)NT + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * NT) {
    const int64_t e = q / nn;
    if (!mask[e]) continue;
    const int64_t rc = q - e * nn;
    const int r = (int)(rc / n), c = (int)(rc - (int64_t)r * n);
    z2 v = A[e * sA + (int64_t)r * lda + c];
    if (S) {
      const z2 w = S[e * sS + (int64_t)r * lds + c];
      v = make_double2(v.x - w.x, v.y - w.y);
    }
    Y[e * sY + (int64_t)r * ldy + c] = v;
  }
}

// Panel kernel of the 32-column panels (read per call): 0 (default) column
// groups (inv_panel_cg, any n <= 512), 4 four threads per row (n <= 256),
// 1 one thread per row.  QX_INV_CW = 4 (default: eight waves) or 8 (four
// waves) sets the column-group width.
int inv_panel_kind() {
  const char* e = std::getenv("QX_INV_PANEL");
  const int v = e ? std::atoi(e) : 0;
  return (v == 1 || v == 4) ? v : 0;
}

int inv_group_width() {
  const char* e = std::getenv("QX_INV_CW");
  return (e && std::atoi(e) == 8) ? 8 : 4;
}

// Update kernel: 16-deep k-chunks with four workgroups per CU from 256
// matrices per call (at 1024 matrices of 256: 6.63 -> 6.12 ms per inverse;
// 128: 1.26 -> 1.19 ms), 32-deep chunks with two below (at 32-128 matrices
// equal or up to 2 % faster).  Both give the same bits.  QX_INV_UPDATE=16|32
// (read per call) forces one.
bool narrow_update(int64_t batch) {
  const char* e = std::getenv("QX_INV_UPDATE");
  const int v = e ? std::atoi(e) : 0;
  if (v == 16 || v == 32) return v == 16;
  return batch >= 256;
}

template <typename T, int NB>
void launch_panel_cg(hipStream_t s, int64_t batch, int n, const PanelArgs& pa) {
  if constexpr (NB == 64) {
    // eight waves of eight columns, four rows per lane (n <= 256)
    hipLaunchKernelGGL((inv_panel_cg<T, 4, 8, 64>), dim3((unsigned)batch), dim3(512), 0, s, pa);
    return;
  } else {
    const bool narrow = inv_group_width() == 4;
    if (n <= 256) {
      if (narrow) hipLaunchKernelGGL((inv_panel_cg<T, 4, 4, 32>), dim3((unsigned)batch), dim3(512), 0, s, pa);
      else hipLaunchKernelGGL((inv_panel_cg<T, 4, 8, 32>), dim3((unsigned)batch), dim3(256), 0, s, pa);
    } else {
      if (narrow) hipLaunchKernelGGL((inv_panel_cg<T, 8, 4, 32>), dim3((unsigned)batch), dim3(512), 0, s, pa);
      else hipLaunchKernelGGL((inv_panel_cg<T, 8, 8, 32>), dim3((unsigned)batch), dim3(256), 0, s, pa);
    }
  }
}

// Look-ahead (default below 256 matrices per call; QX_INV_LOOKAHEAD=0, read
// per call, turns it off): with 32-column panels of eight-wave column
// groups, the update of step p is split into the column tile that holds
// panel p + 1 (its own launch) and the rest, which runs in one launch with
// panel p + 1 (inv_step), so the pivoting chain of the next panel no longer
// waits for the whole update.  The same kernels' arithmetic on the same
// data: the same bits.  32 x 256^2: 557 -> 500 us per inverse, 32 x 416^2
// 1691 -> 1241, 32 x 512^2 2315 -> 1629.  From 256 matrices on, where the
// update streams every matrix from HBM and runs 16-deep chunks with four
// workgroups per CU, the combined launch (one workgroup of 512 threads per
// CU) was slower: 1024 x 256^2 5.77 -> 6.87 ms.
bool inv_lookahead() {
  const char* e = std::getenv("QX_INV_LOOKAHEAD");
  return !(e && std::atoi(e) == 0);
}

// An update launch of the look-ahead (32-deep chunks, as inv_step).
template <typename T>
void launch_update(hipStream_t s, int64_t batch, const UpdArgs& ua) {
  const int64_t blocks = batch * ua.tiles;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "inv: grid too large");
  hipLaunchKernelGGL((inv_update<T, 32, 32, 1>), dim3((unsigned)blocks), dim3(NT), 0, s, ua);
}

template <typename T>
void run_inverse_lookahead(hipStream_t s, int64_t batch, int n, const Mat (&B)[2], int s0,
                           const InvScratch& scr, int* info, const int* mask) {
  constexpr int NB = 32;
  const int npan = (n + NB - 1) / NB, tn = (n + 63) / 64;
  auto panel_args = [&](int p) {
    const int c = (s0 + p) & 1, k0 = p * NB;
    return PanelArgs{n, k0, min(NB, n - k0), B[c].p, B[c].stride, B[c].ld,
                     B[1 - c].p, B[1 - c].stride, B[1 - c].ld, scr.ints, scr.int_stride,
                     info, mask, p & 1};
  };
  launch_panel_cg<T, NB>(s, batch, n, panel_args(0));
  for (int p = 0; p < npan; ++p) {
    const PanelArgs pa = panel_args(p);
    UpdArgs ua{n, pa.k0, pa.nbw, tn * tn, pa.Xc, pa.sXc, pa.ldc, pa.Xn, pa.sXn, pa.ldn,
               scr.ints, scr.int_stride, mask, p & 1, 0, 0, 0};
    if (p + 1 == npan) {
      launch_update<T>(s, batch, ua);
      break;
    }
    // the column tile of panel p + 1 first
    const int cx = (pa.k0 + NB) / 64;
    UpdArgs un = ua;
    un.cfirst = cx;
    un.ncols = 1;
    un.tiles = tn;
    launch_update<T>(s, batch, un);
    // then panel p + 1 beside the other column tiles (n <= 64 has none:
    // total = 0 and the launch is the panel alone)
    UpdArgs ur = ua;
    ur.ncols = tn - 1;
    ur.cskip1 = cx + 1;
    ur.tiles = tn * (tn - 1);
    const int64_t total = batch * ur.tiles;
    const int np = (int)((batch + 7) / 8 * 8);
    const int64_t blocks = np + (total + 1) / 2;
    QX_REQUIRE(blocks < (int64_t(1) << 31), "inv: grid too large");
    const PanelArgs pn = panel_args(p + 1);
    if (n <= 256)
      hipLaunchKernelGGL((inv_step<T, 4>), dim3((unsigned)blocks), dim3(512), 0, s, pn, ur,
                         (int)batch, np, total);
    else
      hipLaunchKernelGGL((inv_step<T, 8>), dim3((unsigned)blocks), dim3(512), 0, s, pn, ur,
                         (int)batch, np, total);
  }
}

template <typename T, int NB, int RPT>
void run_inverse(hipStream_t s, int64_t batch, int n, const Mat& X, const InvScratch& scr,
                 int s0, int* info, const int* mask) {
  const int kind = inv_panel_kind();
  const int npan = (n + NB - 1) / NB;
  const Mat B[2] = {X, Mat(scr.buf, scr.buf_stride, n)};
  int cur = s0;
  const int tn = (n + 63) / 64;
  const bool la = NB == 32 && kind == 0 && inv_group_width() == 4 && npan > 1 &&
                  !narrow_update(batch) && inv_lookahead();
  if (la) run_inverse_lookahead<T>(s, batch, n, B, s0, scr, info, mask);
  for (int p = 0; !la && p < npan; ++p) {
    const int k0 = p * NB, nbw = min(NB, n - k0), nxt = 1 - cur;
    PanelArgs pa{n, k0, nbw, B[cur].p, B[cur].stride, B[cur].ld,
                 B[nxt].p, B[nxt].stride, B[nxt].ld, scr.ints, scr.int_stride, info, mask};
    if constexpr (NB == 64) {
      launch_panel_cg<T, 64>(s, batch, n, pa);
    } else {
      if (NB == 32 && kind == 0)
        launch_panel_cg<T, 32>(s, batch, n, pa);
      else if (NB == 32 && RPT == 1 && kind == 4)
        hipLaunchKernelGGL((inv_panel4<T>), dim3((unsigned)batch), dim3(NT4), 0, s, pa);
      else
        hipLaunchKernelGGL((inv_panel<T, NB, RPT>), dim3((unsigned)batch), dim3(NT), 0, s, pa);
    }
    if (n > nbw) {
      UpdArgs ua{n, k0, nbw, tn * tn, B[cur].p, B[cur].stride, B[cur].ld,
                 B[nxt].p, B[nxt].stride, B[nxt].ld, scr.ints, scr.int_stride, mask};
      const int64_t blocks = batch * tn * tn;
      QX_REQUIRE(blocks < (int64_t(1) << 31), "inv: grid too large");
      if (narrow_update(batch))
        hipLaunchKernelGGL((inv_update<T, NB, 16, 4>), dim3((unsigned)blocks), dim3(NT), 0, s, ua);
      else
        hipLaunchKernelGGL((inv_update<T, NB, 32, 1>), dim3((unsigned)blocks), dim3(NT), 0, s, ua);
    }
    cur = nxt;
  }
  // the parity of s0 leaves the reduced matrix in the scratch buffer B[1]
  const int64_t total = batch * (int64_t)n * n;
  int64_t blocks = (total + NT - 1) / NT;
  if (blocks > 65536 * 4) blocks = 65536 * 4;
  hipLaunchKernelGGL(inv_gather, dim3((unsigned)blocks), dim3(NT), 0, s, n, B[1].p,
                     B[1].stride, B[1].ld, X.p, X.stride, X.ld, scr.ints, scr.int_stride, total,
                     mask);
}

}  // namespace

// Panel width: 32 columns; with the older panel kernels 16 above n = 256
// (where two rows per thread double their per-column work).  QX_INV_NB=16|32|64
// overrides for n <= 256 (16 runs the one-thread-per-row kernel; 64 the
// column-group panel of eight waves of eight columns, whose rank-64 updates
// pass over the matrix half as often).
int inv_panel_width(int n) {
  const char* e = std::getenv("QX_INV_NB");  // read per call
  const int v = e ? std::atoi(e) : 0;
  const int forced_nb = (v == 16 || v == 32 || v == 64) ? v : 0;
  if (n > 256) return inv_panel_kind() == 0 ? 32 : 16;
  if (forced_nb == 64 && inv_panel_kind() != 0) return 32;
  return forced_nb ? forced_nb : 32;
}

Mat zinv_input(int n, const Mat& X, const InvScratch& scr) {
  const int NB = inv_panel_width(n);
  const int npan = (n + NB - 1) / NB;
  // the panels ping-pong between X and the scratch buffer, the last one
  // writing the scratch buffer: an odd count starts in X, an even one in it
  return (npan % 2) ? X : Mat(scr.buf, scr.buf_stride, n);
}

void zinv(hipStream_t s, int64_t batch, int n, const Mat& A, const Mat& S, const Mat& X,
          int* info, int precision, const InvScratch& scr, const int* mask) {
  QX_REQUIRE(n >= 1 && n <= kInvMaxN, "inv: block size must be in [1, 512]");
  QX_REQUIRE(!A.zero() && !X.zero(), "inv: null operand");
  QX_REQUIRE(scr.buf && scr.ints && scr.buf_stride >= (int64_t)n * n &&
                 scr.int_stride >= 3 * (int64_t)n,
             "inv: scratch too small");
  QX_REQUIRE(X.p != scr.buf, "inv: output aliases the scratch buffer");
  QX_REQUIRE(precision == QX_PREC_C128 || precision == QX_PREC_C64 ||
                 precision == QX_PREC_C128_3M,
             "inv: precision must be QX_PREC_C128, QX_PREC_C64 or QX_PREC_C128_3M");
  if (batch <= 0) return;
  QX_REQUIRE(batch < (int64_t(1) << 31), "inv: batch too large");
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);

  // n <= 32 in complex128: one wave per matrix, one launch
  if (precision != QX_PREC_C64 && inv_small(s, batch, n, A, S, X, info, mask)) {
    if (timer) timer->end(s, ev0, KernelTimer::INV, 8.0 * (double)n * n * n * batch);
    return;
  }

  const int NB = inv_panel_width(n);
  const bool big = n > 256 && NB == 16;
  const int npan = (n + NB - 1) / NB;
  // start in the buffer that makes the last panel write the scratch buffer
  const int s0 = 1 ^ (npan % 2);
  const Mat W(scr.buf, scr.buf_stride, n);
  const Mat start = s0 == 0 ? X : W;
  const bool in_place =
      (A.p == start.p) && S.zero() && A.ld == start.ld && A.stride == start.stride;
  if (!in_place) {
    if (mask) {
      const int64_t total = batch * (int64_t)n * n;
      int64_t blocks = (total + NT - 1) / NT;
      if (blocks > 65536 * 4) blocks = 65536 * 4;
      hipLaunchKernelGGL(inv_masked_start, dim3((unsigned)blocks), dim3(NT), 0, s, n, A.p,
                         A.stride, A.ld, S.p, S.stride, S.ld, start.p, start.stride, start.ld,
                         total, mask);
    } else {
      ew_axpby(s, batch, n, n, Cplx{1.0, 0.0}, A, Cplx{-1.0, 0.0}, S, start);
    }
  }

  if (precision == QX_PREC_C64) {
    if (big) run_inverse<float, 16, 2>(s, batch, n, X, scr, s0, info, mask);
    else if (NB == 16) run_inverse<float, 16, 1>(s, batch, n, X, scr, s0, info, mask);
    else if (NB == 64) run_inverse<float, 64, 1>(s, batch, n, X, scr, s0, info, mask);
    else run_inverse<float, 32, 1>(s, batch, n, X, scr, s0, info, mask);
  } else {
    if (big) run_inverse<double, 16, 2>(s, batch, n, X, scr, s0, info, mask);
    else if (NB == 16) run_inverse<double, 16, 1>(s, batch, n, X, scr, s0, info, mask);
    else if (NB == 64) run_inverse<double, 64, 1>(s, batch, n, X, scr, s0, info, mask);
    else run_inverse<double, 32, 1>(s, batch, n, X, scr, s0, info, mask);
  }
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::INV, 8.0 * (double)n * n * n * batch);
}

}  // namespace qx
