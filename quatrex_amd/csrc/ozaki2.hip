// Complex128 GEMM on int8 matrix cores by modular arithmetic (Ozaki scheme II:
// Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique").
//
// C = alpha op(A) op(B) + beta (Cin [+ E - E^H])  [, Ct = -C^H]
//
// 1. Scaling (oz2_split_kernel).  Every line (row of op(A), column of op(B))
//    gets a power of two 2^e with max(|re|, |im|) < 2^e over the line, and its
//    entries become integers a' = rint(a 2^(T - e)), |a'| <= 2^T, exact in
//    FP64 for T <= 52.  The truncation error is at most 2^(-T-1) of the
//    line's largest entry.
// 2. Residues.  For pairwise coprime moduli m_1..m_N <= 256 (product M) the
//    symmetric residues of a'_re and a'_im modulo m_l are int8 values; they
//    are written as planes [l][re, im][line][k_pad], k contiguous and
//    zero-padded to a multiple of 64.
// 3. Products (oz2_mod_kernel).  Per modulus the four real products of the
//    complex product, Re Re, Im Im, Re Im and Im Re, run on
//    v_mfma_i32_16x16x64_i8 with exact int32 accumulation (|P| <= k 2^14);
//    modulo m_l, Re C' = P_rr - P_ii and Im C' = P_ri + P_ir.  (The
//    Karatsuba form would save one product in four, but its third operand
//    plane costs more in residue traffic than the int8 matrix cores save.)
// 4. Chinese remainder reconstruction (oz2_crt_kernel).  With
//    y_l = (M / m_l)^(-1) mod m_l, the exact integer product satisfies
//    C' / M = sum_l v_l / m_l (mod 1), v_l = (x_l y_l) mod m_l, x_l the
//    residue of C' from step 3.  The sum is formed in FP64 (each term in
//    [-1/2, 1/2], brought back to [-1/2, 1/2] every four moduli), its
//    fractional part f in [-1/2, 1/2) gives C' = f M, and
//    C = C' 2^(e_row + e_col - 2T).  T is the largest integer with
//    4 k_pad 2^(2T) < M, so |Re C'|, |Im C'| <= 2 k 2^(2T) < M / 2 and C' is
//    recovered uniquely.
//
// Error: the truncation of step 1 (relative 2^-T per line, e.g. 2^-42 with
// 12 moduli at k = 256) dominates; the FP64 reconstruction adds ~N 2^-52 of
// M, i.e. below 2^-44 of the bound 2 k 2^(2T).
//
// MFMA 16x16x64 i8 operands: lane l passes 16 int8 values of row l % 16 of A
// (column l % 16 of B) at the same 16 values of k, k-group l / 16; result
// element e of lane l is D[4 (l / 16) + e][l % 16].
#include <cmath>
#include <cstdint>

#include "qx_internal.h"

namespace qx {
namespace {

constexpr int kOzMaxModuli = 20;
// pairwise coprime, largest first: 2^8, 3*5*17, 11*23, 251, 13*19, then
// primes and 7*31
constexpr int kModuli[kOzMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                       223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

struct OzConsts {
  int nmod;
  int T;           // integer bits per scaled entry
  double M;        // product of the moduli (rounded to FP64)
  double m[kOzMaxModuli];
  double inv_m[kOzMaxModuli];
  double y[kOzMaxModuli];  // (M / m_l)^-1 mod m_l
  int lo[kOzMaxModuli];    // symmetric residue range [lo, lo + m - 1]
};

typedef int v4i __attribute__((ext_vector_type(4)));

// Symmetric residue of the integer-valued double x modulo m, in [lo, lo + m - 1].
__device__ __forceinline__ int sym_mod(double x, double m, double inv_m, int lo) {
  const double q = rint(x * inv_m);
  int r = (int)fma(-q, m, x);
  const int mi = (int)m;
  r = r > lo + mi - 1 ? r - mi : r;
  r = r < lo ? r + mi : r;
  return r;
}

// One workgroup per 8 lines of one matrix, covering the whole of k in
// chunks of 256.  Thread t owns one line and 8 consecutive values of k per
// chunk: for k-contiguous storage of the lines (element (line, k) at
// X[line * ld + k]) line t / 32 and k offset 8 (t % 32), so the 32 threads of
// a line read and write 256 contiguous values; for line-contiguous storage
// (element at X[k * ld + line]) line t % 8 and k offset 8 (t / 8), so 8
// consecutive threads read 8 consecutive lines.  The first pass finds each
// line's largest magnitude, the second writes the residues.
__global__ __launch_bounds__(256) void oz2_split_kernel(const z2* X, int64_t sX, int ld, int trans,
                                                        int conj, int lines, int k, int k_pad,
                                                        OzConsts c, int8_t* res, int64_t s_res,
                                                        int* expo, int64_t s_expo) {
  constexpr int LPW = 8, EPT = 8;
  __shared__ double red[256];
  const int groups = (lines + LPW - 1) / LPW;
  const int64_t e = blockIdx.x / groups;
  const int line0 = (int)(blockIdx.x - e * groups) * LPW;
  const int t = threadIdx.x;
  const int ll = trans ? (t % LPW) : (t / (256 / LPW));
  const int kq = EPT * (trans ? (t / LPW) : (t % (256 / LPW)));
  const int line = line0 + ll;
  const bool valid = line < lines;
  const z2* Xe = X + e * sX;
  auto load = [&](int kk) -> z2 {
    if (!valid || kk >= k) return make_double2(0.0, 0.0);
    return trans ? Xe[(int64_t)kk * ld + line] : Xe[(int64_t)line * ld + kk];
  };
  double mx = 0.0;
  for (int kc = 0; kc < k_pad; kc += 256) {
    if (kc + kq >= k_pad) continue;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const z2 v = load(kc + kq + j);
      mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
    }
  }
  red[t] = mx;
  __syncthreads();
  // the 32 threads of this line
  mx = 0.0;
#pragma unroll
  for (int q = 0; q < 256 / LPW; ++q)
    mx = fmax(mx, red[trans ? (ll + LPW * q) : ((256 / LPW) * ll + q)]);
  // smallest ex with mx < 2^ex (0 for an all-zero line)
  const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
  if (valid && kq == 0) expo[e * s_expo + line] = ex;
  if (!valid) return;
  const int sh = c.T - ex;
  const double sgn = conj ? -1.0 : 1.0;
  const int64_t plane = (int64_t)lines * k_pad;
  int8_t* R = res + e * s_res + (int64_t)line * k_pad;
  typedef int v2i __attribute__((ext_vector_type(2)));
  for (int kc = 0; kc < k_pad; kc += 256) {
    if (kc + kq >= k_pad) continue;
    double ar[EPT], ai[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const z2 v = load(kc + kq + j);
      ar[j] = rint(ldexp(v.x, sh));
      ai[j] = rint(ldexp(sgn * v.y, sh));
    }
    for (int l = 0; l < c.nmod; ++l) {
      const double m = c.m[l], im = c.inv_m[l];
      const int lo = c.lo[l];
      v2i pr, pi;
#pragma unroll
      for (int w = 0; w < EPT / 4; ++w) {
        unsigned br = 0, bi = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          br |= (unsigned)(sym_mod(ar[4 * w + b], m, im, lo) & 0xFF) << (8 * b);
          bi |= (unsigned)(sym_mod(ai[4 * w + b], m, im, lo) & 0xFF) << (8 * b);
        }
        pr[w] = (int)br;
        pi[w] = (int)bi;
      }
      const int64_t o = (int64_t)(2 * l) * plane + kc + kq;
      *reinterpret_cast<v2i*>(R + o) = pr;
      *reinterpret_cast<v2i*>(R + o + plane) = pi;
    }
  }
}

struct Oz2Args {
  int m, n, k_pad;
  double ar, ai, br, bi;
  const int8_t* Ra;  // [batch][l][2][m][k_pad]
  int64_t sRa;
  const int8_t* Rb;  // [batch][l][2][n][k_pad]
  int64_t sRb;
  const int* ea;  // [batch][m]
  const int* eb;  // [batch][n]
  z2* C;
  int64_t sC;
  int ldc;
  const z2* Cin;
  int64_t sCin;
  int ldcin;
  const z2* E;  // optional: beta term Cin + E - E^H (m == n)
  int64_t sE;
  int lde;
  z2* Ct;  // optional second output -C^H
  int64_t sCt;
  int ldct;
  int tilesM, tilesN;
  int beta_zero;
  int8_t* V;  // [matrix][tile][modulus][re, im][4096] residues of C'
  OzConsts c;
};

__device__ inline int oz2_xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx, rem = nblocks % nx;
  const int x = bid % nx, j = bid / nx;
  return x * per + min(x, rem) + j;
}

// Residues of the product modulo one modulus per workgroup: a 64 x 64
// output tile of one matrix, 4 waves (2 x 2, 32 x 32 each: 2 x 2 MFMA tiles),
// grid (matrix, tile, modulus).  Operand fragments are read straight from
// global memory (L2), up to four 64-wide k-steps ahead of the products that
// use them; the many short workgroups of the grid hide the rest of the
// latency.  NK = k_pad / 64 is a template parameter (1..8).  The
// residues v_l = (x_l y_l) mod m_l of Re C' and Im C' leave as int8 in the
// MFMA result order: per (matrix, tile, modulus, re / im) 256 threads x 16
// values, one 16-byte store per thread.
template <int NK>
__global__ __launch_bounds__(256) void oz2_mod_kernel(Oz2Args g) {
  const int tiles = g.tilesM * g.tilesN;
  const int64_t id = oz2_xcd_remap(blockIdx.x, gridDim.x);
  const int l = (int)(id % g.c.nmod);
  const int64_t et = id / g.c.nmod;  // matrix * tiles + tile
  const int64_t e = et / tiles;
  const int t = (int)(et - e * tiles);
  const int m0 = (t / g.tilesN) * 64, n0 = (t % g.tilesN) * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int kg = 16 * (lane >> 4);
  const int64_t planeA = (int64_t)g.m * g.k_pad, planeB = (int64_t)g.n * g.k_pad;
  // wave-uniform plane bases and 32-bit per-lane offsets (one address
  // register per fragment); planes hold < 2^31 bytes (k_pad <= 512)
  const int8_t* Ar = g.Ra + e * g.sRa + (int64_t)(2 * l) * planeA;
  const int8_t* Ai = Ar + planeA;
  const int8_t* Br = g.Rb + e * g.sRb + (int64_t)(2 * l) * planeB;
  const int8_t* Bi = Br + planeB;
  // fragment rows / columns, clamped into the matrix (ragged edges)
  int offA[2], offB[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int r = min(m0 + wm * 32 + i * 16 + (lane & 15), g.m - 1);
    const int cc = min(n0 + wn * 32 + i * 16 + (lane & 15), g.n - 1);
    offA[i] = r * g.k_pad + kg;
    offB[i] = cc * g.k_pad + kg;
  }
  auto fetch = [&](int kk, v4i (&a)[2][2], v4i (&b)[2][2]) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[0][i] = *reinterpret_cast<const v4i*>(Ar + offA[i] + 64 * kk);
      a[1][i] = *reinterpret_cast<const v4i*>(Ai + offA[i] + 64 * kk);
      b[0][i] = *reinterpret_cast<const v4i*>(Br + offB[i] + 64 * kk);
      b[1][i] = *reinterpret_cast<const v4i*>(Bi + offB[i] + 64 * kk);
    }
  };
  v4i prr[2][2], pii[2][2], pri[2][2], pir[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) prr[i][j] = pii[i][j] = pri[i][j] = pir[i][j] = v4i{0, 0, 0, 0};
  auto step = [&](const v4i (&a)[2][2], const v4i (&b)[2][2]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        prr[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[0][i], b[0][j], prr[i][j], 0, 0, 0);
        pii[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[1][i], b[1][j], pii[i][j], 0, 0, 0);
        pri[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[0][i], b[1][j], pri[i][j], 0, 0, 0);
        pir[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[1][i], b[0][j], pir[i][j], 0, 0, 0);
      }
  };
  // NK = k_pad / 64 k-steps, in groups of up to four whose loads are all
  // issued before the first of their products (straight-line code, so each
  // product waits for exactly its own fragments)
#pragma unroll
  for (int k0 = 0; k0 < NK; k0 += 4) {
    constexpr int G = NK < 4 ? NK : 4;
    v4i fa[G][2][2], fb[G][2][2];
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (k0 + q < NK) fetch(k0 + q, fa[q], fb[q]);
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (k0 + q < NK) step(fa[q], fb[q]);
  }
  const double m = g.c.m[l], im = g.c.inv_m[l], y = g.c.y[l];
  v4i vr, vi;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      unsigned br = 0, bi = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double tr = (double)(prr[i][j][q] - pii[i][j][q]) * y;
        const double ti = (double)(pri[i][j][q] + pir[i][j][q]) * y;
        // symmetric residues, |v| <= m / 2 <= 128 (the 128 of m = 256 wraps
        // to -128, the same residue)
        const int ir = (int)fma(-rint(tr * im), m, tr), ii = (int)fma(-rint(ti * im), m, ti);
        br |= (unsigned)(ir & 0xFF) << (8 * q);
        bi |= (unsigned)(ii & 0xFF) << (8 * q);
      }
      vr[2 * i + j] = (int)br;
      vi[2 * i + j] = (int)bi;
    }
  int8_t* V = g.V + ((et * g.c.nmod + l) * 2) * 4096;
  reinterpret_cast<v4i*>(V)[threadIdx.x] = vr;
  reinterpret_cast<v4i*>(V + 4096)[threadIdx.x] = vi;
}

// Chinese remainder reconstruction and epilogue: one workgroup per 64 x 64
// output tile, each thread owning the 16 elements it held in oz2_mod_kernel.
__global__ __launch_bounds__(256) void oz2_crt_kernel(Oz2Args g) {
  const int tiles = g.tilesM * g.tilesN;
  const int64_t et = oz2_xcd_remap(blockIdx.x, gridDim.x);
  const int64_t e = et / tiles;
  const int t = (int)(et - e * tiles);
  const int m0 = (t / g.tilesN) * 64, n0 = (t % g.tilesN) * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w >> 1, wn = w & 1;
  double sr[16], si[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) sr[q] = si[q] = 0.0;
  const int8_t* V = g.V + et * g.c.nmod * 2 * 4096;
  for (int l = 0; l < g.c.nmod; ++l) {
    const v4i vr = reinterpret_cast<const v4i*>(V + (2 * l) * 4096)[threadIdx.x];
    const v4i vi = reinterpret_cast<const v4i*>(V + (2 * l + 1) * 4096)[threadIdx.x];
    const double im = g.c.inv_m[l];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int br = (int)(signed char)((unsigned)vr[q >> 2] >> (8 * (q & 3)));
      const int bi = (int)(signed char)((unsigned)vi[q >> 2] >> (8 * (q & 3)));
      sr[q] = fma((double)br, im, sr[q]);
      si[q] = fma((double)bi, im, si[q]);
    }
    if ((l & 3) == 3) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        sr[q] -= rint(sr[q]);
        si[q] -= rint(si[q]);
      }
    }
  }
  const int* ea = g.ea + e * g.m;
  const int* eb = g.eb + e * g.n;
  z2* C = g.C + e * g.sC;
  const z2* Ci = g.Cin + e * g.sCin;
  const z2* Ee = g.E ? g.E + e * g.sE : nullptr;
  z2* Ct = g.Ct ? g.Ct + e * g.sCt : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
        const int x = 4 * (2 * i + j) + q;
        if (row < g.m && col < g.n) {
          const double fr = sr[x] - rint(sr[x]);
          const double fi = si[x] - rint(si[x]);
          const int ex = ea[row] + eb[col] - 2 * g.c.T;
          const double vr = ldexp(fr * g.c.M, ex), vi = ldexp(fi * g.c.M, ex);
          double outr = fma(g.ar, vr, -g.ai * vi), outi = fma(g.ar, vi, g.ai * vr);
          if (!g.beta_zero) {
            z2 s = Ci[(int64_t)row * g.ldcin + col];
            if (Ee) {
              const z2 a = Ee[(int64_t)row * g.lde + col], b = Ee[(int64_t)col * g.lde + row];
              s = make_double2((a.x - b.x) + s.x, (a.y + b.y) + s.y);
            }
            outr = fma(g.br, s.x, fma(-g.bi, s.y, outr));
            outi = fma(g.br, s.y, fma(g.bi, s.x, outi));
          }
          C[(int64_t)row * g.ldc + col] = make_double2(outr, outi);
          if (Ct) Ct[(int64_t)col * g.ldct + row] = make_double2(-outr, outi);
        }
      }
}

int64_t pad64(int k) { return ((int64_t)k + 63) / 64 * 64; }

// Constants for `nmod` moduli at inner dimension k.
OzConsts make_consts(int nmod, int k) {
  OzConsts c{};
  c.nmod = nmod;
  double log2M = 0.0, M = 1.0;
  for (int l = 0; l < nmod; ++l) {
    const int m = kModuli[l];
    c.m[l] = m;
    c.inv_m[l] = 1.0 / m;
    c.lo[l] = -(m / 2);
    log2M += std::log2((double)m);
    M *= m;
    // (M / m_l) mod m_l, then its inverse modulo m_l (m_l < 2^9: brute force)
    int64_t p = 1;
    for (int j = 0; j < nmod; ++j)
      if (j != l) p = (p * kModuli[j]) % m;
    int inv = 0;
    for (int v = 1; v < m; ++v)
      if ((p * v) % m == 1) {
        inv = v;
        break;
      }
    QX_REQUIRE(inv > 0, "ozaki2: moduli are not coprime");
    c.y[l] = inv;
  }
  c.M = M;
  // largest T with 4 k_pad 2^(2T) < M
  const double kp = (double)pad64(k);
  c.T = (int)std::floor((log2M - std::log2(4.0 * kp)) / 2.0 - 1e-9);
  QX_REQUIRE(c.T >= 8 && c.T <= 52, "ozaki2: moduli give no usable integer width at this k");
  return c;
}

}  // namespace

size_t oz2_workspace_bytes(int64_t batch, int m, int n, int k, int nmod) {
  const int64_t kp = pad64(k);
  const int64_t tiles = (int64_t)((m + 63) / 64) * ((n + 63) / 64);
  return (size_t)(batch * 2 * nmod * (int64_t)(m + n) * kp) +
         (size_t)(batch * (m + n)) * sizeof(int) + 256 +
         (size_t)(batch * tiles * nmod * 2 * 4096) + 256;
}

void zgemm_oz2(hipStream_t s, int64_t batch, int m, int n, int k, int opA, int opB, Cplx alpha,
               Cplx beta, const GemmOperands& x, int nmod, void* ws, size_t ws_bytes) {
  QX_REQUIRE(!x.C.zero(), "zgemm_oz2: output is a zero matrix");
  QX_REQUIRE(opA == QX_OP_N || opA == QX_OP_C, "zgemm_oz2: bad op_a");
  QX_REQUIRE(opB == QX_OP_N || opB == QX_OP_C, "zgemm_oz2: bad op_b");
  QX_REQUIRE(nmod >= 8 && nmod <= kOzMaxModuli, "zgemm_oz2: moduli count must be in [8, 20]");
  QX_REQUIRE(x.E.zero() || (m == n && !x.Cin.zero()), "zgemm_oz2: E needs m == n and Cin");
  if (batch <= 0 || m <= 0 || n <= 0) return;
  const bool beta_zero = (beta.re == 0.0 && beta.im == 0.0) || x.Cin.zero();
  if (x.A.zero() || x.B.zero() || k <= 0) {
    // op(A) op(B) == 0: C = beta Cin (Cin + E - E^H)
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
  QX_REQUIRE(k <= 512, "zgemm_oz2: k must be <= 512");
  QX_REQUIRE(ws_bytes >= oz2_workspace_bytes(batch, m, n, k, nmod),
             "zgemm_oz2: workspace too small");
  const OzConsts c = make_consts(nmod, k);
  const int k_pad = (int)pad64(k);
  int8_t* Ra = static_cast<int8_t*>(ws);
  const int64_t sRa = (int64_t)2 * nmod * m * k_pad, sRb = (int64_t)2 * nmod * n * k_pad;
  int8_t* Rb = Ra + batch * sRa;
  int* ea = reinterpret_cast<int*>(Rb + batch * sRb);
  int* eb = ea + batch * m;
  // residues of C', 256-byte aligned after the exponents
  const uintptr_t vbase = (reinterpret_cast<uintptr_t>(eb + batch * n) + 255) & ~uintptr_t(255);
  int8_t* V = reinterpret_cast<int8_t*>(vbase);
  KernelTimer* timer = active_timer;
  hipEvent_t ev0 = nullptr;
  if (timer) timer->begin(s, &ev0);
  // line i of op(A): row i of A, or conj of column i of A for A^H;
  // line j of op(B): column j of B, or conj of row j of B for B^H
  const int64_t ga = batch * ((m + 7) / 8), gb = batch * ((n + 7) / 8);
  QX_REQUIRE(ga < (int64_t(1) << 31) && gb < (int64_t(1) << 31), "zgemm_oz2: grid too large");
  hipLaunchKernelGGL(oz2_split_kernel, dim3((unsigned)ga), dim3(256), 0, s, x.A.p, x.A.stride,
                     x.A.ld, opA == QX_OP_C ? 1 : 0, opA == QX_OP_C ? 1 : 0, m, k, k_pad, c, Ra,
                     sRa, ea, (int64_t)m);
  hipLaunchKernelGGL(oz2_split_kernel, dim3((unsigned)gb), dim3(256), 0, s, x.B.p, x.B.stride,
                     x.B.ld, opB == QX_OP_N ? 1 : 0, opB == QX_OP_C ? 1 : 0, n, k, k_pad, c, Rb,
                     sRb, eb, (int64_t)n);
  Oz2Args g{};
  g.m = m;
  g.n = n;
  g.k_pad = k_pad;
  g.ar = alpha.re;
  g.ai = alpha.im;
  g.br = beta.re;
  g.bi = beta.im;
  g.Ra = Ra;
  g.sRa = sRa;
  g.Rb = Rb;
  g.sRb = sRb;
  g.ea = ea;
  g.eb = eb;
  g.C = x.C.p;
  g.sC = x.C.stride;
  g.ldc = x.C.ld;
  const Mat& ci = x.Cin.zero() ? x.C : x.Cin;
  g.Cin = ci.p;
  g.sCin = ci.stride;
  g.ldcin = ci.ld;
  g.E = x.E.p;
  g.sE = x.E.stride;
  g.lde = x.E.ld;
  g.Ct = x.Ct.p;
  g.sCt = x.Ct.stride;
  g.ldct = x.Ct.ld;
  g.tilesM = (m + 63) / 64;
  g.tilesN = (n + 63) / 64;
  g.beta_zero = beta_zero ? 1 : 0;
  g.V = V;
  g.c = c;
  const int64_t blocks = batch * g.tilesM * g.tilesN;
  QX_REQUIRE(blocks * nmod < (int64_t(1) << 31), "zgemm_oz2: grid too large");
  const dim3 mgrid((unsigned)(blocks * nmod));
  switch (k_pad / 64) {
#define QX_OZ_MOD(NK)                                                   \
  case NK:                                                              \
    hipLaunchKernelGGL(oz2_mod_kernel<NK>, mgrid, dim3(256), 0, s, g);  \
    break;
    QX_OZ_MOD(1)
    QX_OZ_MOD(2)
    QX_OZ_MOD(3)
    QX_OZ_MOD(4)
    QX_OZ_MOD(5)
    QX_OZ_MOD(6)
    QX_OZ_MOD(7)
    QX_OZ_MOD(8)
#undef QX_OZ_MOD
    default:
      throw Error{QX_EINVAL, "zgemm_oz2: k must be <= 512"};
  }
  hipLaunchKernelGGL(oz2_crt_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
  if (timer) timer->end(s, ev0, KernelTimer::GEMM, 8.0 * m * n * (double)k * batch);
}

}  // namespace qx
