// Mode classification of the Spectral open-boundary condition, fused: for
// every eigenpair (w_j, v_j) of the lead's quadratic eigenproblem
// (a_ji / w + a_ii + a_ij w) v = 0, given the products a_n V, one thread
// computes
//   the residual   || sum_n (a_n v_j) w_j^n || / || v_j ||  (/ |w_j| when
//                  normalised)                      (spectral.py:198-243),
//   the group velocity dE/dk_j = -sum_{n != 0} i n (v_j^H a_n v_j) w_j^n
//                                                   (spectral.py:160-196),
//   k_j = -i log w_j, and the reflected-mode test of spectral.py:263-320:
//   propagating (|Im k| < min_decay, Re dE/dk < 0 and |Re dE/dk| /
//   (|Im dE/dk| + eps) above min_propagation) or decaying (Im k below
//   -min_decay, or below -eta_decay for a mode that does not propagate;
//   above -max_decay), with a residual below the tolerance.
// NaN and infinities in dE/dk and k count as zero, as in the reference.
// V and the products are (n, b, m) row-major: thread (e, j) walks column j,
// so the lanes of a wave read consecutive columns of one row.
#include <cfloat>
#include <cmath>

#include "qx_internal.h"

namespace qx {
namespace {

struct ModeArgs {
  int64_t n;
  int b, m;
  const z2* ws;
  const z2* vs;
  const z2* am;  // a_ji V (power -1)
  const z2* a0;  // a_ii V
  const z2* ap;  // a_ij V (power +1)
  double min_decay, max_decay, min_prop, res_tol, eta_decay;
  int res_norm;
  unsigned char* mask;
  double* res;  // nullable
  z2* dk;       // nullable
};

__device__ inline double finite_or_zero(double x) { return isfinite(x) ? x : 0.0; }

__global__ __launch_bounds__(256) void spectral_modes_kernel(ModeArgs g) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= g.n * g.m) return;
  const int64_t e = t / g.m;
  const int j = (int)(t - e * g.m);
  const z2 w = g.ws[t];
  double wir, wii;  // 1 / w
  complex_recip(w.x, w.y, wir, wii);
  const int64_t base = e * (int64_t)g.b * g.m + j;
  double vn = 0.0, pn = 0.0, qmr = 0.0, qmi = 0.0, qpr = 0.0, qpi = 0.0;
  for (int r = 0; r < g.b; ++r) {
    const int64_t o = base + (int64_t)r * g.m;
    const z2 v = g.vs[o], x = g.am[o], y = g.a0[o], z = g.ap[o];
    vn = fma(v.x, v.x, fma(v.y, v.y, vn));
    // x / w + y + z w
    const double pr = (x.x * wir - x.y * wii) + y.x + (z.x * w.x - z.y * w.y);
    const double pi = (x.x * wii + x.y * wir) + y.y + (z.x * w.y + z.y * w.x);
    pn = fma(pr, pr, fma(pi, pi, pn));
    // conj(v) x, conj(v) z
    qmr = fma(v.x, x.x, fma(v.y, x.y, qmr));
    qmi = fma(v.x, x.y, fma(-v.y, x.x, qmi));
    qpr = fma(v.x, z.x, fma(v.y, z.y, qpr));
    qpi = fma(v.x, z.y, fma(-v.y, z.x, qpi));
  }
  const double wabs = hypot(w.x, w.y);
  double res = sqrt(pn) / sqrt(vn);
  if (g.res_norm) res /= wabs;
  // dE/dk = i qm / w - i qp w
  const double tmr = qmr * wir - qmi * wii, tmi = qmr * wii + qmi * wir;
  const double tpr = qpr * w.x - qpi * w.y, tpi = qpr * w.y + qpi * w.x;
  const double dkr = finite_or_zero(-tmi + tpi), dki = finite_or_zero(tmr - tpr);
  // Im k for k = -i log w = arg w - i log |w|
  const double ki = finite_or_zero(-log(wabs));
  const double ratio = fabs(dkr) / (fabs(dki) + DBL_EPSILON);
  const bool prop = fabs(ki) < g.min_decay && g.min_prop < ratio && dkr < 0.0;
  bool dec = ki < -g.min_decay || (g.min_prop >= ratio && ki < -g.eta_decay);
  dec = dec && ki > -g.max_decay;
  g.mask[t] = ((prop || dec) && res < g.res_tol) ? 1 : 0;
  if (g.res) g.res[t] = res;
  if (g.dk) g.dk[t] = make_double2(dkr, dki);
}

}  // namespace

void spectral_modes(hipStream_t s, int64_t n, int b, int m, const z2* ws, const z2* vs,
                    const z2* am, const z2* a0, const z2* ap, double min_decay,
                    double max_decay, double min_propagation, double residual_tolerance,
                    double eta_decay, bool residual_normalization, unsigned char* mask,
                    double* res, z2* dk) {
  QX_REQUIRE(n >= 0 && b >= 1 && m >= 1, "spectral_modes: bad sizes");
  QX_REQUIRE(ws && vs && am && a0 && ap && mask, "spectral_modes: null argument");
  if (n == 0) return;
  ModeArgs g{n, b, m, ws, vs, am, a0, ap, min_decay, max_decay, min_propagation,
             residual_tolerance, eta_decay, residual_normalization ? 1 : 0, mask, res, dk};
  const int64_t blocks = (n * m + 255) / 256;
  QX_REQUIRE(blocks < (int64_t(1) << 31), "spectral_modes: too many modes");
  hipLaunchKernelGGL(spectral_modes_kernel, dim3((unsigned)blocks), dim3(256), 0, s, g);
  QX_HIP_CHECK(hipGetLastError());
}

}  // namespace qx
