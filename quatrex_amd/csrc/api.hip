// C ABI of libquatrex_amd (declared in include/quatrex_amd.h), plus the
// open-boundary drivers (Sancho-Rubio decimation, Beyn contour moments).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <thread>

#include "qx_internal.h"

namespace qx {
size_t rgf_workspace_bytes(const qx_btd_layout* la, bool solve, int prec);
void rgf_selected_inv(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                      const void* const* obc_r, const qx_btd_layout* lx, void* x, int prec);
void rgf_selected_solve(qx_handle* h, hipStream_t st, const qx_btd_layout* la, const void* a,
                        const qx_btd_layout* lsl, const void* sl, const qx_btd_layout* lsg,
                        const void* sg, const void* const* obc_r, const void* const* obc_l,
                        const void* const* obc_g, const qx_btd_layout* lxl, void* xl,
                        const qx_btd_layout* lxg, void* xg, const qx_btd_layout* lxr, void* xr,
                        void* current, void* corner, double* density, double* ldos,
                        uint32_t flags, int prec);

namespace {

__global__ void fill_int_kernel(int* p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

const Cplx kOne{1.0, 0.0}, kMinus{-1.0, 0.0}, kZero{0.0, 0.0};

// Sancho-Rubio decimation (sancho_rubio.py:33-86).  Products are formed left
// to right as in the reference (alpha @ inverse @ beta == (alpha g) beta);
// the repeated alpha g beta of the reference is computed once.
int sancho_rubio(qx_handle* h, hipStream_t st, int64_t n, int b, const void* a_ii,
                 const void* a_ij, const void* a_ji, void* x_ii, int max_iter, double tol,
                 int* iters) {
  QX_REQUIRE(n >= 0 && b >= 1 && b <= kInvMaxN, "sancho_rubio: bad sizes");
  QX_REQUIRE(a_ii && a_ij && a_ji && x_ii, "sancho_rubio: null argument");
  if (iters) *iters = 0;
  if (n == 0) return QX_OK;
  const int64_t bb = (int64_t)b * b;
  const size_t slot = (size_t)(n * bb) * sizeof(z2);
  // 10 matrix slots, inverse scratch (buffer + pivot ints), norms
  z2* ws = static_cast<z2*>(h->ensure_ws(slot * 12 + (size_t)n * sizeof(double) * 2));
  auto M = [&](int k) { return Mat(ws + (int64_t)k * n * bb, bb, b); };
  const InvScratch scr{ws + 10 * n * bb, bb, reinterpret_cast<int*>(ws + 11 * n * bb), 4 * bb};
  double* dnorm = reinterpret_cast<double*>(ws + 12 * n * bb);
  int* info = h->ensure_info(n);
  const Mat Aii(a_ii, bb, b), Aij(a_ij, bb, b), Aji(a_ji, bb, b), X(x_ii, bb, b);
  int eps = 0, eps_s = 1, alpha = 2, beta = 3, g = 4, t1 = 5, t2 = 6, p = 7, na = 8, nb = 9;
  ew_axpby(st, n, b, b, kOne, Aii, kZero, Mat(), M(eps));
  ew_axpby(st, n, b, b, kOne, Aii, kZero, Mat(), M(eps_s));
  ew_axpby(st, n, b, b, kOne, Aji, kZero, Mat(), M(alpha));
  ew_axpby(st, n, b, b, kOne, Aij, kZero, Mat(), M(beta));
  std::vector<double> hn(n);
  bool converged = false;
  const int prec = QX_PREC_C128;
  for (int it = 0; it < max_iter; ++it) {
    if (iters) *iters = it + 1;
    zinv(st, n, b, M(eps), Mat(), M(g), info, prec, scr);
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kOne, M(alpha), M(g), kZero, M(t1), prec);  // a g
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kOne, M(beta), M(g), kZero, M(t2), prec);   // b g
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kOne, M(t1), M(beta), kZero, M(p), prec);   // a g b
    // eps -= a g b + b g a ; eps_s -= a g b
    ew_axpby(st, n, b, b, kOne, M(eps), kMinus, M(p), M(eps));
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kMinus, M(t2), M(alpha), kOne, M(eps), prec);
    ew_axpby(st, n, b, b, kOne, M(eps_s), kMinus, M(p), M(eps_s));
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kOne, M(t1), M(alpha), kZero, M(na), prec);  // a g a
    zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, kOne, M(t2), M(beta), kZero, M(nb), prec);   // b g b
    std::swap(alpha, na);
    std::swap(beta, nb);
    absnorm2(st, n, b, M(alpha), M(beta), dnorm);
    QX_HIP_CHECK(hipMemcpyAsync(hn.data(), dnorm, n * sizeof(double), hipMemcpyDeviceToHost, st));
    QX_HIP_CHECK(hipStreamSynchronize(st));
    const double delta = *std::max_element(hn.begin(), hn.end()) / 2.0;
    if (delta < tol) {
      converged = true;
      break;
    }
  }
  zinv(st, n, b, M(eps_s), Mat(), X, info, prec, scr);
  return converged ? QX_OK : QX_NOT_CONVERGED;
}

// Beyn contour moments (beyn.py:177-235 with operator.py:113-179).
int beyn_moments(qx_handle* h, hipStream_t st, int64_t n, int b, int nq, double r_o,
                 double r_i, const void* a_ji, const void* a_ii, const void* a_ij, void* q0,
                 void* q1, int prec) {
  QX_REQUIRE(n >= 0 && b >= 1 && b <= kInvMaxN && nq >= 1, "beyn: bad sizes");
  QX_REQUIRE(a_ji && a_ii && a_ij && q0 && q1, "beyn: null argument");
  if (n == 0) return QX_OK;
  const int64_t bb = (int64_t)b * b;
  // nodes and weights: z = r exp(2 pi i (k + 1.5) / nq), w = 1/nq
  std::vector<z2> host(6 * nq);
  z2 *zo = host.data(), *zi = zo + nq, *w0o = zi + nq, *w1o = w0o + nq, *w0i = w1o + nq,
     *w1i = w0i + nq;
  for (int k = 0; k < nq; ++k) {
    const double ph = 2.0 * M_PI * (k + 1.5) / nq;
    const double c = std::cos(ph), s = std::sin(ph), w = 1.0 / nq;
    zo[k] = make_double2(r_o * c, r_o * s);
    zi[k] = make_double2(r_i * c, r_i * s);
    w0o[k] = make_double2(w * zo[k].x, w * zo[k].y);
    w1o[k] = make_double2(w * (zo[k].x * zo[k].x - zo[k].y * zo[k].y), w * 2 * zo[k].x * zo[k].y);
    w0i[k] = make_double2(-w * zi[k].x, -w * zi[k].y);
    w1i[k] = make_double2(-w * (zi[k].x * zi[k].x - zi[k].y * zi[k].y), -w * 2 * zi[k].x * zi[k].y);
  }
  // energies per chunk: operator, inverse and inverse scratch stacks of
  // chunk * nq matrices each, plus the pivot ints
  const int64_t per_e = 4 * nq * bb;
  int64_t chunk = std::max<int64_t>(1, (int64_t(1) << 28) / per_e);  // <= 4 GiB
  chunk = std::min(chunk, n);
  const size_t dev_aux = 6 * nq * sizeof(z2);
  z2* ws = static_cast<z2*>(h->ensure_ws(dev_aux + (size_t)(chunk * per_e) * sizeof(z2)));
  z2* dz = ws;
  z2* T = ws + 6 * nq;
  z2* Ti = T + chunk * nq * bb;
  const InvScratch scr{Ti + chunk * nq * bb, bb, reinterpret_cast<int*>(Ti + 2 * chunk * nq * bb),
                       4 * bb};
  QX_HIP_CHECK(hipMemcpyAsync(dz, host.data(), dev_aux, hipMemcpyHostToDevice, st));
  int* info = h->ensure_info(chunk * nq);
  const Mat Q0(q0, bb, b), Q1(q1, bb, b);
  ew_axpby(st, n, b, b, kZero, Mat(), kZero, Mat(), Q0);
  ew_axpby(st, n, b, b, kZero, Mat(), kZero, Mat(), Q1);
  for (int64_t e0 = 0; e0 < n; e0 += chunk) {
    const int64_t ne = std::min(chunk, n - e0);
    const Mat Aji(static_cast<const z2*>(a_ji) + e0 * bb, bb, b);
    const Mat Aii(static_cast<const z2*>(a_ii) + e0 * bb, bb, b);
    const Mat Aij(static_cast<const z2*>(a_ij) + e0 * bb, bb, b);
    const Mat q0c(static_cast<z2*>(q0) + e0 * bb, bb, b), q1c(static_cast<z2*>(q1) + e0 * bb, bb, b);
    for (int side = 0; side < 2; ++side) {
      const z2* z = dz + side * nq;
      contour_operator(st, ne, nq, b, Aji, Aii, Aij, z, Mat(T, bb, b));
      zinv(st, ne * nq, b, Mat(T, bb, b), Mat(), Mat(Ti, bb, b), info, prec, scr);
      const z2* w0 = dz + (side == 0 ? 2 : 4) * nq;
      const z2* w1 = dz + (side == 0 ? 3 : 5) * nq;
      contour_accumulate(st, ne, nq, b, Mat(Ti, bb, b), w0, w1, q0c, q1c);
    }
  }
  return QX_OK;
}

}  // namespace

thread_local KernelTimer* active_timer = nullptr;
thread_local OzWorkspace active_oz_ws;
thread_local int active_gemm_tile = 0;

thread_local int split_part = 0;

void run_split(qx_handle* h, hipStream_t st, int64_t S,
               const std::function<void(hipStream_t, int64_t, int64_t)>& f) {
  int P = (int)std::min<int64_t>(h->parts, std::max<int64_t>(1, S / 4));
  if (P <= 1) {
    f(st, 0, S);
    return;
  }
  while ((int)h->streams.size() < P - 1) {
    hipStream_t s2;
    hipEvent_t ev;
    QX_HIP_CHECK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    QX_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    h->streams.push_back(s2);
    h->join.push_back(ev);
  }
  if (!h->fork) QX_HIP_CHECK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
  QX_HIP_CHECK(hipEventRecord(h->fork, st));
  for (int k = 1; k < P; ++k) QX_HIP_CHECK(hipStreamWaitEvent(h->streams[k - 1], h->fork, 0));
  std::vector<int64_t> e0(P + 1, 0);
  for (int k = 0; k < P; ++k) e0[k + 1] = e0[k] + S / P + (k < S % P ? 1 : 0);
  std::vector<std::exception_ptr> err(P);
  KernelTimer* timer = active_timer;
  const int tile = active_gemm_tile;
  std::vector<std::thread> threads;
  for (int k = 1; k < P; ++k) {
    threads.emplace_back([&, k]() {
      try {
        active_timer = timer;
        active_gemm_tile = tile;
        QX_HIP_CHECK(hipSetDevice(h->device));
        split_part = k;
        f(h->streams[k - 1], e0[k], e0[k + 1] - e0[k]);
      } catch (...) {
        err[k] = std::current_exception();
      }
      active_timer = nullptr;
    });
  }
  try {
    split_part = 0;
    f(st, 0, e0[1]);
  } catch (...) {
    err[0] = std::current_exception();
  }
  for (auto& t : threads) t.join();
  for (int k = 1; k < P; ++k) {
    QX_HIP_CHECK(hipEventRecord(h->join[k - 1], h->streams[k - 1]));
    QX_HIP_CHECK(hipStreamWaitEvent(st, h->join[k - 1], 0));
  }
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

hipEvent_t KernelTimer::get() {
  if (next == pool.size()) {
    hipEvent_t e;
    QX_HIP_CHECK(hipEventCreate(&e));
    pool.push_back(e);
  }
  return pool[next++];
}

void KernelTimer::begin(hipStream_t s, hipEvent_t* ev) {
  std::lock_guard<std::mutex> lock(mu);
  *ev = get();
  QX_HIP_CHECK(hipEventRecord(*ev, s));
}

void KernelTimer::end(hipStream_t s, hipEvent_t a, int kind, double flops) {
  std::lock_guard<std::mutex> lock(mu);
  hipEvent_t b = get();
  QX_HIP_CHECK(hipEventRecord(b, s));
  recs.push_back(Rec{a, b, kind, flops});
}

KernelTimer::~KernelTimer() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace qx

void* qx_handle::ensure_ws(size_t bytes) {
  if (bytes > ws_bytes) {
    if (ws) QX_HIP_CHECK(hipFree(ws));
    ws = nullptr;
    ws_bytes = 0;
    if (hipMalloc(&ws, bytes) != hipSuccess) {
      ws = nullptr;
      (void)hipGetLastError();
      throw qx::Error{QX_ENOMEM, "workspace allocation of " + std::to_string(bytes) + " bytes failed"};
    }
    ws_bytes = bytes;
  }
  return ws;
}

int* qx_handle::ensure_info(size_t n) {
  if (n > info_len) {
    const size_t len = std::max<size_t>(n, 1024);
    int* p = nullptr;
    if (hipMalloc(&p, len * sizeof(int)) != hipSuccess)
      throw qx::Error{QX_ENOMEM, "info allocation failed"};
    hipLaunchKernelGGL(qx::fill_int_kernel, dim3(64), dim3(256), 0, nullptr, p, (int64_t)len, 0);
    QX_HIP_CHECK(hipDeviceSynchronize());
    if (info) {
      // carry over flags raised before the growth
      std::vector<int> old(info_len), cur(len, 0);
      QX_HIP_CHECK(hipMemcpy(old.data(), info, info_len * sizeof(int), hipMemcpyDeviceToHost));
      int any = 0;
      for (int v : old) any |= v;
      if (any) {
        cur[0] = any;
        QX_HIP_CHECK(hipMemcpy(p, cur.data(), len * sizeof(int), hipMemcpyHostToDevice));
      }
      QX_HIP_CHECK(hipFree(info));
    }
    info = p;
    info_len = len;
  }
  return info;
}

int* qx_handle::ensure_mask(size_t n) {
  if (n > mask_len) {
    const size_t len = std::max<size_t>(n, 1024);
    int* p = nullptr;
    if (hipMalloc(&p, len * sizeof(int)) != hipSuccess)
      throw qx::Error{QX_ENOMEM, "mask allocation failed"};
    if (mask) {
      QX_HIP_CHECK(hipDeviceSynchronize());
      QX_HIP_CHECK(hipFree(mask));
    }
    mask = p;
    mask_len = len;
  }
  return mask;
}

namespace {

// Per-call thread-local state taken from the handle: kernel timer and GEMM
// tile policy (helper threads of run_split inherit both).
struct TimerScope {
  explicit TimerScope(qx_handle* h) {
    qx::active_timer = h ? h->timer : nullptr;
    qx::active_gemm_tile = h ? h->gemm_tile : 0;
  }
  ~TimerScope() {
    qx::active_timer = nullptr;
    qx::active_gemm_tile = 0;
  }
};

template <typename F>
int guarded(qx_handle* h, F&& f) {
  try {
    TimerScope scope(h);
    if (h) {
      int cur = -1;
      if (hipGetDevice(&cur) == hipSuccess && cur != h->device) QX_HIP_CHECK(hipSetDevice(h->device));
    }
    return f();
  } catch (const qx::Error& e) {
    if (h) h->last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (h) h->last_error = e.what();
    return QX_EINVAL;
  }
}

}  // namespace

extern "C" {

int qx_abi_version(void) { return QX_ABI_VERSION; }

int qx_create(qx_handle** handle, int device) {
  if (!handle) return QX_EINVAL;
  *handle = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return QX_EHIP;
  try {
    qx::noalloc_init(device);
  } catch (const qx::Error& e) {
    return e.code;
  }
  qx_handle* h = new qx_handle();
  h->device = device;
  if (const char* env = std::getenv("QX_RGF_STREAMS")) {
    const int p = std::atoi(env);
    if (p >= 1 && p <= 8) h->parts = p;
  }
  *handle = h;
  return QX_OK;
}

int qx_timer_start(qx_handle* h, void* stream) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    if (!h->timer) h->timer = new qx::KernelTimer();
    h->timer->recs.clear();
    h->timer->next = 0;
    h->timer->t0 = h->timer->get();
    QX_HIP_CHECK(hipEventRecord(h->timer->t0, static_cast<hipStream_t>(stream)));
    return QX_OK;
  });
}

int qx_timer_stop(qx_handle* h, void* stream, double* stats) {
  return guarded(h, [&]() {
    QX_REQUIRE(h && stats, "timer_stop: null argument");
    for (int i = 0; i < 4 * qx::KernelTimer::NKIND; ++i) stats[i] = 0.0;
    if (!h->timer) return QX_OK;
    QX_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    QX_HIP_CHECK(hipDeviceSynchronize());
    std::vector<std::pair<double, double>> iv[qx::KernelTimer::NKIND];
    for (const auto& r : h->timer->recs) {
      float ms = 0.f, s0 = 0.f, s1 = 0.f;
      QX_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      QX_HIP_CHECK(hipEventElapsedTime(&s0, h->timer->t0, r.a));
      QX_HIP_CHECK(hipEventElapsedTime(&s1, h->timer->t0, r.b));
      stats[4 * r.kind + 0] += 1.0;
      stats[4 * r.kind + 1] += ms;
      stats[4 * r.kind + 2] += r.flops;
      iv[r.kind].emplace_back(s0, s1);
    }
    // wall time covered by at least one launch of the family
    for (int k = 0; k < qx::KernelTimer::NKIND; ++k) {
      std::sort(iv[k].begin(), iv[k].end());
      double covered = 0.0, lo = -1.0, hi = -1.0;
      for (const auto& p : iv[k]) {
        if (p.first > hi) {
          if (hi > lo) covered += hi - lo;
          lo = p.first;
          hi = p.second;
        } else if (p.second > hi) {
          hi = p.second;
        }
      }
      if (hi > lo) covered += hi - lo;
      stats[4 * k + 3] = covered;
    }
    delete h->timer;
    h->timer = nullptr;
    return QX_OK;
  });
}

int qx_destroy(qx_handle* h) {
  if (!h) return QX_OK;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  delete h->timer;
  for (hipStream_t s2 : h->streams) (void)hipStreamDestroy(s2);
  for (hipEvent_t ev : h->join) (void)hipEventDestroy(ev);
  if (h->fork) (void)hipEventDestroy(h->fork);
  if (h->ws) (void)hipFree(h->ws);
  if (h->info) (void)hipFree(h->info);
  if (h->mask) (void)hipFree(h->mask);
  delete h;
  return QX_OK;
}

const char* qx_last_error(const qx_handle* h) { return h ? h->last_error.c_str() : "null handle"; }

int qx_workspace_query(qx_handle* h, const qx_btd_layout* la, int precision, size_t* bytes) {
  (void)precision;
  return guarded(h, [&]() {
    QX_REQUIRE(la && bytes, "workspace_query: null argument");
    *bytes = qx::rgf_workspace_bytes(la, true, precision);
    return QX_OK;
  });
}

int qx_sync_check(qx_handle* h, void* stream) {
  return guarded(h, [&]() {
    hipStream_t st = static_cast<hipStream_t>(stream);
    QX_HIP_CHECK(hipStreamSynchronize(st));
    if (!h->info) return QX_OK;
    std::vector<int> v(h->info_len);
    QX_HIP_CHECK(hipMemcpy(v.data(), h->info, h->info_len * sizeof(int), hipMemcpyDeviceToHost));
    int any = 0, fallbacks = 0;
    for (int x : v) {
      any |= x;
      fallbacks += (x & 2) ? 1 : 0;
    }
    if (any) {
      hipLaunchKernelGGL(qx::fill_int_kernel, dim3(64), dim3(256), 0, st, h->info,
                         (int64_t)h->info_len, 0);
      QX_HIP_CHECK(hipStreamSynchronize(st));
      if (any & 1) {
        h->last_error = "Singular matrix";
        return QX_ESINGULAR;
      }
      h->last_error = std::to_string(fallbacks) +
                      " matrices had a complex64 block inverse redone in FP64 (residual of the "
                      "Newton-Schulz check above tolerance or not finite)";
      return QX_FP64_FALLBACK;
    }
    return QX_OK;
  });
}

int qx_rgf_selected_inv(qx_handle* h, void* stream, const qx_btd_layout* la, const void* a,
                        const void* const* obc_r, const qx_btd_layout* lx, void* x,
                        int precision) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    qx::rgf_selected_inv(h, static_cast<hipStream_t>(stream), la, a, obc_r, lx, x, precision);
    return QX_OK;
  });
}

int qx_rgf_selected_solve(qx_handle* h, void* stream, const qx_btd_layout* la, const void* a,
                          const qx_btd_layout* lsl, const void* sl, const qx_btd_layout* lsg,
                          const void* sg, const void* const* obc_r, const void* const* obc_l,
                          const void* const* obc_g, const qx_btd_layout* lxl, void* xl,
                          const qx_btd_layout* lxg, void* xg, const qx_btd_layout* lxr, void* xr,
                          void* current, void* corner, double* density, double* ldos,
                          uint32_t flags, int precision) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    qx::rgf_selected_solve(h, static_cast<hipStream_t>(stream), la, a, lsl, sl, lsg, sg, obc_r,
                           obc_l, obc_g, lxl, xl, lxg, xg, lxr, xr, current, corner, density, ldos,
                           flags, precision);
    return QX_OK;
  });
}

int qx_zgemm_batched(qx_handle* h, void* stream, int64_t batch, int m, int n, int k, int op_a,
                     int op_b, const double* alpha, const void* a, int lda, int64_t stride_a,
                     const void* b, int ldb, int64_t stride_b, const double* beta, void* c,
                     int ldc, int64_t stride_c, int precision) {
  return guarded(h, [&]() {
    QX_REQUIRE(alpha && beta, "zgemm: null alpha/beta");
    struct OzScope {
      ~OzScope() { qx::active_oz_ws = qx::OzWorkspace{}; }
    } oz_scope;
    if (const size_t bytes = qx::split_workspace_bytes(precision, batch, m, n, k)) {
      QX_REQUIRE(h, "null handle");
      qx::active_oz_ws = qx::OzWorkspace{h->ensure_ws(bytes), h->ws_bytes};
    }
    qx::zgemm(static_cast<hipStream_t>(stream), batch, m, n, k, op_a, op_b,
              qx::Cplx{alpha[0], alpha[1]}, qx::Mat(a, stride_a, lda), qx::Mat(b, stride_b, ldb),
              qx::Cplx{beta[0], beta[1]}, qx::Mat(c, stride_c, ldc), precision);
    return QX_OK;
  });
}

int qx_inv_batched(qx_handle* h, void* stream, int64_t batch, int n, const void* a, int lda,
                   int64_t stride_a, const void* s, int lds, int64_t stride_s, void* x, int ldx,
                   int64_t stride_x, int precision) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(batch >= 0 && n >= 1, "inv: bad sizes");
    int* info = h->ensure_info(batch);
    const int64_t nn = (int64_t)n * n;
    qx::z2* ws = static_cast<qx::z2*>(h->ensure_ws((size_t)(batch * nn) * sizeof(qx::z2) +
                                                   (size_t)(batch * 3 * n) * sizeof(int)));
    const qx::InvScratch scr{ws, nn, reinterpret_cast<int*>(ws + batch * nn), 3 * (int64_t)n};
    qx::zinv(static_cast<hipStream_t>(stream), batch, n, qx::Mat(a, stride_a, lda),
             qx::Mat(s, stride_s, lds), qx::Mat(x, stride_x, ldx), info, precision, scr);
    return QX_OK;
  });
}

int qx_sancho_rubio(qx_handle* h, void* stream, int64_t n, int b, const void* a_ii,
                    const void* a_ij, const void* a_ji, void* x_ii, int max_iterations, double tol,
                    int* iterations) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    return qx::sancho_rubio(h, static_cast<hipStream_t>(stream), n, b, a_ii, a_ij, a_ji, x_ii,
                            max_iterations, tol, iterations);
  });
}

int qx_svd_batched(qx_handle* h, void* stream, int64_t n, int rows, int cols, const void* a,
                   int lda, int64_t stride_a, void* u, double* s, void* v, int max_sweeps,
                   double tol, int* sweeps) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(n >= 0 && rows >= 1 && cols >= 1, "svd: bad sizes");
    if (n == 0) return QX_OK;
    const size_t bytes = qx::svd_workspace_bytes(n, rows, cols);
    void* ws = h->ensure_ws(bytes);
    qx::svd_tall(static_cast<hipStream_t>(stream), n, rows, cols, qx::Mat(a, stride_a, lda),
                 qx::Mat(u, (int64_t)rows * cols, cols), s, cols,
                 qx::Mat(v, (int64_t)cols * cols, cols), max_sweeps, tol, ws, h->ws_bytes,
                 sweeps);
    return QX_OK;
  });
}

int qx_spectral_modes(qx_handle* h, void* stream, int64_t n, int b, int m, const void* ws,
                      const void* vs, const void* av_ji, const void* av_ii, const void* av_ij,
                      double min_decay, double max_decay, double min_propagation,
                      double residual_tolerance, double eta_decay, int residual_normalization,
                      unsigned char* mask, double* residual, void* de_dk) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    qx::spectral_modes(static_cast<hipStream_t>(stream), n, b, m,
                       static_cast<const qx::z2*>(ws), static_cast<const qx::z2*>(vs),
                       static_cast<const qx::z2*>(av_ji), static_cast<const qx::z2*>(av_ii),
                       static_cast<const qx::z2*>(av_ij), min_decay, max_decay, min_propagation,
                       residual_tolerance, eta_decay, residual_normalization != 0, mask, residual,
                       static_cast<qx::z2*>(de_dk));
    return QX_OK;
  });
}

int qx_beyn_moments(qx_handle* h, void* stream, int64_t n, int b, int n_quad, double r_o,
                    double r_i, const void* a_ji, const void* a_ii, const void* a_ij, void* q0,
                    void* q1, int precision) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    return qx::beyn_moments(h, static_cast<hipStream_t>(stream), n, b, n_quad, r_o, r_i, a_ji,
                            a_ii, a_ij, q0, q1, precision);
  });
}

int qx_contact_self_energy(qx_handle* h, void* stream, int64_t n, int b, const void* m10,
                           const void* g, const void* m01, const double* occupancies,
                           void* sigma_r, void* sigma_l, void* sigma_g) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(n >= 0 && b >= 1, "contact_self_energy: bad sizes");
    QX_REQUIRE(m10 && g && m01 && occupancies && sigma_r && sigma_l && sigma_g,
               "contact_self_energy: null argument");
    if (n == 0) return QX_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t bb = (int64_t)b * b;
    qx::z2* t = static_cast<qx::z2*>(h->ensure_ws((size_t)(n * bb) * sizeof(qx::z2)));
    const qx::Cplx one{1.0, 0.0}, zero{0.0, 0.0};
    qx::zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, one, qx::Mat(g, bb, b), qx::Mat(m01, bb, b), zero,
              qx::Mat(t, bb, b), QX_PREC_C128);
    qx::zgemm(st, n, b, b, b, QX_OP_N, QX_OP_N, one, qx::Mat(m10, bb, b), qx::Mat(t, bb, b),
              zero, qx::Mat(sigma_r, bb, b), QX_PREC_C128);
    qx::ew_contact_sigma(st, n, b, qx::Mat(sigma_r, bb, b), occupancies,
                         qx::Mat(sigma_l, bb, b), qx::Mat(sigma_g, bb, b));
    return QX_OK;
  });
}

int qx_assemble_system_matrix(qx_handle* h, void* stream, int64_t n_stack, int64_t nnz,
                              const void* z, const void* s, const void* t, const void* sigma,
                              int64_t sigma_stride, void* a, int64_t a_stride) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(n_stack >= 0 && nnz >= 0, "assemble_system_matrix: bad sizes");
    QX_REQUIRE(a && (z || !s), "assemble_system_matrix: null argument");
    qx::ew_assemble(static_cast<hipStream_t>(stream), n_stack, nnz,
                    static_cast<const qx::z2*>(z), static_cast<const qx::z2*>(s),
                    static_cast<const qx::z2*>(t), static_cast<const qx::z2*>(sigma),
                    sigma_stride, static_cast<qx::z2*>(a), a_stride);
    return QX_OK;
  });
}

int qx_set_tuning(qx_handle* h, int streams, int gemm_tile) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(streams >= 1 && streams <= 8, "set_tuning: streams must be in [1, 8]");
    QX_REQUIRE(gemm_tile == QX_GEMM_AUTO || gemm_tile == QX_GEMM_REG_64x64 ||
                   gemm_tile == QX_GEMM_REG_64x128 || gemm_tile == QX_GEMM_REG_128x64 ||
                   gemm_tile == QX_GEMM_LDS_64 || gemm_tile == QX_GEMM_LDS_128,
               "set_tuning: gemm_tile must be one of QX_GEMM_* (0, 1, 2, 3, 64, 128)");
    h->parts = streams;
    h->gemm_tile = gemm_tile;
    return QX_OK;
  });
}

int qx_zgemm_ozaki(qx_handle* h, void* stream, int64_t batch, int m, int n, int k, int op_a,
                   int op_b, const double* alpha, const void* a, int lda, int64_t stride_a,
                   const void* b, int ldb, int64_t stride_b, const double* beta, void* c,
                   int ldc, int64_t stride_c, int slices) {
  return guarded(h, [&]() {
    QX_REQUIRE(h, "null handle");
    QX_REQUIRE(alpha && beta, "zgemm_ozaki: null alpha/beta");
    QX_REQUIRE(batch >= 0 && m >= 0 && n >= 0 && k >= 1, "zgemm_ozaki: bad sizes");
    const size_t bytes = qx::ozaki_workspace_bytes(batch, m, n, k, slices < 2 ? 2 : slices);
    void* ws = h->ensure_ws(bytes);
    qx::zgemm_ozaki(static_cast<hipStream_t>(stream), batch, m, n, k, op_a, op_b,
                    qx::Cplx{alpha[0], alpha[1]}, qx::Mat(a, stride_a, lda),
                    qx::Mat(b, stride_b, ldb), qx::Cplx{beta[0], beta[1]},
                    qx::Mat(c, stride_c, ldc), qx::Mat(c, stride_c, ldc), slices, ws, h->ws_bytes);
    return QX_OK;
  });
}

}  // extern "C"
