// qx_inv_batched_noalloc: the batched inverse with every scratch buffer in
// a fixed device array of this file, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The array `scratch` holds, per call, the inverse's n x n ping-pong buffer
// and 3 n pivot ints per matrix plus one status flag per matrix.  Each GPU
// holds its own copy, at its own address, from the first handle created on
// it until the process ends.  Host code never passes the array's name to a
// kernel (on the host that name is only a stand-in for the symbol): the
// device address of each GPU's copy is looked up once, when the first handle
// on that GPU is created, and kept in `scratch_ptr`.
//
// Every call on one GPU shares that GPU's copy and leaves the last call's
// contents in it; the flags are cleared by a fill kernel at the start of each
// call.  Calls that overlap in time on one GPU (different streams, or host
// threads without ordering) race on the array.
#include <mutex>

#include "qx_internal.h"

namespace qx {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // QX_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == QX_NOALLOC_SCRATCH_BYTES, "header and library disagree");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxDevices] = {};

__global__ void noalloc_fill_int(int* p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (called from qx_create).
void noalloc_init(int device) {
  QX_REQUIRE(device >= 0 && device < kMaxDevices, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return;
  int prev = -1;
  QX_HIP_CHECK(hipGetDevice(&prev));
  if (prev != device) QX_HIP_CHECK(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) QX_HIP_CHECK(hipSetDevice(prev));
  if (e != hipSuccess || !p)
    throw Error{QX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  scratch_ptr[device] = p;
}

}  // namespace qx

extern "C" int qx_inv_batched_noalloc(qx_handle* h, void* stream, int64_t batch, int n,
                                      const void* a, int lda, int64_t stride_a, const void* s,
                                      int lds, int64_t stride_s, void* x, int ldx,
                                      int64_t stride_x, int* info, int precision) {
  try {
    if (!h) return QX_EINVAL;
    // no kernel timer here: its events are created on demand
    qx::active_timer = nullptr;
    qx::active_gemm_tile = h->gemm_tile;
    QX_REQUIRE(batch >= 0 && n >= 1, "inv_noalloc: bad sizes");
    QX_REQUIRE(n <= qx::kInvMaxN, "inv_noalloc: block size must be in [1, 512]");
    const int64_t nn = (int64_t)n * n;
    // layout: [batch][n x n complex] | [batch][3 n ints] | [batch flags]
    const size_t buf_bytes = qx::align256((size_t)(batch * nn) * sizeof(qx::z2));
    const size_t int_bytes = qx::align256((size_t)(batch * 3 * n) * sizeof(int));
    const size_t flag_bytes = qx::align256((size_t)batch * sizeof(int));
    const size_t need = buf_bytes + int_bytes + flag_bytes;
    if (need > qx::kScratchBytes)
      throw qx::Error{QX_EINVAL, "inv_noalloc: " + std::to_string(batch) + " matrices of " +
                                     std::to_string(n) + " need " + std::to_string(need) +
                                     " bytes of scratch, more than the fixed " +
                                     std::to_string(qx::kScratchBytes)};
    int cur = -1;
    QX_HIP_CHECK(hipGetDevice(&cur));
    if (cur != h->device) QX_HIP_CHECK(hipSetDevice(h->device));
    void* base = nullptr;
    {
      std::lock_guard<std::mutex> lock(qx::scratch_mu);
      base = qx::scratch_ptr[h->device];
    }
    QX_REQUIRE(base, "inv_noalloc: scratch not set up on this device");
    if (batch == 0) return QX_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char* p = static_cast<unsigned char*>(base);
    const qx::InvScratch scr{reinterpret_cast<qx::z2*>(p), nn,
                             reinterpret_cast<int*>(p + buf_bytes), 3 * (int64_t)n};
    int* flags = info ? info : reinterpret_cast<int*>(p + buf_bytes + int_bytes);
    const unsigned grid = (unsigned)std::min<int64_t>((batch + 255) / 256, 1024);
    hipLaunchKernelGGL(qx::noalloc_fill_int, dim3(grid), dim3(256), 0, st, flags, batch, 0);
    QX_HIP_CHECK(hipGetLastError());
    qx::zinv(st, batch, n, qx::Mat(a, stride_a, lda), qx::Mat(s, stride_s, lds),
             qx::Mat(x, stride_x, ldx), flags, precision, scr);
    return QX_OK;
  } catch (const qx::Error& e) {
    h->last_error = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    h->last_error = e.what();
    return QX_EINVAL;
  }
}
