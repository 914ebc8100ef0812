"""ctypes binding of libquatrex_amd.so (the C ABI in include/quatrex_amd.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already mapped (both use the soname ``libamdhip64.so.7``).
There is no fallback: if the library is missing or no HIP device is visible,
every solver call raises.
"""

import ctypes
import os
import threading
import warnings
from pathlib import Path

import numpy as np
import torch

# QX_LIBRARY selects another build of the same library (e.g. a tuning variant).
LIB_PATH = Path(os.environ.get("QX_LIBRARY") or
                Path(__file__).resolve().parent / "libquatrex_amd.so")

QX_OK = 0
QX_NOT_CONVERGED = 1
QX_FP64_FALLBACK = 2
QX_EINVAL = -1
QX_ENOMEM = -2
QX_EHIP = -3
QX_ESINGULAR = -4

ABI_VERSION = 3  # QX_ABI_VERSION of include/quatrex_amd.h
PREC_C128 = 0
PREC_C64 = 1
PREC_MIXED = 2
PREC_C128_3M = 3

def PREC_OZAKI2(n: int) -> int:
    """QX_PREC_OZAKI2(n): complex128 data, block products on int8 matrix cores
    by modular arithmetic with n moduli (8..20)."""
    return 16 + n


PRECISIONS = {"c128": PREC_C128, "complex128": PREC_C128, "c64": PREC_C64, "complex64": PREC_C64,
              "c128-3m": PREC_C128_3M}
PRECISIONS.update({f"oz2-{n}": PREC_OZAKI2(n) for n in range(8, 21)})


def PREC_OZAKI_BF16(s: int) -> int:
    """QX_PREC_OZAKI_BF16(s): complex128 data, block products on bf16 matrix
    cores from s 7-bit digit slices per operand line (2..7)."""
    return 40 + s


PRECISIONS.update({f"bf16x{s}": PREC_OZAKI_BF16(s) for s in range(2, 8)})
# RGF also accepts the mixed mode (complex64 products, FP64 Newton-Schulz
# step on every block inverse).
RGF_PRECISIONS = dict(PRECISIONS, mixed=PREC_MIXED)

OP_N = 0
# GEMM kernel choices for set_tuning (QX_GEMM_* of the C header)
GEMM_AUTO, GEMM_REG_64x64, GEMM_REG_64x128, GEMM_REG_128x64 = 0, 1, 2, 3
GEMM_LDS_64, GEMM_LDS_128 = 64, 128
OP_C = 2

SOLVE_RETARDED = 1
SOLVE_CURRENT = 2
SOLVE_RETARDED_OFFDIAG = 4
SOLVE_SKIP_LOWER_L = 8
SOLVE_SKIP_LOWER_G = 16

EXPORTED_SYMBOLS = (
    "qx_abi_version", "qx_create", "qx_destroy", "qx_last_error", "qx_workspace_query",
    "qx_sync_check", "qx_rgf_selected_inv", "qx_rgf_selected_solve", "qx_zgemm_batched",
    "qx_inv_batched", "qx_sancho_rubio", "qx_beyn_moments", "qx_timer_start", "qx_timer_stop",
    "qx_contact_self_energy", "qx_assemble_system_matrix", "qx_set_tuning", "qx_zgemm_ozaki",
    "qx_inv_batched_noalloc", "qx_svd_batched", "qx_spectral_modes",
)


class BtdLayout(ctypes.Structure):
    """Mirror of ``qx_btd_layout``."""

    _fields_ = [
        ("n_blocks", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n_stack", ctypes.c_int64),
        ("stack_stride", ctypes.c_int64),
        ("block_sizes", ctypes.c_void_p),
        ("diag", ctypes.c_void_p),
        ("upper", ctypes.c_void_p),
        ("lower", ctypes.c_void_p),
    ]


def make_layout(block_sizes, n_stack, stack_stride, diag, upper, lower):
    """Build a BtdLayout; returns (struct, keepalive) -- keep both alive."""
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    d = np.ascontiguousarray(diag, dtype=np.int64)
    u = np.ascontiguousarray(upper, dtype=np.int64)
    lo = np.ascontiguousarray(lower, dtype=np.int64)
    if len(u) == 0:
        u = np.full(1, -1, dtype=np.int64)
        lo = np.full(1, -1, dtype=np.int64)
    lay = BtdLayout(len(bs), 0, int(n_stack), int(stack_stride), bs.ctypes.data,
                    d.ctypes.data, u.ctypes.data, lo.ctypes.data)
    return lay, (bs, d, u, lo)


_lib = None
_lib_lock = threading.Lock()
_handles = {}


def load() -> ctypes.CDLL:
    """Load the library (once).  Raises ImportError when it is not built."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m quatrex_amd.build` "
                "(hipcc, gfx950). quatrex_amd has no CPU fallback.")
        lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        lay_p = ctypes.POINTER(BtdLayout)
        lib.qx_abi_version.restype = i32
        lib.qx_create.argtypes = [ctypes.POINTER(vp), i32]
        lib.qx_destroy.argtypes = [vp]
        lib.qx_last_error.argtypes = [vp]
        lib.qx_last_error.restype = ctypes.c_char_p
        lib.qx_workspace_query.argtypes = [vp, lay_p, i32, ctypes.POINTER(ctypes.c_size_t)]
        lib.qx_sync_check.argtypes = [vp, vp]
        lib.qx_rgf_selected_inv.argtypes = [vp, vp, lay_p, vp, vp, lay_p, vp, i32]
        lib.qx_rgf_selected_solve.argtypes = [
            vp, vp, lay_p, vp, lay_p, vp, lay_p, vp, vp, vp, vp, lay_p, vp, lay_p, vp,
            lay_p, vp, vp, vp, vp, vp, ctypes.c_uint32, i32]
        lib.qx_zgemm_batched.argtypes = [
            vp, vp, i64, i32, i32, i32, i32, i32, ctypes.POINTER(f64), vp, i32, i64, vp, i32,
            i64, ctypes.POINTER(f64), vp, i32, i64, i32]
        lib.qx_inv_batched.argtypes = [vp, vp, i64, i32, vp, i32, i64, vp, i32, i64, vp, i32,
                                       i64, i32]
        lib.qx_inv_batched_noalloc.argtypes = [vp, vp, i64, i32, vp, i32, i64, vp, i32, i64,
                                               vp, i32, i64, vp, i32]
        lib.qx_svd_batched.argtypes = [vp, vp, i64, i32, i32, vp, i32, i64, vp, vp, vp, i32, f64,
                                       vp]
        lib.qx_spectral_modes.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp, vp, vp, f64, f64,
                                          f64, f64, f64, i32, vp, vp, vp]
        lib.qx_sancho_rubio.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, i32, f64,
                                        ctypes.POINTER(ctypes.c_int)]
        lib.qx_beyn_moments.argtypes = [vp, vp, i64, i32, i32, f64, f64, vp, vp, vp, vp, vp, i32]
        lib.qx_contact_self_energy.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp, vp, vp, vp]
        lib.qx_assemble_system_matrix.argtypes = [vp, vp, i64, i64, vp, vp, vp, vp, i64, vp,
                                                   i64]
        lib.qx_set_tuning.argtypes = [vp, i32, i32]
        lib.qx_zgemm_ozaki.argtypes = [
            vp, vp, i64, i32, i32, i32, i32, i32, ctypes.POINTER(f64), vp, i32, i64, vp, i32,
            i64, ctypes.POINTER(f64), vp, i32, i64, i32]
        lib.qx_timer_start.argtypes = [vp, vp]
        lib.qx_timer_stop.argtypes = [vp, vp, ctypes.POINTER(f64)]
        for name in EXPORTED_SYMBOLS:
            getattr(lib, name).restype = getattr(lib, name).restype or i32
        if lib.qx_abi_version() != ABI_VERSION:
            raise ImportError("libquatrex_amd ABI version mismatch; rebuild the library")
        _lib = lib
        return lib


def require_device_tensor(t: torch.Tensor, what: str) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{what} must be a HIP device tensor (torch 'cuda' device); "
                         "quatrex_amd has no CPU fallback")


_override = threading.local()


def new_handle(device: torch.device) -> ctypes.c_void_p:
    """A separate library handle (own workspace and helper streams)."""
    lib = load()
    device = torch.device(device)
    idx = device.index if device.index is not None else torch.cuda.current_device()
    h = ctypes.c_void_p()
    rc = lib.qx_create(ctypes.byref(h), idx)
    if rc != QX_OK:
        raise RuntimeError(f"qx_create failed on device {idx} (code {rc})")
    return h


def workspace_bytes(h, block_sizes, n_stack: int, precision: str = "c128") -> int:
    """Bytes of handle workspace an RGF selected_solve of ``n_stack``
    matrices with these block sizes grows the handle to (qx_workspace_query)."""
    n = len(block_sizes)
    lay, keep = make_layout(block_sizes, n_stack, 0, np.zeros(n), np.zeros(max(n - 1, 0)),
                            np.zeros(max(n - 1, 0)))
    out = ctypes.c_size_t(0)
    check(load().qx_workspace_query(h, ctypes.byref(lay), int(RGF_PRECISIONS[precision]),
                                    ctypes.byref(out)), h, "workspace_query")
    return int(out.value)


def set_tuning(h, streams: int, gemm_tile: int = 0) -> None:
    """Sub-batch streams of one RGF call and GEMM kernel (GEMM_*; 0 automatic)."""
    check(load().qx_set_tuning(h, int(streams), int(gemm_tile)), h, "set_tuning")


class use_handle:
    """Route this thread's library calls through handle ``h`` (context manager)."""

    def __init__(self, h):
        self.h = h

    def __enter__(self):
        self.prev = getattr(_override, "h", None)
        _override.h = self.h
        return self.h

    def __exit__(self, *exc):
        _override.h = self.prev


def capturing(device=None) -> bool:
    """True while the current stream records a graph (no host syncs then)."""
    return torch.cuda.is_current_stream_capturing()


def handle(device: torch.device) -> ctypes.c_void_p:
    """This thread's library handle: the ``use_handle`` override if any, else
    the per-device default handle (created on first use)."""
    h = getattr(_override, "h", None)
    if h is not None:
        return h
    lib = load()
    idx = device.index if device.index is not None else torch.cuda.current_device()
    h = _handles.get(idx)
    if h is None:
        h = ctypes.c_void_p()
        rc = lib.qx_create(ctypes.byref(h), idx)
        if rc != QX_OK:
            raise RuntimeError(f"qx_create failed on device {idx} (code {rc})")
        _handles[idx] = h
    return h


def stream_ptr(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def check(rc: int, h, what: str) -> int:
    """Map a C-ABI status to the reference's Python error behaviour."""
    if rc == QX_OK:
        return rc
    msg = load().qx_last_error(h).decode(errors="replace")
    if rc == QX_NOT_CONVERGED:
        return rc
    if rc == QX_FP64_FALLBACK:
        warnings.warn(f"{what}: {msg}", RuntimeWarning, stacklevel=3)
        return rc
    if rc == QX_EINVAL:
        raise ValueError(f"{what}: {msg}")
    if rc == QX_ENOMEM:
        raise MemoryError(f"{what}: {msg}")
    if rc == QX_ESINGULAR:
        raise np.linalg.LinAlgError("Singular matrix")
    raise RuntimeError(f"{what}: {msg} (code {rc})")


def sync_check(device: torch.device, what: str) -> None:
    """Wait for the device stream and raise on recorded device errors.

    Skipped while a graph is being captured; the caller checks after replay.
    """
    if capturing():
        return
    h = handle(device)
    check(load().qx_sync_check(h, stream_ptr(device)), h, what)


class KernelTimer:
    """Context manager timing the library's GEMM / inverse launches with HIP events.

    ``stats`` after exit: dict with launches, milliseconds and algorithmic
    flops per kernel family ("gemm", "inv").
    """

    def __init__(self, device):
        self.device = torch.device(device)
        self.stats = None

    def __enter__(self):
        h = handle(self.device)
        check(load().qx_timer_start(h, stream_ptr(self.device)), h, "timer_start")
        return self

    def __exit__(self, *exc):
        h = handle(self.device)
        out = (ctypes.c_double * 8)()
        check(load().qx_timer_stop(h, stream_ptr(self.device), out), h, "timer_stop")
        self.stats = {
            name: {"launches": int(out[4 * k]), "ms": out[4 * k + 1], "flops": out[4 * k + 2],
                   "busy_ms": out[4 * k + 3]}
            for k, name in enumerate(("gemm", "inv"))
        }
        return False


def warn_not_converged(message: str) -> None:
    warnings.warn(message, RuntimeWarning, stacklevel=3)
