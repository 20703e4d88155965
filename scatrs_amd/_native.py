"""Loads the HIP library (libscatrs_hip.so, built in-tree for gfx950) and binds its C ABI.

There is no CPU fallback: if the library is missing or no GPU is visible, every entry point
raises NativeLibraryError. Device memory and streams come from PyTorch-ROCm; the library
itself takes plain device pointers.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

from .errors import (InvalidInputError, NativeLibraryError, NotConvergedError,
                     OutsideValidatedDomainError)

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libscatrs_hip.so"

OK, INVALID_INPUT, OUTSIDE_VALIDATED_DOMAIN, NOT_CONVERGED, DEVICE_ERROR = 0, 1, 2, 3, 4
ABI_VERSION = 3

_lib = None
_lock = threading.Lock()
_tls = threading.local()

# Every symbol include/scatrs_hip.h declares.
SYMBOLS = (
    "scatrs_abi_version",
    "scatrs_create",
    "scatrs_destroy",
    "scatrs_set_stream",
    "scatrs_last_error",
    "scatrs_set_profiling",
    "scatrs_kernel_times",
    "scatrs_set_angular_kernel",
    "scatrs_set_tuning",
    "scatrs_get_tuning",
    "scatrs_set_phase_cosines",
    "scatrs_default_n_max",
    "scatrs_mie_efficiencies",
    "scatrs_mie_efficiencies_noalloc",
    "scatrs_mie_efficiencies_host",
    "scatrs_mie_terms",
    "scatrs_mie_coefficients",
    "scatrs_layered_terms",
    "scatrs_layered_efficiencies",
    "scatrs_layered_coefficients",
    "scatrs_ensemble_fixed",
    "scatrs_ensemble_fixed_vjp",
    "scatrs_ensemble_fixed_graph",
    "scatrs_graph_launch",
    "scatrs_graph_status",
    "scatrs_graph_last_error",
    "scatrs_graph_destroy",
    "scatrs_ensemble_nodes",
    "scatrs_ensemble_nodes_tangent",
    "scatrs_ensemble_nodes_vjp",
    "scatrs_size_range",
    "scatrs_size_log_density",
    "scatrs_gauss_legendre",
    "scatrs_size_message",
)


def load() -> C.CDLL:
    """The library, loaded once; raises NativeLibraryError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise NativeLibraryError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`"
                " or `python -m scatrs_amd.build`")
        try:
            # torch's HIP runtime first: the library then binds to it, where loading the
            # library first would bring a second HIP and HSA runtime into the process, and the
            # one initialized second finds no device
            import torch  # noqa: F401

            L = C.CDLL(str(LIB_PATH))
        except OSError as e:
            raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
        L.scatrs_abi_version.restype = i32
        L.scatrs_create.argtypes = [i32, vp, C.POINTER(vp)]
        L.scatrs_destroy.argtypes = [vp]
        L.scatrs_set_stream.argtypes = [vp, vp]
        L.scatrs_last_error.argtypes = [vp]
        L.scatrs_last_error.restype = C.c_char_p
        L.scatrs_set_profiling.argtypes = [vp, i32]
        L.scatrs_kernel_times.argtypes = [vp, C.POINTER(dbl), i32]
        L.scatrs_set_angular_kernel.argtypes = [vp, i32]
        L.scatrs_set_tuning.argtypes = [vp, i32, i64]
        L.scatrs_get_tuning.argtypes = [vp, i32]
        L.scatrs_get_tuning.restype = i64
        L.scatrs_set_phase_cosines.argtypes = [vp, vp, i64]
        L.scatrs_default_n_max.argtypes = [dbl]
        L.scatrs_default_n_max.restype = i64
        L.scatrs_mie_efficiencies.argtypes = [vp, vp, vp, i64, i64, i32, vp, vp,
                                              C.POINTER(i64)]
        L.scatrs_mie_efficiencies_noalloc.argtypes = [vp, vp, vp, i64, i64, i32, vp, vp,
                                                      C.POINTER(i64)]
        L.scatrs_mie_efficiencies_host.argtypes = [vp, vp, vp, i64, i64, i32, vp, vp,
                                                   C.POINTER(i64)]
        L.scatrs_mie_terms.argtypes = [vp, vp, vp, i64, i64, i32, C.POINTER(i64), C.POINTER(i64)]
        L.scatrs_mie_coefficients.argtypes = [vp, vp, vp, i64, i64, i32, i64, vp, vp, vp, vp,
                                              C.POINTER(i64)]
        L.scatrs_layered_terms.argtypes = [vp, vp, vp, i64, i32, i64, i32, i32, C.POINTER(i64),
                                           C.POINTER(i64)]
        L.scatrs_layered_efficiencies.argtypes = [vp, vp, vp, i64, i32, i64, i32, vp, vp,
                                                  C.POINTER(i64)]
        L.scatrs_layered_coefficients.argtypes = [vp, vp, vp, i64, i32, i64, i32, i64, vp, vp, vp,
                                                  vp, C.POINTER(i64)]
        if hasattr(L, "scatrs_ensemble_fixed"):
            L.scatrs_ensemble_fixed.argtypes = [vp, i64, vp, vp, vp, vp, vp, i32, vp, vp, i64,
                                                i64, i64, i32, i32, vp, vp, vp,
                                                C.POINTER(i64)]
        L.scatrs_ensemble_fixed_graph.argtypes = [vp, i64, vp, vp, vp, vp, vp, i32, vp, vp, i64,
                                                  i64, i64, i32, i32, vp, vp, vp,
                                                  C.POINTER(i64), C.POINTER(vp)]
        L.scatrs_graph_launch.argtypes = [vp, vp]
        L.scatrs_graph_status.argtypes = [vp, vp]
        L.scatrs_graph_last_error.argtypes = [vp]
        L.scatrs_graph_last_error.restype = C.c_char_p
        L.scatrs_graph_destroy.argtypes = [vp]
        if hasattr(L, "scatrs_ensemble_fixed_vjp"):
            L.scatrs_ensemble_fixed_vjp.argtypes = [vp, i64, vp, vp, vp, vp, vp, i32, vp, vp,
                                                    i64, i64, i64, i32, i32, i32, vp, vp, vp, vp,
                                                    C.POINTER(i64)]
        L.scatrs_ensemble_nodes_tangent.argtypes = [vp, i64, i64, vp, vp, i32, vp, vp, i64, i64,
                                                    i32, i32, i32, vp, vp, vp, C.POINTER(i64)]
        if hasattr(L, "scatrs_ensemble_nodes"):
            L.scatrs_ensemble_nodes.argtypes = [vp, i64, i64, vp, vp, i32, vp, vp, i64, i64, i32,
                                                i32, vp, vp, vp, C.POINTER(i64)]
        if hasattr(L, "scatrs_size_range"):
            L.scatrs_size_range.argtypes = [i32, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
            L.scatrs_size_log_density.argtypes = [i64, C.POINTER(C.c_int32), C.POINTER(dbl),
                                                  C.POINTER(dbl), C.POINTER(dbl), vp]
        L.scatrs_gauss_legendre.argtypes = [i64, C.POINTER(dbl), C.POINTER(dbl)]
        if hasattr(L, "scatrs_size_message"):
            L.scatrs_size_message.argtypes = [i32, C.POINTER(dbl), C.c_char_p, i64]
        if hasattr(L, "scatrs_ensemble_nodes_vjp"):
            L.scatrs_ensemble_nodes_vjp.argtypes = [vp, i64, i64, vp, vp, vp, i32, vp, vp, i64,
                                                    i64, i32, i32, vp, vp, vp, vp,
                                                    C.POINTER(i64)]
        if L.scatrs_abi_version() != ABI_VERSION:
            raise NativeLibraryError("libscatrs_hip.so has a different ABI version")
        _lib = L
        return L


class Context:
    """A library context on one device, bound to torch's current stream at each call."""

    def __init__(self, device: int) -> None:
        self.lib = load()
        self.device = device
        h = C.c_void_p()
        rc = self.lib.scatrs_create(device, None, C.byref(h))
        if rc != OK:
            raise NativeLibraryError(f"scatrs_create failed on device {device} (code {rc})")
        self.handle = h

    def set_profiling(self, on: bool) -> None:
        self.lib.scatrs_set_profiling(self.handle, int(bool(on)))

    def kernel_times(self) -> list[float]:
        """Durations (ms) of the dominant kernel of each call since profiling was set."""
        buf = (C.c_double * 1024)()
        n = self.lib.scatrs_kernel_times(self.handle, buf, 1024)
        return [buf[i] for i in range(n)]

    def set_phase_cosines(self, mu) -> None:
        """The cosines of the ensemble phase matrix (None: the Gauss-Legendre rule's)."""
        import numpy as np

        if mu is None:
            self.check(self.lib.scatrs_set_phase_cosines(self.handle, None, 0))
            return
        a = np.ascontiguousarray(mu, np.float64).reshape(-1)
        self.check(self.lib.scatrs_set_phase_cosines(
            self.handle, a.ctypes.data_as(C.c_void_p), a.size))

    def set_tuning(self, key: int, value: int) -> None:
        """A launch parameter of the context (scatrs_set_tuning)."""
        self.check(self.lib.scatrs_set_tuning(self.handle, int(key), int(value)))

    def get_tuning(self, key: int) -> int:
        return int(self.lib.scatrs_get_tuning(self.handle, int(key)))

    def set_angular_kernel(self, kind: int) -> None:
        """0: FP64 matrix cores (default); 1: the plain FP64 vector kernel; 2: fp16 matrix
        cores with two-slice split operands (phase tables to ~1e-7 of their largest element)."""
        self.check(self.lib.scatrs_set_angular_kernel(self.handle, int(kind)))

    def bind_stream(self, stream_ptr: int) -> None:
        self.lib.scatrs_set_stream(self.handle, C.c_void_p(stream_ptr))

    def check(self, rc: int) -> None:
        if rc == OK:
            return
        msg = self.lib.scatrs_last_error(self.handle).decode()
        if rc == INVALID_INPUT:
            raise InvalidInputError(msg)
        if rc == OUTSIDE_VALIDATED_DOMAIN:
            raise OutsideValidatedDomainError(msg)
        if rc == NOT_CONVERGED:
            raise NotConvergedError(msg)
        raise NativeLibraryError(msg or f"device error (code {rc})")

    def __del__(self) -> None:
        try:
            if getattr(self, "handle", None):
                self.lib.scatrs_destroy(self.handle)
        except Exception:
            pass


def release_contexts() -> None:
    """Destroys this thread's contexts (each waits for its device work first); the next
    context() call creates a new one."""
    ctxs = getattr(_tls, "ctxs", None) or {}
    _tls.ctxs = {}
    for ctx in ctxs.values():
        h, ctx.handle = ctx.handle, None
        if h:
            ctx.lib.scatrs_destroy(h)


def context(device: int | None = None) -> Context:
    """This thread's context on `device` (default: torch's current device), bound to torch's
    current stream."""
    import torch

    if not torch.cuda.is_available():
        raise NativeLibraryError("no GPU is visible: the HIP library needs an MI355X")
    if device is None:
        device = torch.cuda.current_device()
    ctxs = getattr(_tls, "ctxs", None)
    if ctxs is None:
        ctxs = _tls.ctxs = {}
    ctx = ctxs.get(device)
    if ctx is None:
        ctx = ctxs[device] = Context(device)
    ctx.bind_stream(torch.cuda.current_stream(device).cuda_stream)
    return ctx
