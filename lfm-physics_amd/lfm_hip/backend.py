"""``HipBackend``: the MI355X backend object for the reference's backend slot.

It has the method set ``lfm.core.Evolver`` calls on a backend
(lfm/core/backends/numpy_backend.py:21-388, cupy_backend.py:48-363):
``name``, ``allocate``, ``create_boundary_mask``, ``step_real``,
``step_complex``, ``step_color``, ``to_numpy``, ``from_numpy``,
``synchronize``.  Arrays are flat float32 torch tensors on the GPU; every
step enqueues hand-written HIP kernels through liblfm_hip's C ABI on torch's
current stream and returns without synchronising, like the CuPy backend.

The step methods honour the reference's four-output contract exactly:
``psi_out = (1-m) Psi'``, ``psi_prev_out = (1-m) Psi``,
``chi_out = m chi0 + (1-m) max(chi', -chi0)``, ``chi_prev_out = chi``.

Buffers handed out (``from_numpy``, ``allocate``, ``create_boundary_mask``)
are :class:`DeviceArray` tensors, which add the two NumPy/CuPy array methods
the reference's callers use on backend arrays and torch spells differently:
``copy()`` (the freeze-chi snapshot, lfm/core/evolver.py:161-166) and
``fill(v)`` (barrier enforcement, lfm/experiment/barrier.py:321-324).
``reshape``, ``[:] =``, slice assignment and in-place arithmetic are torch's
own and already behave as the reference expects (views share storage).

There is no CPU path: without a GPU or without the built library the
constructor raises :class:`GpuUnavailableError`, an ``ImportError`` as the
reference's ``get_backend("gpu")`` raises when its GPU stack is missing
(lfm/core/backends/__init__.py:57-59).
"""
from __future__ import annotations

import contextlib
import ctypes
import gc

import numpy as np
import torch

from . import _abi
from .fields import boundary_mask_host


class GpuUnavailableError(ImportError, RuntimeError):
    """No ROCm GPU or no built liblfm_hip.so (an ImportError, as the
    reference's backend registry raises for a missing GPU stack)."""


class DeviceArray(torch.Tensor):
    """A torch tensor on the GPU with the NumPy/CuPy spellings of the two
    array methods the reference calls on backend buffers.  Views (reshape,
    slicing) stay DeviceArrays and share storage, so a write through any of
    them is seen by the next step."""

    def copy(self) -> "DeviceArray":
        """A new device array with the same values (ndarray.copy)."""
        return self.clone()

    def fill(self, value) -> None:
        """Set every element to `value` in place (ndarray.fill)."""
        self.fill_(value)

    def astype(self, dtype) -> "DeviceArray":
        """A copy cast to `dtype` (a NumPy dtype or a torch dtype)."""
        if not isinstance(dtype, torch.dtype):
            dtype = _TORCH_DTYPES[np.dtype(dtype)]
        return self.to(dtype=dtype, copy=True)


_TORCH_DTYPES = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
                 np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
                 np.dtype(np.complex64): torch.complex64, np.dtype(np.bool_): torch.bool}


def as_device_array(t: torch.Tensor) -> DeviceArray:
    """The same storage viewed as a DeviceArray (no copy)."""
    return t if isinstance(t, DeviceArray) else t.as_subclass(DeviceArray)


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def _plain_tensors(method):
    """Run a step method with torch's subclass dispatch off: the argument
    checks and pointer reads on DeviceArray buffers then cost what they cost
    on plain tensors (about 0.1 instead of 2 us each)."""
    import functools

    @functools.wraps(method)
    def wrapper(*args, **kwargs):
        with torch._C.DisableTorchFunctionSubclass():
            return method(*args, **kwargs)
    return wrapper


def _require(t, n: int, what: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: expected a torch.Tensor on the GPU, got {type(t).__name__}")
    if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError(f"{what}: need a contiguous float32 GPU tensor "
                         f"(got {t.dtype} on {t.device}, contiguous={t.is_contiguous()})")
    if t.numel() < n:
        raise ValueError(f"{what}: has {t.numel()} elements, needs {n}")


def params(N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w=0.0, kappa_c=0.0,
           epsilon_cc=0.0, kappa_string=0.0, kappa_tube=0.0, flags=0, tile=0,
           chunk=0) -> _abi.LfmParams:
    """lfm_params_t with every scalar rounded to float32, as the CuPy backend
    passes them (cupy_backend.py:160-165,213-219,298-309)."""
    return _abi.LfmParams(int(N), int(flags), float(dt2), float(kappa), float(lambda_self),
                          float(chi0), float(e0_sq), float(epsilon_w), float(kappa_c),
                          float(epsilon_cc), float(kappa_string), float(kappa_tube),
                          int(tile), int(chunk))


@contextlib.contextmanager
def graph_capture(graph, stream):
    """Capture the enclosed launches on `stream` into `graph` with Python's
    cyclic garbage collector held off.  A collection during the capture could
    free an unreachable object that owns another CUDAGraph (an Evolver in a
    reference cycle, say), and destroying a graph while a stream captures is
    not permitted: the destructor's error would abort the process.  Dead
    cycles are left to the collector's next run after the capture."""
    enabled = gc.isenabled()
    gc.disable()
    try:
        with torch.cuda.stream(stream):
            with torch.cuda.graph(graph, stream=stream):
                yield
    finally:
        if enabled:
            gc.enable()


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


class ScvPlan:
    """Owner of one lfm_scv plan (batched rocFFT R2C/C2R for ncolors fields)."""

    def __init__(self, N: int, ncolors: int = 3, method: int = 0):
        self._lib = _abi.load()
        h = ctypes.c_void_p()
        _abi.check(self._lib.lfm_scv_plan_create_ex(N, ncolors, method, ctypes.byref(h)),
                   "lfm_scv_plan_create_ex")
        self.handle = h
        self.N, self.nc = N, ncolors
        self.method = int(self._lib.lfm_scv_plan_method(h))

    def apply(self, pr, pi, gamma, D, sa_out):
        _abi.check(self._lib.lfm_scv_apply(self.handle, pr.data_ptr(), pi.data_ptr(),
                                           float(gamma), float(D), sa_out.data_ptr(),
                                           stream_handle()), "lfm_scv_apply")

    def apply_psisq(self, e, gamma, D, sa_out):
        _abi.check(self._lib.lfm_scv_apply_psisq(self.handle, e.data_ptr(), float(gamma),
                                                 float(D), sa_out.data_ptr(), stream_handle()),
                   "lfm_scv_apply_psisq")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.lfm_scv_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


# device masks by (device, N, boundary_fraction), shared by every backend
_MASKS: dict[tuple, torch.Tensor] = {}


class HipBackend:
    """GPU compute backend on hand-written gfx950 kernels (see module doc)."""

    def __init__(self, device: int | str | None = None) -> None:
        if not torch.cuda.is_available():
            raise GpuUnavailableError(
                "HipBackend needs a ROCm GPU; torch.cuda.is_available() is False")
        try:
            self._lib = _abi.load()
        except ImportError as e:
            raise GpuUnavailableError(str(e)) from e
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        self._scv: dict[int, ScvPlan] = {}
        self._sa_raw: dict[int, torch.Tensor] = {}

    @property
    def name(self) -> str:
        return "hip"

    # -- memory --------------------------------------------------------------
    def allocate(self, N: int, n_psi_arrays: int, chi0: float) -> dict[str, torch.Tensor]:
        n = N ** 3
        z = lambda: as_device_array(torch.zeros(n_psi_arrays * n, dtype=torch.float32,
                                                device=self.device))
        c = lambda: as_device_array(torch.full((n,), float(chi0), dtype=torch.float32,
                                               device=self.device))
        return {"psi_A": z(), "psi_prev_A": z(), "chi_A": c(), "chi_prev_A": c(),
                "psi_B": z(), "psi_prev_B": z(), "chi_B": c(), "chi_prev_B": c()}

    def create_boundary_mask(self, N: int, boundary_fraction: float) -> DeviceArray:
        """The reference's cos^2 mask on the device.  Built on the host once
        per (N, boundary_fraction) and kept; each call returns a fresh copy,
        so callers may modify theirs."""
        key = (str(self.device), int(N), float(boundary_fraction))
        m = _MASKS.get(key)
        if m is None:
            m = _MASKS[key] = self.from_numpy(boundary_mask_host(N, boundary_fraction))
        return as_device_array(m.clone())

    def to_numpy(self, arr) -> np.ndarray:
        return arr.detach().to("cpu").numpy()

    def from_numpy(self, arr) -> DeviceArray:
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32))
        return as_device_array(torch.from_numpy(a).to(self.device))

    def synchronize(self) -> None:
        torch.cuda.synchronize(self.device)

    # -- steps -----------------------------------------------------------------
    @_plain_tensors
    def step_real(self, psi_in, psi_prev_in, chi_in, chi_prev_in, boundary_mask, psi_out,
                  psi_prev_out, chi_out, chi_prev_out, N, dt2, kappa, lambda_self, chi0,
                  e0_sq) -> None:
        n = N ** 3
        for t, w in ((psi_in, "psi_in"), (psi_prev_in, "psi_prev_in"), (chi_in, "chi_in"),
                     (chi_prev_in, "chi_prev_in"), (psi_out, "psi_out"),
                     (psi_prev_out, "psi_prev_out"), (chi_out, "chi_out"),
                     (chi_prev_out, "chi_prev_out")):
            _require(t, n, w)
        flags = 0
        if boundary_mask is not None:
            _require(boundary_mask, n, "boundary_mask")
            flags |= _abi.LFM_F_MASK
        p = params(N, dt2, kappa, lambda_self, chi0, e0_sq, flags=flags)
        rc = self._lib.lfm_step_real_f32(
            _ptr(psi_in), _ptr(psi_prev_in), _ptr(chi_in), _ptr(chi_prev_in), _ptr(boundary_mask),
            _ptr(psi_out), _ptr(psi_prev_out), _ptr(chi_out), _ptr(chi_prev_out),
            ctypes.byref(p), stream_handle())
        _abi.check(rc, "lfm_step_real_f32")

    @_plain_tensors
    def step_complex(self, psi_r_in, psi_r_prev_in, psi_i_in, psi_i_prev_in, chi_in,
                     chi_prev_in, boundary_mask, psi_r_out, psi_r_prev_out, psi_i_out,
                     psi_i_prev_out, chi_out, chi_prev_out, N, dt2, kappa, lambda_self, chi0,
                     e0_sq, epsilon_w) -> None:
        n = N ** 3
        for t, w in ((psi_r_in, "psi_r_in"), (psi_r_prev_in, "psi_r_prev_in"),
                     (psi_i_in, "psi_i_in"), (psi_i_prev_in, "psi_i_prev_in"),
                     (chi_in, "chi_in"), (chi_prev_in, "chi_prev_in"),
                     (psi_r_out, "psi_r_out"), (psi_r_prev_out, "psi_r_prev_out"),
                     (psi_i_out, "psi_i_out"), (psi_i_prev_out, "psi_i_prev_out"),
                     (chi_out, "chi_out"), (chi_prev_out, "chi_prev_out")):
            _require(t, n, w)
        flags = 0
        if boundary_mask is not None:
            _require(boundary_mask, n, "boundary_mask")
            flags |= _abi.LFM_F_MASK
        p = params(N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w, flags=flags)
        rc = self._lib.lfm_step_complex_f32(
            _ptr(psi_r_in), _ptr(psi_r_prev_in), _ptr(psi_i_in), _ptr(psi_i_prev_in),
            _ptr(chi_in), _ptr(chi_prev_in), _ptr(boundary_mask), _ptr(psi_r_out),
            _ptr(psi_r_prev_out), _ptr(psi_i_out), _ptr(psi_i_prev_out), _ptr(chi_out),
            _ptr(chi_prev_out), ctypes.byref(p), stream_handle())
        _abi.check(rc, "lfm_step_complex_f32")

    def scv_plan(self, N: int) -> ScvPlan:
        if N not in self._scv:
            self._scv[N] = ScvPlan(N, 3)
        return self._scv[N]

    @_plain_tensors
    def step_color(self, psi_r_in, psi_r_prev_in, psi_i_in, psi_i_prev_in, chi_in, chi_prev_in,
                   boundary_mask, psi_r_out, psi_r_prev_out, psi_i_out, psi_i_prev_out,
                   chi_out, chi_prev_out, N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w,
                   kappa_c=0.0, epsilon_cc=0.0, kappa_string=0.0, kappa_tube=0.0,
                   sa_fields_in=None, sa_fields_out=None, sa_gamma=0.1, sa_d=4.9,
                   dt=0.02) -> None:
        """3-colour step.  When kappa_tube > 0 and ``sa_fields_out`` is given,
        S_a is recomputed from the input Psi (numpy_backend.py:341,
        cupy_backend.py:261) and its clipped value lands in ``sa_fields_out``;
        ``sa_fields_in`` is never read, as in the reference."""
        n = N ** 3
        for t, w in ((psi_r_in, "psi_r_in"), (psi_r_prev_in, "psi_r_prev_in"),
                     (psi_i_in, "psi_i_in"), (psi_i_prev_in, "psi_i_prev_in"),
                     (psi_r_out, "psi_r_out"), (psi_r_prev_out, "psi_r_prev_out"),
                     (psi_i_out, "psi_i_out"), (psi_i_prev_out, "psi_i_prev_out")):
            _require(t, 3 * n, w)
        for t, w in ((chi_in, "chi_in"), (chi_prev_in, "chi_prev_in"), (chi_out, "chi_out"),
                     (chi_prev_out, "chi_prev_out")):
            _require(t, n, w)
        flags = 0
        if boundary_mask is not None:
            _require(boundary_mask, n, "boundary_mask")
            flags |= _abi.LFM_F_MASK
        sa = None
        if kappa_tube > 0 and sa_fields_out is not None:
            _require(sa_fields_out, 3 * n, "sa_fields_out")
            sa = self._sa_raw.get(N)
            if sa is None:
                sa = self._sa_raw[N] = torch.empty(3 * n, dtype=torch.float32, device=self.device)
            self.scv_plan(N).apply(psi_r_in, psi_i_in, sa_gamma, sa_d, sa)
        else:
            kappa_tube = 0.0   # no S_a fields: the SCV term is zero, as in the reference
        p = params(N, dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w, kappa_c, epsilon_cc,
                   kappa_string, kappa_tube, flags=flags)
        rc = self._lib.lfm_step_color_f32(
            _ptr(psi_r_in), _ptr(psi_r_prev_in), _ptr(psi_i_in), _ptr(psi_i_prev_in),
            _ptr(chi_in), _ptr(chi_prev_in), _ptr(boundary_mask), _ptr(psi_r_out),
            _ptr(psi_r_prev_out), _ptr(psi_i_out), _ptr(psi_i_prev_out), _ptr(chi_out),
            _ptr(chi_prev_out), _ptr(sa), _ptr(sa_fields_out) if sa is not None else None,
            None, ctypes.byref(p), stream_handle())
        _abi.check(rc, "lfm_step_color_f32")
