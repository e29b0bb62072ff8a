"""Beam propagation model: the reference's ``bpm_forward`` / ``bpm_reconstruct`` with
``backend="hip"``, a device-resident ``BpmPlan`` for callers that keep their arrays in device
memory, and ``BpmFunction``, the model as a ``torch.autograd.Function`` whose backward is the
hand-written adjoint.

Mirrors python/qoraholo/__init__.py:241-295 and include/qoraholo/bpm.hpp of the reference.
The volume passed to the device is the contrast x = ri - n_medium, formed in float32 as the
reference's src/api.cpp:159-163 does; results get n_medium added back in float32
(src/api.cpp:199).

Device options beyond the reference: ``chunk`` (illuminations per device pass; 0 picks the
largest that keeps the adjoint's buffers within half the free device memory) and
``checkpoint`` (slices per recomputed segment of the adjoint; 0 keeps every envelope, as the
reference's CUDA backend does, src/cuda/bpm.cu:139-141; -1, the default, keeps every envelope
when one pass's illuminations then fit within half the free device memory and otherwise
recomputes them from checkpoints every 16 slices).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native

DEFAULT_CHECKPOINT = -1


def _backend(backend: str) -> None:
    if backend != "hip":
        raise ValueError(f"unknown backend {backend!r}: this package provides backend='hip'")


def _illum(illuminations) -> np.ndarray:
    return np.ascontiguousarray(np.reshape(illuminations, (-1, 2)), dtype=np.float32)


def _opts(obliquity: bool, chunk: int, checkpoint: int) -> _native.BpmOptions_t:
    return _native.BpmOptions_t(int(bool(obliquity)), int(chunk), int(checkpoint))


def _volume(ri, who: str) -> np.ndarray:
    v = np.ascontiguousarray(ri, dtype=np.float32)
    if v.ndim != 3 or v.shape[1] != v.shape[2]:
        raise ValueError(f"{who}: volume must be (nz, n, n)")
    return v


def bpm_tables(n: int, nz: int, illuminations, geometry, obliquity: bool = True):
    """(step, last, q, phase, lipschitz): the per-illumination tables the model uses, as the
    reference's src/bpm.cpp:10-60 computes them (double on the host, stored in float)."""
    il = _illum(illuminations)
    count = il.shape[0]
    step = np.empty((count, n, n), np.complex64)
    last = np.empty((count, n, n), np.complex64)
    q = np.empty((count, n, n), np.float32)
    phase = ctypes.c_float()
    lip = ctypes.c_float()
    _native.check(_native.lib().qh_bpm_tables(
        int(n), int(nz), _native.ptr(il), count, ctypes.byref(geometry._c()), int(bool(obliquity)),
        _native.ptr(step), _native.ptr(last), _native.ptr(q), ctypes.byref(phase),
        ctypes.byref(lip)))
    return step, last, q, phase.value, lip.value


def bpm_forward(ri, illuminations, geometry, obliquity: bool = True, backend: str = "hip",
                chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
    """Refractive index (nz, n, n) to field ratios (count, n, n) by multi-slice beam propagation.

    One slice per voxel: a thin phase screen, then angular-spectrum propagation in the medium.
    The output is refocused to z = 0 and band-limited to the detection NA like
    ``forward_fields``. ``obliquity`` weights the scattered part of each slice by the obliquity
    of the scattered wave (the weak-object limit is then the first Born approximation at every
    illumination angle).
    """
    _backend(backend)
    v = _volume(ri, "bpm_forward")
    x = np.ascontiguousarray(v - np.float32(geometry.n_medium), dtype=np.float32)
    return bpm_forward_contrast(x, illuminations, geometry, obliquity, chunk, checkpoint)


def bpm_forward_contrast(x, illuminations, geometry, obliquity: bool = True, chunk: int = 0,
                         checkpoint: int = DEFAULT_CHECKPOINT):
    """The model on the contrast x = ri - n_medium (the reference's cuda::bpm_forward)."""
    x = _volume(x, "bpm_forward")
    il = _illum(illuminations)
    nz, n, _ = x.shape
    out = np.empty((il.shape[0], n, n), np.complex64)
    _native.check(_native.lib().qh_hip_bpm_forward(
        _native.ptr(x), nz, n, _native.ptr(il), il.shape[0], ctypes.byref(geometry._c()),
        ctypes.byref(_opts(obliquity, chunk, checkpoint)), _native.ptr(out)))
    return out


def bpm_vjp(x, r, illuminations, geometry, obliquity: bool = True, chunk: int = 0,
            checkpoint: int = DEFAULT_CHECKPOINT):
    """Re(J^H r): the adjoint of the model linearised at the contrast x, applied to r
    (count, n, n) (the reference's cuda::bpm_vjp, src/backends.hpp:124)."""
    x = _volume(x, "bpm_vjp")
    il = _illum(illuminations)
    nz, n, _ = x.shape
    r = np.ascontiguousarray(r, dtype=np.complex64)
    if r.shape != (il.shape[0], n, n):
        raise ValueError("bpm_vjp: r must be (count, n, n)")
    grad = np.empty_like(x)
    _native.check(_native.lib().qh_hip_bpm_vjp(
        _native.ptr(x), _native.ptr(r), nz, n, _native.ptr(il), il.shape[0],
        ctypes.byref(geometry._c()), ctypes.byref(_opts(obliquity, chunk, checkpoint)),
        _native.ptr(grad)))
    return grad


def bpm_jvp(x, h, illuminations, geometry, obliquity: bool = True, chunk: int = 0,
            checkpoint: int = DEFAULT_CHECKPOINT):
    """J h: the model linearised at the contrast x, applied to a direction h (nz, n, n); the
    tangent fields (count, n, n) (the reference's cpu::bpm_jvp, src/cpu/bpm.cpp:141-176)."""
    x = _volume(x, "bpm_jvp")
    h = _volume(h, "bpm_jvp")
    if h.shape != x.shape:
        raise ValueError("bpm_jvp: h must have the shape of x")
    il = _illum(illuminations)
    nz, n, _ = x.shape
    out = np.empty((il.shape[0], n, n), np.complex64)
    _native.check(_native.lib().qh_hip_bpm_jvp(
        _native.ptr(x), _native.ptr(h), nz, n, _native.ptr(il), il.shape[0],
        ctypes.byref(geometry._c()), ctypes.byref(_opts(obliquity, chunk, checkpoint)),
        _native.ptr(out)))
    return out


class BpmTables:
    """The tables of one volume shape and illumination set (the reference's BpmTables,
    src/backends.hpp:61-70): step, last (count, n, n) complex64, q (count, n, n) float32, the
    screen phase and the Lipschitz constant. The ``*_tables`` calls run the model on them, as
    the reference's backends take them (src/backends.hpp:119-124)."""

    def __init__(self, step, last, q, phase: float, lipschitz: float, nz: int):
        self.step = np.ascontiguousarray(step, dtype=np.complex64)
        self.last = np.ascontiguousarray(last, dtype=np.complex64)
        self.q = np.ascontiguousarray(q, dtype=np.float32)
        if self.step.ndim != 3 or self.step.shape != self.last.shape or self.step.shape != self.q.shape:
            raise ValueError("BpmTables: step, last and q must be (count, n, n)")
        self.phase = float(phase)
        self.lipschitz = float(lipschitz)
        self.nz = int(nz)
        self.count, self.n, _ = self.step.shape

    @classmethod
    def build(cls, n: int, nz: int, illuminations, geometry, obliquity: bool = True):
        step, last, q, phase, lip = bpm_tables(n, nz, illuminations, geometry, obliquity)
        return cls(step, last, q, phase, lip, nz)

    def _c(self):
        return _native.BpmTableView_t(self.n, self.nz, self.count, self.phase,
                                      self.step.ctypes.data, self.last.ctypes.data,
                                      self.q.ctypes.data, self.lipschitz)

    def _x(self, x, who):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.nz, self.n, self.n):
            raise ValueError(f"{who}: volume must be (nz, n, n) of the tables")
        return x

    def _fields(self, f, who):
        f = np.ascontiguousarray(f, dtype=np.complex64)
        if f.shape != (self.count, self.n, self.n):
            raise ValueError(f"{who}: fields must be (count, n, n) of the tables")
        return f

    def forward(self, x, chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
        x = self._x(x, "bpm_forward")
        out = np.empty((self.count, self.n, self.n), np.complex64)
        _native.check(_native.lib().qh_hip_bpm_forward_tables(
            _native.ptr(x), ctypes.byref(self._c()), ctypes.byref(_opts(True, chunk, checkpoint)),
            _native.ptr(out)))
        return out

    def vjp(self, x, r, chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
        x = self._x(x, "bpm_vjp")
        r = self._fields(r, "bpm_vjp")
        grad = np.empty_like(x)
        _native.check(_native.lib().qh_hip_bpm_vjp_tables(
            _native.ptr(x), _native.ptr(r), ctypes.byref(self._c()),
            ctypes.byref(_opts(True, chunk, checkpoint)), _native.ptr(grad)))
        return grad

    def jvp(self, x, h, chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
        x = self._x(x, "bpm_jvp")
        h = self._x(h, "bpm_jvp")
        out = np.empty((self.count, self.n, self.n), np.complex64)
        _native.check(_native.lib().qh_hip_bpm_jvp_tables(
            _native.ptr(x), _native.ptr(h), ctypes.byref(self._c()),
            ctypes.byref(_opts(True, chunk, checkpoint)), _native.ptr(out)))
        return out

    def reconstruct(self, fields, x, iterations: int = 20, batch: int = 0, step: float = 1.0,
                    tv_weight: float = 0.0, tv_inner: int = 10, chunk: int = 0,
                    checkpoint: int = DEFAULT_CHECKPOINT):
        """(contrast, loss) from the starting contrast x, as cuda::bpm_reconstruct."""
        f = self._fields(fields, "bpm_reconstruct")
        x = self._x(x, "bpm_reconstruct").copy()
        solve = _native.BpmSolveOptions_t(int(iterations), int(batch), float(step),
                                          float(tv_weight), int(tv_inner))
        loss = np.zeros(max(int(iterations), 0), np.float64)
        _native.check(_native.lib().qh_hip_bpm_reconstruct_tables(
            _native.ptr(f), _native.ptr(x), ctypes.byref(self._c()),
            ctypes.byref(_opts(True, chunk, checkpoint)), ctypes.byref(solve),
            _native.ptr(loss) if loss.size else None))
        return x, loss


def bpm_reconstruct(fields, illuminations, geometry, initial, iterations: int = 20, batch: int = 0,
                    step: float = 1.0, tv_weight: float = 0.0, tv_inner: int = 10,
                    obliquity: bool = True, backend: str = "hip", return_loss: bool = False,
                    chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
    """Field ratios (count, n, n) to refractive index by gradient descent on the BPM.

    ``initial`` (nz, n, n) is the starting volume, e.g. a Rytov reconstruction. Each iteration
    back-propagates the residual for the gradient, then applies n >= n_medium and, with
    ``tv_weight > 0``, a TV proximal step. ``batch=0`` uses all illuminations per step with
    FISTA momentum; otherwise one iteration is a pass over batches of that many illuminations.
    ``step`` multiplies the automatic step 1 / L. With ``return_loss`` the data term of every
    iteration is returned as well.
    """
    _backend(backend)
    f = np.ascontiguousarray(fields, dtype=np.complex64)
    init = _volume(initial, "bpm_reconstruct")
    if f.ndim != 3 or f.shape[1:] != init.shape[1:]:
        raise ValueError("bpm_reconstruct: fields (count, n, n) and initial (nz, n, n) must share n")
    il = _illum(illuminations)
    count, n, _ = f.shape
    if il.shape[0] != count:
        raise ValueError("bpm_reconstruct: need one illumination per field")
    nz = init.shape[0]
    x = np.ascontiguousarray(init - np.float32(geometry.n_medium), dtype=np.float32)
    solve = _native.BpmSolveOptions_t(int(iterations), int(batch), float(step), float(tv_weight),
                                      int(tv_inner))
    loss = np.zeros(max(int(iterations), 0), np.float64)
    _native.check(_native.lib().qh_hip_bpm_reconstruct(
        _native.ptr(f), count, n, _native.ptr(il), ctypes.byref(geometry._c()), _native.ptr(x),
        nz, ctypes.byref(_opts(obliquity, chunk, checkpoint)), ctypes.byref(solve),
        _native.ptr(loss) if loss.size else None))
    ri = (x + np.float32(geometry.n_medium)).astype(np.float32)
    return (ri, loss) if return_loss else ri


def _addr(t) -> int:
    """Device address of a contiguous torch tensor (or an integer address)."""
    if isinstance(t, int):
        return t
    if not t.is_contiguous():
        raise ValueError("BpmPlan: tensors must be contiguous")
    return int(t.data_ptr())


def _stream(stream):
    """hipStream_t of a torch stream or an address; None: torch's current stream."""
    if stream is None:
        import torch

        stream = torch.cuda.current_stream()
    s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
    return ctypes.c_void_p(s)


class BpmPlan:
    """Device-resident BPM of one volume shape and illumination set.

    The tables and work buffers stay in device memory; ``forward``, ``vjp`` and ``gradient``
    take torch tensors on the GPU (or device addresses) and run on torch's current stream
    (or ``stream``: a torch stream or a hipStream_t address). Volumes are contrasts x = ri - n_medium, float32
    (nz, n, n); fields and residuals complex64 (count, n, n). One thread per plan.
    """

    def __init__(self, nz: int, n: int, illuminations, geometry, obliquity: bool = True,
                 chunk: int = 0, checkpoint: int = DEFAULT_CHECKPOINT):
        il = _illum(illuminations)
        self.nz, self.n, self.count = int(nz), int(n), il.shape[0]
        self.geometry = geometry
        self._h = None
        h = ctypes.c_void_p()
        _native.check(_native.lib().qh_hip_bpm_plan_create(
            self.nz, self.n, _native.ptr(il), self.count, ctypes.byref(geometry._c()),
            ctypes.byref(_opts(obliquity, chunk, checkpoint)), ctypes.byref(h)))
        self._destroy = _native.lib().qh_hip_bpm_plan_destroy
        self._h = h
        phase, lip, ch, nbytes = ctypes.c_float(), ctypes.c_float(), ctypes.c_int(), ctypes.c_size_t()
        _native.check(_native.lib().qh_hip_bpm_plan_info(h, ctypes.byref(phase), ctypes.byref(lip),
                                                         ctypes.byref(ch), ctypes.byref(nbytes)))
        self.phase, self.lipschitz, self.chunk = phase.value, lip.value, ch.value
        self._bytes = nbytes
        cp = ctypes.c_int()
        _native.check(_native.lib().qh_hip_bpm_plan_checkpoint(h, ctypes.byref(cp)))
        self.checkpoint = cp.value  # as resolved: 0 when every envelope is kept

    def __del__(self):
        # The destroy function is held by the object: module globals may already be gone when
        # the interpreter shuts down.
        if getattr(self, "_h", None) and getattr(self, "_destroy", None):
            self._destroy(self._h)
            self._h = None

    @property
    def device_bytes(self) -> int:
        nbytes = ctypes.c_size_t()
        _native.check(_native.lib().qh_hip_bpm_plan_info(self._h, None, None, None,
                                                         ctypes.byref(nbytes)))
        return nbytes.value

    def _check_x(self, x):
        if hasattr(x, "shape") and tuple(x.shape) != (self.nz, self.n, self.n):
            raise ValueError("BpmPlan: x must be (nz, n, n)")

    def _check_fields(self, f, who):
        if hasattr(f, "shape") and tuple(f.shape) != (self.count, self.n, self.n):
            raise ValueError(f"BpmPlan: {who} must be (count, n, n)")

    def forward(self, x, out=None, stream=None):
        """Field ratios (count, n, n) at z = 0 of the contrast x."""
        import torch

        self._check_x(x)
        if out is None:
            out = torch.empty((self.count, self.n, self.n), dtype=torch.complex64, device=x.device)
        self._check_fields(out, "out")
        _native.check(_native.lib().qh_hip_bpm_forward_dev(self._h, _addr(x), _addr(out),
                                                           _stream(stream)))
        return out

    def vjp(self, x, r, out=None, stream=None):
        """Re(J^H r) at the contrast x."""
        import torch

        self._check_x(x)
        self._check_fields(r, "r")
        if out is None:
            out = torch.empty((self.nz, self.n, self.n), dtype=torch.float32, device=x.device)
        self._check_x(out)
        _native.check(_native.lib().qh_hip_bpm_vjp_dev(self._h, _addr(x), _addr(r), _addr(out),
                                                       _stream(stream)))
        return out

    def jvp(self, x, h, out=None, stream=None):
        """J h at the contrast x: tangent fields (count, n, n)."""
        import torch

        self._check_x(x)
        self._check_x(h)
        if out is None:
            out = torch.empty((self.count, self.n, self.n), dtype=torch.complex64, device=x.device)
        self._check_fields(out, "out")
        _native.check(_native.lib().qh_hip_bpm_jvp_dev(self._h, _addr(x), _addr(h), _addr(out),
                                                       _stream(stream)))
        return out

    def gradient(self, x, fields, first: int = 0, last: int | None = None, out=None, stream=None):
        """(grad, loss) of the reference's gradient(): Re(J^H (bpm(x) - fields)) / (last - first)
        over illuminations [first, last) and 0.5 / count times the squared residual."""
        import torch

        self._check_x(x)
        self._check_fields(fields, "fields")
        last = self.count if last is None else last
        if out is None:
            out = torch.empty((self.nz, self.n, self.n), dtype=torch.float32, device=x.device)
        loss = ctypes.c_double()
        _native.check(_native.lib().qh_hip_bpm_gradient_dev(
            self._h, _addr(x), _addr(fields), int(first), int(last), _addr(out),
            ctypes.byref(loss), _stream(stream)))
        return out, loss.value


def _function():
    import torch

    class BpmFunction(torch.autograd.Function):
        """y = bpm(x) for a contrast tensor x on the GPU; backward is Re(J^H grad_y), the
        hand-written adjoint (with the checkpointed recomputation of the envelopes)."""

        @staticmethod
        def forward(ctx, x, plan):
            x = x.contiguous()
            ctx.plan = plan
            ctx.save_for_backward(x)
            return plan.forward(x)

        @staticmethod
        def backward(ctx, grad_y):
            (x,) = ctx.saved_tensors
            return ctx.plan.vjp(x, grad_y.contiguous()), None

    return BpmFunction


_BPM_FUNCTION = None


def bpm(x, plan: BpmPlan):
    """Differentiable model: field ratios (count, n, n) complex64 of the contrast tensor x."""
    global _BPM_FUNCTION
    if _BPM_FUNCTION is None:
        _BPM_FUNCTION = _function()
    return _BPM_FUNCTION.apply(x, plan)
