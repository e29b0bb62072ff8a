"""MI355X holotomography backend: QoraHolo's numpy API with ``backend="hip"``.

Mirrors ``python/qoraholo/__init__.py`` of the reference: arrays are numpy, centre-origin,
square with even sides; image functions take a 2D image or a 3D stack (count, n, n) and return
the same rank; lengths in micrometres, spatial frequencies in cycles per micrometre. The
compute runs in ``libqoraholo_hip.so`` (hand-written gfx950 kernels behind the C-ABI of
``include/qoraholo_hip.h``). ``field_ratio`` and ``find_sideband`` are host functions, as in the
reference. Only ``backend="hip"`` is served here; the reference's own CPU backend is not part
of this package.

With PyTorch in the same process, import torch first: torch carries its own HIP runtime under
the same soname, and the library then shares it; loaded the other way round the process would
hold two runtimes.
"""

from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass

import numpy as np

from . import _native
from ._native import LogicError, OutOfDeviceMemory
from .bpm import BpmPlan, BpmTables, bpm, bpm_forward, bpm_jvp, bpm_reconstruct, bpm_tables, bpm_vjp
from .phantom import Ellipsoid, Sphere, forward_fields, render_phantom, synthesize_holograms

__all__ = [
    "BpmPlan",
    "BpmTables",
    "Ellipsoid",
    "Geometry",
    "LogicError",
    "OutOfDeviceMemory",
    "Pipeline",
    "Server",
    "Sideband",
    "Sphere",
    "bpm",
    "bpm_forward",
    "bpm_jvp",
    "bpm_reconstruct",
    "bpm_tables",
    "bpm_vjp",
    "circular_scan",
    "cuda_available",
    "extract_fields",
    "field_ratio",
    "estimate_illuminations",
    "find_sideband",
    "find_sidebands",
    "forward_fields",
    "hip_available",
    "kspace",
    "reconstruct",
    "render_phantom",
    "rytov",
    "synthesize_holograms",
    "unwrap",
]


@dataclass
class Geometry:
    """include/qoraholo/odt.hpp:13-18."""

    wavelength: float = 0.532
    n_medium: float = 1.337
    na: float = 1.2
    pixel: float = 0.1

    def _c(self) -> _native.Geometry_t:
        return _native.Geometry_t(self.wavelength, self.n_medium, self.na, self.pixel)


@dataclass
class Sideband:
    """include/qoraholo/field.hpp:13-19: peak offset from DC in pixels."""

    iy: int = 0
    ix: int = 0
    y: float = 0.0
    x: float = 0.0
    magnitude: float = 0.0


def hip_available() -> bool:
    """True when the library is built and a HIP device is visible."""
    try:
        return bool(_native.lib().qh_hip_device_present())
    except (OSError, RuntimeError):
        return False


def cuda_available() -> bool:
    """This package has no CUDA backend."""
    return False


def _backend(backend: str) -> None:
    if backend != "hip":
        raise ValueError(f"unknown backend {backend!r}: this package provides backend='hip'")


def _stack(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim == 2:
        return a[None], True
    if a.ndim != 3:
        raise ValueError(f"expected a 2D image or a 3D stack, got {a.ndim} dimensions")
    return a, False


def _unstack(a, single):
    return a[0] if single else a


def _illum(illuminations) -> np.ndarray:
    return np.ascontiguousarray(np.reshape(illuminations, (-1, 2)), dtype=np.float32)


def circular_scan(count: int, na_illumination: float, geometry: Geometry) -> np.ndarray:
    """Normal incidence, then count - 1 directions evenly spaced on a circle: (count, 2)."""
    if count < 1:
        raise ValueError("circular_scan: count must be positive")
    out = np.zeros((count, 2), np.float32)
    _native.check(_native.lib().qh_circular_scan(int(count), float(na_illumination),
                                                 ctypes.byref(geometry._c()), _native.ptr(out)))
    return out


def find_sideband(hologram, dc_exclusion: float = 0.0) -> Sideband:
    """Peak of the +1 sideband (x > 0 half plane) of the first hologram, in pixels from DC.

    Host function (src/cpu/field.cpp:35-60): spectrum of the (-1)^(x+y)-modulated hologram,
    largest magnitude outside the DC disc, refined by Jacobsen's three-bin estimator with
    Candan's correction.
    """
    h, _ = _stack(hologram, np.float32)
    h = h[0]
    n = h.shape[-1]
    if h.shape[0] != n or n % 2 or n < 4:
        raise ValueError("find_sideband: holograms must be square with an even side")
    dc = np.float32(dc_exclusion)
    if dc <= 0:
        dc = np.float32(n) / np.float32(8)
    k = np.arange(n)
    sign = np.where((k[:, None] + k[None, :]) % 2 == 0, np.float32(1), np.float32(-1))
    s = np.fft.fft2((h * sign).astype(np.complex64)).astype(np.complex64)
    c = n // 2
    ys = np.arange(1, n - 1)[:, None]
    xs = np.arange(c, n - 1)[None, :]
    dy = ys - c
    dx = xs - c
    valid = ~((dx == 0) & (dy <= 0)) & ((dy * dy + dx * dx).astype(np.float32) > dc * dc)
    mag = np.abs(s[1:n - 1, c:n - 1]).astype(np.float32)
    mag = np.where(valid, mag, np.float32(0))
    flat = int(np.argmax(mag))
    best = float(mag.flat[flat])
    if best == 0:
        raise RuntimeError("find_sideband: no peak outside the DC region")
    y = 1 + flat // mag.shape[1]
    x = c + flat % mag.shape[1]

    def refine(lo, mid, hi):
        den = np.complex64(2) * mid - lo - hi
        if np.abs(den) == 0:
            return np.float32(0)
        cc = np.float32(np.pi) / np.float32(n)
        d = np.float32(((lo - hi) / den).real) * np.float32(np.tan(cc)) / cc
        return np.float32(min(max(d, -0.5), 0.5))

    ry = refine(s[y - 1, x], s[y, x], s[y + 1, x])
    rx = refine(s[y, x - 1], s[y, x], s[y, x + 1])
    return Sideband(int(y - c), int(x - c), float(np.float32(y - c) + ry),
                    float(np.float32(x - c) + rx), best)


def find_sidebands(holograms, dc_exclusion: float = 0.0) -> list[Sideband]:
    """The sideband of every hologram of a stack, found on the device.

    ``find_sideband`` (src/cpu/field.cpp:35-60) per image: spectra and peak search run on the
    device, the sub-bin refinement on the host; the reference runs it on the host for one
    hologram at a time.
    """
    h, _ = _stack(holograms, np.float32)
    count, ny, n = h.shape
    if ny != n or n % 2 or n < 4:
        raise ValueError("find_sideband: holograms must be square with an even side")
    out = np.empty((count, 5), np.float32)
    _native.check(_native.lib().qh_hip_find_sidebands(_native.ptr(h), count, n,
                                                      float(dc_exclusion), _native.ptr(out)))
    return [Sideband(int(r[0]), int(r[1]), float(r[2]), float(r[3]), float(r[4])) for r in out]


def estimate_illuminations(backgrounds, hologram_geometry: Geometry, dc_exclusion: float = 0.0):
    """(carrier, illuminations): the carrier is the sideband of background 0 and illumination a
    is (s_a - carrier) / (N pixel) in cycles per micrometre, as the reference's
    examples/readme_figure.py:39-46 estimates the angles; on the device."""
    h, _ = _stack(backgrounds, np.float32)
    count, ny, n = h.shape
    if ny != n or n % 2 or n < 4:
        raise ValueError("estimate_illuminations: holograms must be square with an even side")
    out = np.empty((count, 2), np.float32)
    c = np.empty(5, np.float32)
    _native.check(_native.lib().qh_hip_estimate_illuminations(
        _native.ptr(h), count, n, float(hologram_geometry.pixel), float(dc_exclusion),
        _native.ptr(out), _native.ptr(c)))
    return Sideband(int(c[0]), int(c[1]), float(c[2]), float(c[3]), float(c[4])), out


def _crop(n: int, center, radius: float, output_size: int) -> _native.CropPlan_t:
    plan = _native.CropPlan_t()
    _native.check(_native.lib().qh_resolve_crop(float(center[0]), float(center[1]), float(radius),
                                                int(output_size), int(n), ctypes.byref(plan)))
    return plan


def extract_fields(holograms, center=None, radius: float = 0.0, dc_exclusion: float = 0.0,
                   output_size: int = 0, backend: str = "hip"):
    """Off-axis hologram(s) to complex field(s) (src/api.cpp:62-81).

    ``center`` is the sideband centre (y, x) in pixels from DC; detected when omitted. For
    multi-angle data pass the reference carrier so the illumination tilt stays in the field.
    ``radius`` defaults to |center| / 3, ``output_size`` to the hologram side.
    """
    _backend(backend)
    h, single = _stack(holograms, np.float32)
    count, ny, nx = h.shape
    if ny != nx or nx < 4 or nx % 2:
        raise ValueError("extract_fields: images must be square with an even side")
    if center is None:
        sb = find_sideband(h, dc_exclusion)
        center = (float(sb.iy), float(sb.ix))
    plan = _crop(nx, center, radius, output_size)
    out = np.empty((count, plan.out, plan.out), np.complex64)
    _native.check(_native.lib().qh_hip_extract_fields(_native.ptr(h), count, nx,
                                                      ctypes.byref(plan), _native.ptr(out)))
    return _unstack(out, single)


def field_ratio(sample, background):
    """sample / background; background is one field or one per sample field (host, on the
    library's copy threads, with the reference's std::complex<float> division)."""
    s, single = _stack(sample, np.complex64)
    b, _ = _stack(background, np.complex64)
    if s.shape[1:] != b.shape[1:] or (b.shape[0] != 1 and b.shape[0] != s.shape[0]):
        raise ValueError("field_ratio: background shape does not match")
    out = np.empty_like(s)
    _native.check(_native.lib().qh_field_ratio(_native.ptr(s), s.shape[0], _native.ptr(b),
                                               b.shape[0], s.shape[1] * s.shape[2],
                                               _native.ptr(out)))
    return _unstack(out, single)


def unwrap(phase, backend: str = "hip"):
    """Least-squares unwrapping (Ghiglia & Romero 1994), congruent to the input mod 2 pi."""
    _backend(backend)
    p, single = _stack(phase, np.float32)
    nz, ny, nx = p.shape
    out = np.empty_like(p)
    _native.check(_native.lib().qh_hip_unwrap(_native.ptr(p), nz, ny, nx, _native.ptr(out)))
    return _unstack(out, single)


def rytov(ratio, backend: str = "hip"):
    """log|r| + i unwrap(arg r), with the mean phase on the image border in (-pi, pi]."""
    _backend(backend)
    r, single = _stack(ratio, np.complex64)
    nz, ny, nx = r.shape
    out = np.empty_like(r)
    _native.check(_native.lib().qh_hip_rytov(_native.ptr(r), nz, ny, nx, _native.ptr(out)))
    return _unstack(out, single)


_SOLVERS = {"gp": 0, "tv": 1}
_TRANSFORMS = {"r2c": 0, "c2c": 1}
_MAPPINGS = {"gather": 0, "scatter": 1}


def _options(nz, iterations, solver, gp_transform, tv_weight, tv_inner, mapping):
    for name, table, v in (("solver", _SOLVERS, solver), ("gp_transform", _TRANSFORMS, gp_transform),
                           ("mapping", _MAPPINGS, mapping)):
        if v not in table:
            raise ValueError(f"unknown {name} {v!r}")
    return _native.OdtOptions_t(int(nz), int(iterations), _SOLVERS[solver],
                                _TRANSFORMS[gp_transform], float(tv_weight), int(tv_inner),
                                _MAPPINGS[mapping])


def _grid(n: int, nz: int, geometry: Geometry) -> _native.OdtGrid_t:
    g = _native.OdtGrid_t()
    _native.check(_native.lib().qh_make_grid(int(n), int(nz) if nz > 0 else int(n),
                                             ctypes.byref(geometry._c()), ctypes.byref(g)))
    return g


def reconstruct(fields, illuminations, geometry: Geometry, nz: int = 0, iterations: int = 0,
                solver: str = "gp", gp_transform: str = "r2c", tv_weight: float = 1e-2,
                tv_inner: int = 10, mapping: str = "gather", backend: str = "hip",
                return_filled: bool = False, timings: dict | None = None):
    """Rytov fields (count, n, n) to a refractive-index volume (nz, n, n).

    As the reference's ``reconstruct`` (src/api.cpp:125-137). With ``timings`` (a dict) the
    device stage times of the call are stored into it.
    """
    _backend(backend)
    f, _ = _stack(fields, np.complex64)
    count, ny, n = f.shape
    if ny != n:
        raise ValueError("reconstruct: images must be square")
    il = _illum(illuminations)
    if il.shape[0] != count:
        raise ValueError("reconstruct: need one illumination per field")
    opts = _options(nz, iterations, solver, gp_transform, tv_weight, tv_inner, mapping)
    nzv = nz if nz > 0 else n
    ri = np.empty((nzv, n, n), np.float32)
    filled = ctypes.c_double()
    t = _native.OdtTimings_t()
    grid = _grid(n, nzv, geometry) if n >= 1 and nzv >= 1 else None
    _native.check(_native.lib().qh_hip_reconstruct(
        _native.ptr(f), count, n, _native.ptr(il), ctypes.byref(grid) if grid else None,
        ctypes.byref(geometry._c()), ctypes.byref(opts), _native.ptr(ri), ctypes.byref(filled),
        ctypes.byref(t) if timings is not None else None))
    if timings is not None:
        timings.update({k: getattr(t, k) for k, _ in t._fields_})
    return (ri, filled.value) if return_filled else ri


def kspace(fields, illuminations, geometry: Geometry, nz: int = 0):
    """The Ewald mapping of reconstruct(): normalised k-space potential (nz, n, n) complex64,
    hit counts (nz, n, n) uint16 and the filled fraction."""
    f, _ = _stack(fields, np.complex64)
    count, _, n = f.shape
    il = _illum(illuminations)
    if il.shape[0] != count:
        raise ValueError("kspace: need one illumination per field")
    nzv = nz if nz > 0 else n
    grid = _grid(n, nzv, geometry)
    pot = np.empty((nzv, n, n), np.complex64)
    counts = np.empty((nzv, n, n), np.uint16)
    filled = ctypes.c_double()
    _native.check(_native.lib().qh_hip_map_kspace(_native.ptr(f), count, n, _native.ptr(il),
                                                  ctypes.byref(grid), _native.ptr(pot),
                                                  _native.ptr(counts), ctypes.byref(filled)))
    return pot, counts, filled.value


def _pipeline_options(count, hologram_size, center, illuminations, geometry, radius, output_size,
                      nz, iterations, solver, gp_transform, tv_weight, tv_inner, mapping,
                      upload_chunks):
    """PipelineOptions (include/qoraholo/pipeline.hpp:15-25) as the C-ABI struct, and the
    illumination array it points to (keep both alive together)."""
    if center is None:
        raise ValueError("Pipeline: field.center (the carrier) is required")
    if count < 1:
        raise ValueError("Pipeline: count must be positive")
    il = _illum(illuminations)
    if il.shape[0] != count:
        raise ValueError("Pipeline: need one illumination per hologram")
    illum_c = (_native.Illumination_t * max(1, il.shape[0]))()
    for a in range(il.shape[0]):
        illum_c[a] = _native.Illumination_t(il[a, 0], il[a, 1])
    plan = _crop(hologram_size, center, radius, output_size)
    opts = _options(nz, iterations, solver, gp_transform, tv_weight, tv_inner, mapping)
    o = _native.PipelineOptions_t(int(count), int(hologram_size), plan, illum_c, geometry._c(),
                                  opts, int(upload_chunks))
    return o, illum_c


class Pipeline:
    """Holograms to refractive index with every buffer allocated once on the device.

    As the reference's ``Pipeline`` (include/qoraholo/pipeline.hpp, python/qoraholo/__init__.py
    :146-218): background fields stay in HBM and each ``run`` does one upload of the sample
    holograms and one download of the volume. ``run_device`` takes holograms already in device
    memory (a torch tensor or an address) and writes the volume to device memory.
    """

    def __init__(self, count: int, hologram_size: int, center, illuminations,
                 geometry: Geometry, radius: float = 0.0, output_size: int = 0, nz: int = 0,
                 iterations: int = 0, solver: str = "gp", gp_transform: str = "r2c",
                 tv_weight: float = 1e-2, tv_inner: int = 10, mapping: str = "gather",
                 backend: str = "hip", upload_chunks: int = 1):
        _backend(backend)
        self._handle = None
        o, self._illum_c = _pipeline_options(count, hologram_size, center, illuminations, geometry,
                                             radius, output_size, nz, iterations, solver,
                                             gp_transform, tv_weight, tv_inner, mapping,
                                             upload_chunks)
        handle = ctypes.c_void_p()
        _native.check(_native.lib().qh_hip_pipeline_create(ctypes.byref(o), ctypes.byref(handle)))
        self._destroy = _native.lib().qh_hip_pipeline_destroy
        self._handle = handle
        self.count = count
        self.hologram_size = hologram_size
        nzc, nyc, nxc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _native.check(_native.lib().qh_hip_pipeline_output_shape(
            self._handle, ctypes.byref(nzc), ctypes.byref(nyc), ctypes.byref(nxc)))
        self.output_shape = (nzc.value, nyc.value, nxc.value)

    def __del__(self):
        # The destroy function is held by the object: module globals may already be gone when
        # the interpreter shuts down.
        if getattr(self, "_handle", None) and getattr(self, "_destroy", None):
            self._destroy(self._handle)
            self._handle = None

    def _host(self, addr, shape):
        buf = (ctypes.c_float * int(np.prod(shape))).from_address(addr)
        return np.frombuffer(buf, np.float32).reshape(shape)

    def set_background(self, backgrounds) -> None:
        """Background holograms (count, N, N); their fields are extracted once and kept."""
        b = np.ascontiguousarray(backgrounds, dtype=np.float32)
        if b.shape != (self.count, self.hologram_size, self.hologram_size):
            raise ValueError("Pipeline::set_background: shape must be (count, N, N)")
        _native.check(_native.lib().qh_hip_pipeline_set_background(self._handle, _native.ptr(b)))

    def input(self) -> np.ndarray:
        """The page-locked input buffer (count, N, N)."""
        return self._host(_native.lib().qh_hip_pipeline_input(self._handle),
                          (self.count, self.hologram_size, self.hologram_size))

    def output(self) -> np.ndarray:
        """The page-locked output buffer (nz, n, n) of the last process()."""
        return self._host(_native.lib().qh_hip_pipeline_output(self._handle), self.output_shape)

    def process(self) -> None:
        _native.check(_native.lib().qh_hip_pipeline_process(self._handle))

    def run(self, holograms) -> np.ndarray:
        """Sample holograms (count, N, N) to the index volume (nz, n, n). The copies through the
        page-locked buffers run on the library's copy threads and overlap the transfers."""
        h = np.ascontiguousarray(holograms, dtype=np.float32)
        if h.shape != (self.count, self.hologram_size, self.hologram_size):
            raise ValueError("Pipeline::run: shape must be (count, N, N)")
        out = np.empty(self.output_shape, np.float32)
        _native.check(_native.lib().qh_hip_pipeline_run(self._handle, _native.ptr(h),
                                                         _native.ptr(out)))
        return out

    def run_device(self, holograms, out, stream=None) -> None:
        """Holograms in device memory -> index volume in device memory, queued on `stream`
        (a torch stream or a hipStream_t address; None: torch's current stream when torch is
        loaded, else the null stream)."""
        src = holograms.data_ptr() if hasattr(holograms, "data_ptr") else int(holograms)
        dst = out.data_ptr() if hasattr(out, "data_ptr") else int(out)
        if hasattr(holograms, "numel"):
            if holograms.numel() != self.count * self.hologram_size ** 2:
                raise ValueError("Pipeline::run_device: holograms must be (count, N, N)")
            if out.numel() != int(np.prod(self.output_shape)):
                raise ValueError("Pipeline::run_device: out must be (nz, n, n)")
        if stream is None:
            torch = sys.modules.get("torch")
            stream = torch.cuda.current_stream() if torch is not None else 0
        s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        _native.check(_native.lib().qh_hip_pipeline_process_device(
            self._handle, ctypes.c_void_p(src), ctypes.c_void_p(dst), ctypes.c_void_p(s)))

    def time_fgp(self, on: bool = True) -> None:
        """Record the FGP kernel intervals in ``timings`` (fgp_ms, fgp_steps)."""
        _native.check(_native.lib().qh_hip_pipeline_time_fgp(self._handle, int(bool(on))))

    def use_graph(self, on: bool = True) -> None:
        """Replay each request's device work as one captured hipGraph (the default) or launch
        it kernel by kernel; the results are the same bit for bit."""
        _native.check(_native.lib().qh_hip_pipeline_use_graph(self._handle, int(bool(on))))

    @property
    def filled(self) -> float:
        """Fraction of k-space voxels that received data in the last run."""
        return float(_native.lib().qh_hip_pipeline_filled(self._handle))

    @property
    def timings(self) -> dict:
        """Milliseconds per stage of the last run (device event intervals)."""
        t = _native.PipelineTimings_t()
        _native.check(_native.lib().qh_hip_pipeline_timings(self._handle, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    @property
    def setup_ms(self) -> float:
        return float(_native.lib().qh_hip_pipeline_setup_ms(self._handle))

    @property
    def device_bytes(self) -> int:
        return int(_native.lib().qh_hip_pipeline_device_bytes(self._handle))


class Server:
    """Served requests of several geometries admitted against an HBM budget.

    ``add_geometry`` registers the options of a Pipeline with its background holograms and
    returns a geometry id; ``submit`` queues sample holograms of a geometry (thread-safe) and
    returns a request id; ``wait`` returns (ri, filled, latency_ms). ``add_bpm`` registers a
    BPM inversion geometry (Rytov of the field ratios, a Gerchberg-Papoulis start, then
    ``bpm_reconstruct``), ``submit_bpm`` queues field ratios of one and ``wait`` returns
    (ri, loss, latency_ms) for it. Requests that arrive while others run are admitted into free
    device-resident slots (created while ``budget_bytes`` allows, at most ``max_slots`` in
    flight) and overlap on the device; each result is the one the request gets alone.
    """

    def __init__(self, budget_bytes: int = 0, max_slots: int = 4):
        self._handle = None
        handle = ctypes.c_void_p()
        _native.check(_native.lib().qh_hip_server_create(int(budget_bytes), int(max_slots),
                                                         ctypes.byref(handle)))
        self._destroy = _native.lib().qh_hip_server_destroy
        self._handle = handle
        self._geometries = {}
        self._pending = {}

    def __del__(self):
        # The destroy function is held by the object: module globals may already be gone when
        # the interpreter shuts down.
        if getattr(self, "_handle", None) and getattr(self, "_destroy", None):
            self._destroy(self._handle)
            self._handle = None

    def add_geometry(self, count: int, hologram_size: int, center, illuminations,
                     geometry: Geometry, backgrounds, radius: float = 0.0, output_size: int = 0,
                     nz: int = 0, iterations: int = 0, solver: str = "gp",
                     gp_transform: str = "r2c", tv_weight: float = 1e-2, tv_inner: int = 10,
                     mapping: str = "gather", max_batch: int = 1) -> int:
        """Registers a geometry. With ``center`` or ``illuminations`` None they are estimated
        from the backgrounds on the device (``estimate_illuminations``; the hologram pixel is
        ``geometry.pixel * output_size / hologram_size``). With ``max_batch`` > 1 a slot of the
        geometry takes up to that many queued requests of it at once (one front end for all
        their stacks, one graph replay for the batch); results are the same bits."""
        if center is None or illuminations is None:
            m = output_size if output_size > 0 else hologram_size
            hg = Geometry(geometry.wavelength, geometry.n_medium, geometry.na,
                          geometry.pixel * m / hologram_size)
            c, est = estimate_illuminations(backgrounds, hg)
            center = (float(c.iy), float(c.ix)) if center is None else center
            illuminations = est if illuminations is None else illuminations
        o, illum_c = _pipeline_options(count, hologram_size, center, illuminations, geometry,
                                       radius, output_size, nz, iterations, solver, gp_transform,
                                       tv_weight, tv_inner, mapping, 1)
        b = np.ascontiguousarray(backgrounds, dtype=np.float32)
        if b.shape != (count, hologram_size, hologram_size):
            raise ValueError("Server: backgrounds must be (count, N, N)")
        gid = ctypes.c_int()
        if max_batch > 1:
            _native.check(_native.lib().qh_hip_server_add_geometry_batched(
                self._handle, ctypes.byref(o), _native.ptr(b), int(max_batch), ctypes.byref(gid)))
        else:
            _native.check(_native.lib().qh_hip_server_add_geometry(
                self._handle, ctypes.byref(o), _native.ptr(b), ctypes.byref(gid)))
        m = o.crop.out
        self._geometries[gid.value] = ((count, hologram_size, hologram_size),
                                       (nz if nz > 0 else m, m, m), None)
        return gid.value

    def add_bpm(self, n: int, illuminations, geometry: Geometry, nz: int = 0,
                start_iterations: int = 20, iterations: int = 40, batch: int = 0,
                step: float = 1.0, tv_weight: float = 0.0, tv_inner: int = 10,
                obliquity: bool = True, chunk: int = 0, checkpoint: int = -1) -> int:
        """Registers a BPM inversion geometry: field ratios (count, n, n) to a refractive index
        (nz, n, n) by ``bpm_reconstruct`` with these options, started from the
        ``start_iterations`` Gerchberg-Papoulis reconstruction of their Rytov fields. An
        automatic ``chunk`` (0) and ``checkpoint`` (-1) are resolved from the free memory when
        the geometry's first slot is built and kept for its later slots."""
        il = _illum(illuminations)
        count = il.shape[0]
        nzv = nz if nz > 0 else n
        illum_c = (_native.Illumination_t * max(1, count))()
        for a in range(count):
            illum_c[a] = _native.Illumination_t(float(il[a, 0]), float(il[a, 1]))
        o = _native.BpmRequestOptions_t(
            int(n), int(nzv), int(count),
            ctypes.cast(illum_c, ctypes.POINTER(_native.Illumination_t)), geometry._c(),
            _native.BpmOptions_t(int(bool(obliquity)), int(chunk), int(checkpoint)),
            int(start_iterations),
            _native.BpmSolveOptions_t(int(iterations), int(batch), float(step), float(tv_weight),
                                      int(tv_inner)))
        gid = ctypes.c_int()
        _native.check(_native.lib().qh_hip_server_add_bpm(self._handle, ctypes.byref(o),
                                                          ctypes.byref(gid)))
        self._geometries[gid.value] = ((count, n, n), (nzv, n, n), int(iterations))
        return gid.value

    def submit(self, geometry: int, holograms) -> int:
        if geometry not in self._geometries:
            raise ValueError("Server::submit: unknown geometry")
        shape_in, shape_out, bpm_iterations = self._geometries[geometry]
        if bpm_iterations is not None:
            raise ValueError("Server::submit: a BPM geometry takes submit_bpm")
        h = np.ascontiguousarray(holograms, dtype=np.float32)
        if h.shape != shape_in:
            raise ValueError(f"Server::submit: holograms must be {shape_in}")
        out = np.empty(shape_out, np.float32)
        rid = ctypes.c_int64()
        _native.check(_native.lib().qh_hip_server_submit(self._handle, int(geometry),
                                                         _native.ptr(h), _native.ptr(out),
                                                         ctypes.byref(rid)))
        self._pending[rid.value] = (h, out, None)
        return rid.value

    def submit_bpm(self, geometry: int, fields) -> int:
        """Queues field ratios (count, n, n) of a BPM geometry; returns the request id."""
        if geometry not in self._geometries:
            raise ValueError("Server::submit: unknown geometry")
        shape_in, shape_out, iterations = self._geometries[geometry]
        if iterations is None:
            raise ValueError("Server::submit_bpm: not a BPM geometry")
        f = np.ascontiguousarray(fields, dtype=np.complex64)
        if f.shape != shape_in:
            raise ValueError(f"Server::submit_bpm: fields must be {shape_in}")
        out = np.empty(shape_out, np.float32)
        loss = np.zeros(max(iterations, 1), np.float64)
        rid = ctypes.c_int64()
        _native.check(_native.lib().qh_hip_server_submit_bpm(
            self._handle, int(geometry), _native.ptr(f), _native.ptr(out), _native.ptr(loss),
            ctypes.byref(rid)))
        self._pending[rid.value] = (f, out, loss[:iterations])
        return rid.value

    def poll(self, request: int) -> bool:
        done = ctypes.c_int()
        _native.check(_native.lib().qh_hip_server_poll(self._handle, int(request), ctypes.byref(done)))
        return bool(done.value)

    def wait(self, request: int):
        filled = ctypes.c_double()
        latency = ctypes.c_double()
        rc = _native.lib().qh_hip_server_wait(self._handle, int(request), ctypes.byref(filled),
                                              ctypes.byref(latency))
        _, out, loss = self._pending.pop(int(request), (None, None, None))
        _native.check(rc)
        return out, (filled.value if loss is None else loss), latency.value

    def stats(self) -> dict:
        s = _native.ServerStats_t()
        _native.check(_native.lib().qh_hip_server_stats(self._handle, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self) -> None:
        _native.check(_native.lib().qh_hip_server_reset_stats(self._handle))
