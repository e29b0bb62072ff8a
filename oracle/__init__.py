"""Test infrastructure: the CPU oracle, loaded by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` only, as the checker. The product never imports it.

The oracle is a C++/OpenMP restatement of the reference's CPU backend (``src/cpu/*.cpp``,
``src/common.hpp``, ``src/api.cpp``), templated on the working precision: ``dtype=np.float32``
runs the reference's arithmetic, ``dtype=np.float64`` the FP64 statement of the same algorithm.
The reference itself cannot be built here (its CMake fetches pocketfft and GoogleTest from the
network), so there is no ``oracle/_ref``.

Functions pinned by checks in ``tests/test_oracle.py``: the FFT (brute-force DFT, numpy's
pocketfft), the DCT (scipy's DCT-II, the reference's own pin in ``tests/test_dct.cpp``), unwrap
and Rytov (the reference's known answers in ``tests/test_unwrap.cpp``), field extraction
(``tests/test_field.cpp``), the Ewald bins (closed form re-evaluated in float), reconstruction
with GP and FISTA-TV (``tests/test_odt.cpp`` known answers and invariants).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_qh.so"

_lib = None


def build() -> Path:
    """Compiles the oracle with the committed Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.oqh_last_error.restype = ctypes.c_char_p
    return _lib


def set_threads(threads: int) -> None:
    """OpenMP threads of later oracle calls."""
    lib().oqh_set_threads(int(threads))


def max_threads() -> int:
    return int(lib().oqh_max_threads())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib().oqh_last_error().decode())


def _suffix(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f"
    if dtype == np.float64:
        return "d"
    raise TypeError(f"oracle runs in float32 or float64, not {dtype}")


def _cdt(dtype):
    return np.complex64 if np.dtype(dtype) == np.float32 else np.complex128


def _real(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def fft_c2c(a, axes, forward=True, scale=1.0, dtype=np.float32):
    """Complex FFT over the given axes of a 3D array (pocketfft's c2c convention)."""
    a = np.ascontiguousarray(a, dtype=_cdt(dtype)).copy()
    assert a.ndim == 3
    mask = 0
    for ax in axes:
        mask |= 1 << ax
    s = _suffix(dtype)
    _check(getattr(lib(), f"oqh_fft_c2c_{s}")(_ptr(a), *map(ctypes.c_int, a.shape),
                                              ctypes.c_uint(mask), ctypes.c_int(int(forward)),
                                              _real(dtype)(scale)))
    return a


def rfft3(x, dtype=np.float32):
    x = np.ascontiguousarray(x, dtype=dtype)
    nz, ny, nx = x.shape
    out = np.empty((nz, ny, nx // 2 + 1), _cdt(dtype))
    _check(getattr(lib(), f"oqh_rfft3_{_suffix(dtype)}")(_ptr(x), _ptr(out), nz, ny, nx))
    return out


def irfft3(h, nx, scale=1.0, dtype=np.float32):
    h = np.ascontiguousarray(h, dtype=_cdt(dtype)).copy()
    nz, ny, _ = h.shape
    out = np.empty((nz, ny, nx), dtype)
    _check(getattr(lib(), f"oqh_irfft3_{_suffix(dtype)}")(_ptr(h), _ptr(out), nz, ny, nx,
                                                          _real(dtype)(scale)))
    return out


def dct2(a, axis, inverse=False, dtype=np.float32):
    """Unnormalised DCT-II (or its exact inverse) along axis 2 (x) or 1 (y) of a stack."""
    a = np.ascontiguousarray(a, dtype=dtype).copy()
    nz, ny, nx = a.shape
    _check(getattr(lib(), f"oqh_dct2_{_suffix(dtype)}")(_ptr(a), nz, ny, nx, axis, int(inverse)))
    return a


def find_sideband(hologram, dc_exclusion=0.0):
    """(iy, ix, y, x, magnitude) of the +1 sideband of the first hologram of a stack."""
    h = np.ascontiguousarray(hologram, dtype=np.float32)
    if h.ndim == 3:
        h = h[0]
    out = np.zeros(5, np.float32)
    _check(lib().oqh_find_sideband(_ptr(h), h.shape[-1], ctypes.c_float(dc_exclusion), _ptr(out)))
    return int(out[0]), int(out[1]), float(out[2]), float(out[3]), float(out[4])


def extract_fields(holograms, cy, cx, radius, out_size, dtype=np.float32):
    h = np.ascontiguousarray(holograms, dtype=np.float32)
    count, n, _ = h.shape
    out = np.empty((count, out_size, out_size), _cdt(dtype))
    _check(getattr(lib(), f"oqh_extract_fields_{_suffix(dtype)}")(
        _ptr(h), count, n, int(cy), int(cx), ctypes.c_float(radius), int(out_size), _ptr(out)))
    return out


def field_ratio(sample, background, dtype=np.float32):
    s = np.ascontiguousarray(sample, dtype=_cdt(dtype))
    b = np.ascontiguousarray(background, dtype=_cdt(dtype))
    out = np.empty_like(s)
    plane = s.shape[1] * s.shape[2]
    _check(getattr(lib(), f"oqh_field_ratio_{_suffix(dtype)}")(_ptr(s), s.shape[0], _ptr(b),
                                                               b.shape[0], plane, _ptr(out)))
    return out


def unwrap(psi, dtype=np.float32):
    p = np.ascontiguousarray(psi, dtype=dtype)
    out = np.empty_like(p)
    _check(getattr(lib(), f"oqh_unwrap_{_suffix(dtype)}")(_ptr(p), *p.shape, _ptr(out)))
    return out


def rytov(ratio, dtype=np.float32):
    r = np.ascontiguousarray(ratio, dtype=_cdt(dtype))
    out = np.empty_like(r)
    _check(getattr(lib(), f"oqh_rytov_{_suffix(dtype)}")(_ptr(r), *r.shape, _ptr(out)))
    return out


def reconstruct(fields, illuminations, geometry, nz=0, iterations=0, solver="gp",
                gp_transform="r2c", tv_weight=1e-2, tv_inner=10, dtype=np.float32,
                intermediates=False):
    """Rytov fields -> index volume. ``geometry`` is (wavelength, n_medium, na, pixel).

    Returns (ri, filled), or (ri, filled, kspace, weights) with ``intermediates``: the
    normalised k-space potential and the per-voxel hit counts after the Ewald mapping.
    """
    f = np.ascontiguousarray(fields, dtype=_cdt(dtype))
    count, n, _ = f.shape
    il = np.ascontiguousarray(np.reshape(illuminations, (-1, 2)), dtype=np.float32)
    geom = np.asarray(geometry, dtype=np.float32)
    nzv = nz if nz > 0 else n
    opts = np.array([nz, iterations, {"gp": 0, "tv": 1}[solver], {"r2c": 0, "c2c": 1}[gp_transform],
                     tv_inner], np.int32)
    ri = np.empty((nzv, n, n), dtype)
    ks = np.empty((nzv, n, n), _cdt(dtype)) if intermediates else None
    w = np.empty((nzv, n, n), np.float32) if intermediates else None
    filled = ctypes.c_double()
    _check(getattr(lib(), f"oqh_reconstruct_{_suffix(dtype)}")(
        _ptr(f), count, n, _ptr(il), _ptr(geom), _ptr(opts), ctypes.c_float(tv_weight), _ptr(ri),
        _ptr(ks) if ks is not None else None, _ptr(w) if w is not None else None,
        ctypes.byref(filled)))
    if intermediates:
        return ri, filled.value, ks, w
    return ri, filled.value


def tv_prox(b, lam, inner, dtype=np.float32):
    """One FGP prox call from a zero dual: argmin 0.5|x-b|^2 + lam TV(x), x >= 0."""
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.empty_like(b)
    _check(getattr(lib(), f"oqh_tv_prox_{_suffix(dtype)}")(_ptr(b), _ptr(x), *b.shape,
                                                           _real(dtype)(lam), int(inner)))
    return x


def circular_scan(count, na_illumination, wavelength):
    out = np.zeros((count, 2), np.float32)
    _check(lib().oqh_circular_scan(count, ctypes.c_float(na_illumination),
                                   ctypes.c_float(wavelength), _ptr(out)))
    return out


def ewald_targets(n, nz, geometry, illuminations):
    """(kz, nu_z) per (illumination, column) in float; kz = -1 where the pair is rejected."""
    il = np.ascontiguousarray(np.reshape(illuminations, (-1, 2)), dtype=np.float32)
    geom = np.asarray(geometry, dtype=np.float32)
    count = il.shape[0]
    kz = np.empty((count, n, n), np.int32)
    nuz = np.empty((count, n, n), np.float32)
    _check(lib().oqh_ewald_targets(n, nz, _ptr(geom), _ptr(il), count, _ptr(kz), _ptr(nuz)))
    return kz, nuz


def pipeline(holograms, backgrounds, cy, cx, radius, out_size, illuminations, geometry,
             dtype=np.float32, **odt):
    """The per-call chain the reference's CPU Pipeline runs (src/pipeline.cpp:34-53)."""
    ratio = field_ratio(extract_fields(holograms, cy, cx, radius, out_size, dtype),
                        extract_fields(backgrounds, cy, cx, radius, out_size, dtype), dtype)
    return reconstruct(rytov(ratio, dtype), illuminations, geometry, dtype=dtype, **odt)


# ---- beam propagation model (oracle/bpm.cpp) ----

def _bpm_check(rc: int) -> None:
    if rc != 0:
        lib().oqh_bpm_last_error.restype = ctypes.c_char_p
        raise RuntimeError(lib().oqh_bpm_last_error().decode())


def _bpm_common(illuminations, geometry):
    il = np.ascontiguousarray(np.reshape(illuminations, (-1, 2)), dtype=np.float32)
    return il, np.asarray(geometry, dtype=np.float32)


def bpm_tables(n, nz, illuminations, geometry, obliquity=True):
    """(step, last, q, phase, lipschitz) as the reference stores them (float)."""
    il, geom = _bpm_common(illuminations, geometry)
    count = il.shape[0]
    step = np.empty((count, n, n), np.complex64)
    last = np.empty((count, n, n), np.complex64)
    q = np.empty((count, n, n), np.float32)
    sc = np.empty(2, np.float32)
    _bpm_check(lib().oqh_bpm_tables_f(n, nz, _ptr(il), count, _ptr(geom), int(obliquity),
                                      _ptr(step), _ptr(last), _ptr(q), _ptr(sc)))
    return step, last, q, float(sc[0]), float(sc[1])


def bpm_forward(x, illuminations, geometry, obliquity=True, dtype=np.float32):
    """Contrast x = ri - n_medium (nz, n, n) -> field ratios (count, n, n) at z = 0."""
    x = np.ascontiguousarray(x, dtype=dtype)
    il, geom = _bpm_common(illuminations, geometry)
    nz, n, _ = x.shape
    out = np.empty((il.shape[0], n, n), _cdt(dtype))
    _bpm_check(getattr(lib(), f"oqh_bpm_forward_{_suffix(dtype)}")(
        _ptr(x), nz, n, _ptr(il), il.shape[0], _ptr(geom), int(obliquity), _ptr(out)))
    return out


def bpm_jvp(x, h, illuminations, geometry, obliquity=True, dtype=np.float32):
    """Tangent-linear model J h at x."""
    x = np.ascontiguousarray(x, dtype=dtype)
    h = np.ascontiguousarray(h, dtype=dtype)
    il, geom = _bpm_common(illuminations, geometry)
    nz, n, _ = x.shape
    out = np.empty((il.shape[0], n, n), _cdt(dtype))
    _bpm_check(getattr(lib(), f"oqh_bpm_jvp_{_suffix(dtype)}")(
        _ptr(x), _ptr(h), nz, n, _ptr(il), il.shape[0], _ptr(geom), int(obliquity), _ptr(out)))
    return out


def bpm_vjp(x, r, illuminations, geometry, obliquity=True, dtype=np.float32):
    """Re(J^H r) at x, summed over illuminations in order."""
    x = np.ascontiguousarray(x, dtype=dtype)
    r = np.ascontiguousarray(r, dtype=_cdt(dtype))
    il, geom = _bpm_common(illuminations, geometry)
    nz, n, _ = x.shape
    grad = np.empty_like(x)
    _bpm_check(getattr(lib(), f"oqh_bpm_vjp_{_suffix(dtype)}")(
        _ptr(x), _ptr(r), nz, n, _ptr(il), il.shape[0], _ptr(geom), int(obliquity), _ptr(grad)))
    return grad


def bpm_gradient(x, fields, illuminations, geometry, first=0, last=None, obliquity=True,
                 dtype=np.float32):
    """(gradient of 0.5/(last-first) sum |bpm(x) - fields|^2, data term with 0.5/count)."""
    x = np.ascontiguousarray(x, dtype=dtype)
    f = np.ascontiguousarray(fields, dtype=_cdt(dtype))
    il, geom = _bpm_common(illuminations, geometry)
    nz, n, _ = x.shape
    count = il.shape[0]
    last = count if last is None else last
    grad = np.empty_like(x)
    loss = ctypes.c_double()
    _bpm_check(getattr(lib(), f"oqh_bpm_gradient_{_suffix(dtype)}")(
        _ptr(x), _ptr(f), nz, n, _ptr(il), count, _ptr(geom), int(obliquity), first, last,
        _ptr(grad), ctypes.byref(loss)))
    return grad, loss.value


def bpm_reconstruct(fields, illuminations, geometry, initial, iterations=20, batch=0, step=1.0,
                    tv_weight=0.0, tv_inner=10, obliquity=True, dtype=np.float32):
    """Field ratios -> index volume by the reference's BPM inversion; returns (ri, loss).

    ``initial`` is a refractive-index volume; the contrast is formed in float32 as the
    reference's api.cpp:159-163 does, and n_medium is added back at the end (api.cpp:199)."""
    f = np.ascontiguousarray(fields, dtype=_cdt(dtype))
    il, geom = _bpm_common(illuminations, geometry)
    count, n, _ = f.shape
    nm = np.float32(geom[1])
    x = (np.asarray(initial, np.float32) - nm).astype(dtype)
    x = np.ascontiguousarray(x)
    nz = x.shape[0]
    opts = np.array([iterations, batch, tv_inner, int(obliquity)], np.int32)
    loss = np.zeros(iterations, np.float64)
    _bpm_check(getattr(lib(), f"oqh_bpm_reconstruct_{_suffix(dtype)}")(
        _ptr(f), count, n, _ptr(il), _ptr(geom), _ptr(x), nz, _ptr(opts), ctypes.c_float(step),
        ctypes.c_float(tv_weight), _ptr(loss)))
    ri = (x.astype(np.float32) + nm) if np.dtype(dtype) == np.float32 else x + float(nm)
    return ri, loss
