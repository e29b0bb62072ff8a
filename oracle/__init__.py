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
