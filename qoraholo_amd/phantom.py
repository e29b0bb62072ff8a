"""Synthetic inputs: phantoms, analytic forward fields and off-axis holograms (host, numpy).

Follows the reference's generator (src/phantom.cpp:18-39,84-221): painter's-order shapes
sampled at voxel centres; the first-Born spectrum of each shape placed on every Ewald cap and
inverse transformed (``born``: 1 + u, ``rytov``: exp(u)); holograms |R + u_in ratio|^2 plus
Gaussian noise drawn from std::mt19937_64 through std::normal_distribution<float>. Ellipsoids
(axis-aligned) extend the sphere model: their transform is the ball transform at the
anisotropically scaled frequency, with the same painter's-order parent rule.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _native


@dataclass
class Sphere:
    """Uniform sphere, centre in micrometres from the volume centre (phantom.hpp:15-21)."""

    z: float = 0.0
    y: float = 0.0
    x: float = 0.0
    radius: float = 1.0
    n: float = 1.38

    def semi_axes(self):
        r = float(np.float32(self.radius))
        return r, r, r


@dataclass
class Ellipsoid:
    """Uniform axis-aligned ellipsoid with semi-axes (az, ay, ax) in micrometres."""

    z: float = 0.0
    y: float = 0.0
    x: float = 0.0
    az: float = 1.0
    ay: float = 1.0
    ax: float = 1.0
    n: float = 1.38

    def semi_axes(self):
        return (float(np.float32(self.az)), float(np.float32(self.ay)), float(np.float32(self.ax)))


def _f32(v) -> float:
    return float(np.float32(v))


def _inside(s, z, y, x):
    """Shape membership at points (arrays, micrometres), in double (phantom.cpp:18-23)."""
    dz = z - _f32(s.z)
    dy = y - _f32(s.y)
    dx = x - _f32(s.x)
    if isinstance(s, Sphere):
        r = _f32(s.radius)
        return dz * dz + dy * dy + dx * dx <= r * r
    az, ay, ax = s.semi_axes()
    return (dz / az) ** 2 + (dy / ay) ** 2 + (dx / ax) ** 2 <= 1.0


def _index_at(shapes, count, n_medium, z, y, x):
    n = np.full(np.broadcast(z, y, x).shape, float(n_medium))
    for s in shapes[:count]:
        n = np.where(_inside(s, z, y, x), _f32(s.n), n)
    return n


def render_phantom(shapes, geometry, nz: int, n: int) -> np.ndarray:
    """Refractive index at voxel centres, (nz, n, n) float32, centre-origin (phantom.cpp:84-100)."""
    if nz < 1 or n < 1:
        raise ValueError("render_phantom: empty shape")
    d = _f32(geometry.pixel)
    zc = ((np.arange(nz) - nz // 2) * d)[:, None, None]
    yc = ((np.arange(n) - n // 2) * d)[None, :, None]
    xc = ((np.arange(n) - n // 2) * d)[None, None, :]
    out = np.empty((nz, n, n), np.float32)
    shapes = list(shapes)
    for k in range(nz):
        out[k] = _index_at(shapes, len(shapes), _f32(geometry.n_medium), zc[k], yc[0], xc[0, 0])
    return out


def _ball(x):
    x = np.asarray(x, dtype=np.float64)
    small = x < 1e-3
    xs = np.where(small, 1.0, x)
    return np.where(small, 1.0 - x * x / 10.0, 3.0 * (np.sin(xs) - xs * np.cos(xs)) / (xs * xs * xs))


def partial_waves(x: float, m: float) -> np.ndarray:
    """Coefficients c_l of the scattered wave of a homogeneous sphere for the scalar Helmholtz
    equation (phantom.cpp:41-80): outside u = sum i^l (2l+1) (j_l(kr) + c_l h_l(kr)) P_l,
    inside b_l j_l(mkr), u and du/dr continuous at r = R; x = k R, m = n_sphere / n_medium.

    j_l / j_{l-1} comes from the downward recurrence (started 60 orders above lmax, where the
    upward one is unstable), y_l from the upward one; c_l = (m D j - j') / (h' - m D h) with
    D = j_l'(mx) / j_l(mx)."""
    lmax = int(x + 4.0 * np.cbrt(x) + 10.0)

    def ratios(z):
        rho = np.zeros(lmax + 1)
        r = 0.0
        for l in range(lmax + 60, 0, -1):
            r = 1.0 / ((2.0 * l + 1.0) / z - r)
            if l <= lmax:
                rho[l] = r
        return rho

    rx = ratios(x)
    rm = ratios(m * x)
    c = np.zeros(lmax + 1, np.complex128)
    j_prev, j = math.cos(x) / x, math.sin(x) / x
    y_prev, y = math.sin(x) / x, -math.cos(x) / x
    for l in range(lmax + 1):
        if l > 0:
            y_prev, y = y, (2.0 * l - 1.0) / x * y - y_prev
            j_prev, j = j, rx[l] * j
        h = complex(j, y)
        h_prev = complex(j_prev, y_prev)
        dj = j_prev - (l + 1.0) / x * j
        dh = h_prev - (l + 1.0) / x * h
        dlog = (1.0 / math.tan(m * x) if l == 0 else 1.0 / rm[l]) - (l + 1.0) / (m * x)
        c[l] = (m * dlog * j - dj) / (dh - m * dlog * h)
    return c


def forward_fields(shapes, illuminations, geometry, n: int, model: str = "rytov") -> np.ndarray:
    """Field ratios u / u_in at z = 0, (count, n, n) complex64 (phantom.cpp:102-191).

    born: first Born; rytov: the Born phase as a Rytov phase; exact: one sphere solved by partial
    waves (multiple scattering included), the forward-scattered plane waves at z = 0 through the
    scattering amplitude f(theta) = -i / k sum (2l+1) c_l P_l(cos theta)."""
    if n < 4 or n % 2:
        raise ValueError("forward_fields: n must be even")
    if model not in ("born", "rytov", "exact"):
        raise ValueError("forward_fields: model must be 'born', 'rytov' or 'exact'")
    shapes = list(shapes)
    if model == "exact" and (len(shapes) != 1 or not isinstance(shapes[0], Sphere)):
        raise ValueError("forward_fields: the exact model takes exactly one sphere")
    wl = _f32(geometry.wavelength)
    nm = _f32(geometry.n_medium)
    k0 = 2.0 * math.pi / wl
    fm2 = nm * nm / (wl * wl)
    fna2 = _f32(geometry.na) ** 2 / (wl * wl)
    d = _f32(geometry.pixel)
    df = 1.0 / (n * d)
    km = k0 * nm
    waves = (partial_waves(km * _f32(shapes[0].radius), _f32(shapes[0].n) / nm)
             if model == "exact" else None)
    steps = []
    for i, s in enumerate(shapes):
        parent = float(_index_at(shapes, i, nm, _f32(s.z), _f32(s.y), _f32(s.x)))
        az, ay, ax = s.semi_axes()
        volume = 4.0 / 3.0 * math.pi * az * ay * ax
        steps.append(k0 * k0 * (_f32(s.n) ** 2 - parent * parent) * volume)
    il = np.reshape(np.asarray(illuminations, np.float32), (-1, 2))
    count = il.shape[0]
    k = np.arange(n)
    sb = np.where(k < n // 2, k, k - n).astype(np.float64) * df
    my = sb[:, None] * np.ones((1, n))
    mx = np.ones((n, 1)) * sb[None, :]
    par = np.where((k[:, None] + k[None, :]) % 2 == 0, 1.0, -1.0)
    out = np.zeros((count, n, n), np.complex64)
    for a in range(count):
        fiy, fix = float(il[a, 0]), float(il[a, 1])
        if fiy * fiy + fix * fix >= fm2:
            raise ValueError("forward_fields: illumination beyond n_medium / wavelength")
        fiz = math.sqrt(fm2 - fiy * fiy - fix * fix)
        fy = my + fiy
        fx = mx + fix
        rho2 = fy * fy + fx * fx
        mask = (rho2 <= fna2) & (rho2 < fm2)
        myv, mxv = my[mask], mx[mask]
        nuz = np.sqrt(fm2 - rho2[mask])
        qz = nuz - fiz
        o = np.zeros(nuz.shape, np.complex128)
        if model == "exact":
            sp = shapes[0]
            cos_t = (fy[mask] * fiy + fx[mask] * fix + nuz * fiz) / fm2
            total = np.zeros(nuz.shape, np.complex128)
            p_prev = np.zeros(nuz.shape)
            p = np.ones(nuz.shape)
            for l, cl in enumerate(waves):
                w = 2.0 * l + 1.0
                total = total + w * cl * p
                p_prev, p = p, (w * cos_t * p - l * p_prev) / (l + 1.0)
            ph = -2.0 * math.pi * (qz * _f32(sp.z) + myv * _f32(sp.y) + mxv * _f32(sp.x))
            o = 4.0 * math.pi * (-1j / km) * total * (np.cos(ph) + 1j * np.sin(ph))
        for s, st in zip(shapes if model != "exact" else [], steps):
            ph = -2.0 * math.pi * (qz * _f32(s.z) + myv * _f32(s.y) + mxv * _f32(s.x))
            if isinstance(s, Sphere):
                q = 2.0 * math.pi * np.sqrt(qz * qz + myv * myv + mxv * mxv)
                b = _ball(q * _f32(s.radius))
            else:
                az, ay, ax = s.semi_axes()
                b = _ball(2.0 * math.pi * np.sqrt((qz * az) ** 2 + (myv * ay) ** 2 + (mxv * ax) ** 2))
            o = o + st * b * (np.cos(ph) + 1j * np.sin(ph))
        u = (1j / (4.0 * math.pi * nuz)) * o * par[mask] / (d * d)
        plane = np.zeros((n, n), np.complex64)
        plane[mask] = u.astype(np.complex64)
        out[a] = plane
    out = np.fft.ifft2(out, axes=(1, 2)).astype(np.complex64)
    if model == "rytov":
        return np.exp(out).astype(np.complex64)
    return (np.complex64(1.0) + out).astype(np.complex64)


def gaussian_noise(seed: int, sigma: float, count: int) -> np.ndarray:
    """The noise stream of the reference's synthesiser: std::mt19937_64(seed) drawn through
    std::normal_distribution<float>(0, sigma), in order."""
    out = np.empty(count, np.float32)
    _native.check(_native.lib().qh_gaussian_noise(int(seed), float(sigma), int(count),
                                                  _native.ptr(out)))
    return out


def synthesize_holograms(ratio, illuminations, geometry, carrier=(5.0, 5.0),
                         reference_amplitude: float = 1.0, noise: float = 0.0,
                         seed: int = 1) -> np.ndarray:
    """|R + u_in ratio|^2 + noise, R = a exp(-2 pi i carrier . r) (phantom.cpp:193-221)."""
    r = np.asarray(ratio, np.complex64)
    single = r.ndim == 2
    if single:
        r = r[None]
    il = np.reshape(np.asarray(illuminations, np.float32), (-1, 2))
    if il.shape[0] != r.shape[0]:
        raise ValueError("synthesize_holograms: need one illumination per field")
    count, ny, nx = r.shape
    d = _f32(geometry.pixel)
    ry = ((np.arange(ny) - ny // 2) * d)[:, None]
    rx = ((np.arange(nx) - nx // 2) * d)[None, :]
    cy, cx = _f32(carrier[0]), _f32(carrier[1])
    amp = _f32(reference_amplitude)
    ref_ph = -2.0 * math.pi * (cy * ry + cx * rx)
    ref = amp * (np.cos(ref_ph) + 1j * np.sin(ref_ph))
    sigma = _f32(noise)
    out = np.empty((count, ny, nx), np.float32)
    stream = gaussian_noise(seed, sigma, count * ny * nx).reshape(count, ny, nx) if sigma > 0 else None
    for a in range(count):
        ph = 2.0 * math.pi * (float(il[a, 0]) * ry + float(il[a, 1]) * rx)
        ui = np.cos(ph) + 1j * np.sin(ph)
        u = ui * r[a].astype(np.complex128)
        t = ref + u
        v = (t.real * t.real + t.imag * t.imag).astype(np.float32)
        if stream is not None:
            v = v + stream[a]
        out[a] = v
    return out[0] if single else out
