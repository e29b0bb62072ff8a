"""Synthetic workloads of the benchmark configurations, built with the package's synthesiser.

Geometry throughout: wavelength 0.532 um, medium 1.337, detection NA 1.2, illumination by
``circular_scan`` at NA 0.9 (index 0 is normal incidence). Holograms have 0.055 um pixels and
are cropped to half their side, so fields and voxels are 0.11 um (the reference README's
geometry, ``bench/bench.cpp:102-113``); the carrier is 5 cycles/um on both axes.

* ``readme_cell``: 49 x 512^2 holograms -> 256^3, the synthetic cell of an ellipsoidal body
  (semi-axes x 8, y 7, z 5 um, n 1.360), a nucleus (r 3 um, n 1.375) and two vesicles (r 1 um,
  n 1.390) placed inside the body by seeded rejection sampling; angles estimated from the
  background holograms with ``find_sideband``.
* ``bead``: 49 x 256^2 holograms -> 128^3, one sphere r 3 um n 1.360 at the origin, carrier
  (70.5, 70.5) px, crop radius 32. With 0.055 um hologram pixels the carrier of 5 cycles/um
  sits at 70.4 px and the NA-1.2 pupil at 31.8 bins, which is what a 70.5 px carrier and a
  radius of 32 describe; a literal 0.2 um field pixel would put 48 of the 49 illuminations
  outside the crop.

Noise is additive Gaussian with sigma = 1% of the mean noiseless intensity, seed 0 for the
sample holograms and seed 1 for the backgrounds.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import Geometry, circular_scan, find_sideband, hip_available
from . import estimate_illuminations as device_estimate
from .phantom import Ellipsoid, Sphere, forward_fields, synthesize_holograms

WAVELENGTH = 0.532
N_MEDIUM = 1.337
NA = 1.2
NA_ILL = 0.9
HOLO_PIXEL = 0.055
CARRIER = (5.0, 5.0)


@dataclass
class Workload:
    name: str
    holograms: np.ndarray
    backgrounds: np.ndarray
    center: tuple
    radius: float
    field_size: int
    geometry: Geometry
    holo_geometry: Geometry
    illuminations: np.ndarray
    true_illuminations: np.ndarray
    shapes: list = field(default_factory=list)


def cell_shapes(seed: int = 0) -> list:
    """Ellipsoidal body with a nucleus and two vesicles, each inside the body and disjoint."""
    rng = np.random.default_rng(seed)
    body = Ellipsoid(0.0, 0.0, 0.0, az=5.0, ay=7.0, ax=8.0, n=1.360)
    placed = []

    def fits(c, r):
        # Inside the body with a margin: the ellipsoid shrunk by r contains the centre.
        inside = (c[0] / (body.az - r)) ** 2 + (c[1] / (body.ay - r)) ** 2 + \
                 (c[2] / (body.ax - r)) ** 2 <= 1.0
        disjoint = all(math.dist(c, p[0]) > r + p[1] + 0.2 for p in placed)
        return inside and disjoint

    out = [body]
    for r, n in ((3.0, 1.375), (1.0, 1.390), (1.0, 1.390)):
        while True:
            c = tuple(float(v) for v in rng.uniform(-1.0, 1.0, 3) * (body.az, body.ay, body.ax))
            if fits(c, r):
                break
        placed.append((c, r))
        out.append(Sphere(z=c[0], y=c[1], x=c[2], radius=r, n=n))
    return out


def _noisy(ratio, illum, holo_geom, seed):
    clean = synthesize_holograms(ratio, illum, holo_geom, carrier=CARRIER)
    sigma = 0.01 * float(np.mean(clean, dtype=np.float64))
    return synthesize_holograms(ratio, illum, holo_geom, carrier=CARRIER, noise=sigma, seed=seed)


def estimate_illuminations(backgrounds, holo_n: int, holo_pixel: float) -> tuple:
    """Carrier from the normal-incidence background and each illumination from its background
    peak, (s - c) * df (the reference's examples/readme_figure.py:39-46)."""
    df = 1.0 / (holo_n * holo_pixel)
    c = find_sideband(backgrounds[0])
    est = np.array([[(s.y - c.y) * df, (s.x - c.x) * df]
                    for s in (find_sideband(b) for b in backgrounds)], np.float32)
    return c, est


def readme_cell(n: int = 256, angles: int = 49, seed: int = 0) -> Workload:
    holo_geom = Geometry(WAVELENGTH, N_MEDIUM, NA, HOLO_PIXEL)
    geom = Geometry(WAVELENGTH, N_MEDIUM, NA, 2 * HOLO_PIXEL)
    illum = circular_scan(angles, NA_ILL, geom)
    shapes = cell_shapes(seed)
    ratio = forward_fields(shapes, illum, holo_geom, 2 * n)
    holo = _noisy(ratio, illum, holo_geom, seed)
    bg = _noisy(np.ones_like(ratio), illum, holo_geom, seed + 1)
    c, est = estimate_illuminations(bg, 2 * n, HOLO_PIXEL)
    df = 1.0 / (2 * n * HOLO_PIXEL)
    radius = float(math.ceil(NA / WAVELENGTH / df) + 1)
    return Workload("readme_cell", holo, bg, (float(c.iy), float(c.ix)), radius, n, geom,
                    holo_geom, est, illum, shapes)


def bead(n: int = 128, angles: int = 49, seed: int = 0) -> Workload:
    holo_geom = Geometry(WAVELENGTH, N_MEDIUM, NA, HOLO_PIXEL)
    geom = Geometry(WAVELENGTH, N_MEDIUM, NA, 2 * HOLO_PIXEL)
    illum = circular_scan(angles, NA_ILL, geom)
    shapes = [Sphere(0.0, 0.0, 0.0, 3.0, 1.360)]
    ratio = forward_fields(shapes, illum, holo_geom, 2 * n)
    holo = _noisy(ratio, illum, holo_geom, seed)
    bg = _noisy(np.ones_like(ratio), illum, holo_geom, seed + 1)
    return Workload("bead", holo, bg, (70.5, 70.5), 32.0, n, geom, holo_geom, illum, illum,
                    shapes)


def random_cell(n: int, seed: int, pixel: float = 2 * HOLO_PIXEL) -> list:
    """3-8 spheres with n ~ U(1.345, 1.395), each either inside an earlier one or disjoint from
    all of them (the painter's-order rule of the synthesiser), by seeded rejection sampling in
    a field of view of n voxels of ``pixel`` um (0.11 um by default)."""
    rng = np.random.default_rng(seed)
    half = 0.4 * n * pixel
    count = int(rng.integers(3, 9))
    placed = []
    while len(placed) < count:
        r = float(rng.uniform(0.08, 0.25) * half)
        c = tuple(float(v) for v in rng.uniform(-half + r, half - r, 3))
        ok = True
        for (pc, pr, _) in placed:
            d = math.dist(c, pc)
            if not (d > r + pr + 0.2 or d + r < pr - 0.2):
                ok = False
                break
        if ok:
            placed.append((c, r, float(rng.uniform(1.345, 1.395))))
    # Larger spheres first so that nested ones are painted over their parents.
    placed.sort(key=lambda p: -p[1])
    return [Sphere(z=c[0], y=c[1], x=c[2], radius=r, n=nn) for c, r, nn in placed]


def mixed_cell(n: int, angles: int, seed: int = 0) -> Workload:
    """A random cell of ``random_cell`` at side n (holograms 2n at 0.055 um), angles from the
    backgrounds, crop radius ceil((NA / lambda) / df) + 1."""
    holo_geom = Geometry(WAVELENGTH, N_MEDIUM, NA, HOLO_PIXEL)
    geom = Geometry(WAVELENGTH, N_MEDIUM, NA, 2 * HOLO_PIXEL)
    illum = circular_scan(angles, NA_ILL, geom)
    shapes = random_cell(n, seed)
    ratio = forward_fields(shapes, illum, holo_geom, 2 * n)
    holo = _noisy(ratio, illum, holo_geom, seed)
    bg = _noisy(np.ones_like(ratio), illum, holo_geom, seed + 1)
    if hip_available():
        c, est = device_estimate(bg, holo_geom)
    else:
        c, est = estimate_illuminations(bg, 2 * n, HOLO_PIXEL)
    df = 1.0 / (2 * n * HOLO_PIXEL)
    radius = float(math.ceil(NA / WAVELENGTH / df) + 1)
    return Workload(f"mixed_cell_{n}", holo, bg, (float(c.iy), float(c.ix)), radius, n, geom,
                    holo_geom, est, illum, shapes)


BPM_PIXEL = 0.2


@dataclass
class BpmRequest:
    name: str
    fields: np.ndarray
    geometry: Geometry
    illuminations: np.ndarray
    truth: np.ndarray
    shapes: list = field(default_factory=list)


def bpm_request(n: int, angles: int, seed: int = 0, noise: float = 0.01) -> BpmRequest:
    """An inversion request of the served BPM configuration: a ``random_cell`` of side n at
    0.2 um voxels (n^2 x n), ``angles`` illuminations at NA 0.9, field ratios from the device
    BPM forward of the rendered cell plus complex Gaussian noise of ``noise`` per component
    (seeded). The data come from the discretisation that is inverted; the request is a timing
    and parity workload, not a validation of the model."""
    from .bpm import bpm_forward
    from .phantom import render_phantom

    geom = Geometry(WAVELENGTH, N_MEDIUM, NA, BPM_PIXEL)
    illum = circular_scan(angles, NA_ILL, geom)
    shapes = random_cell(n, seed, BPM_PIXEL)
    truth = render_phantom(shapes, geom, n, n)
    fields = bpm_forward(truth, illum, geom)
    rng = np.random.default_rng(seed + 100)
    fields = fields + (noise * (rng.standard_normal(fields.shape, np.float32)
                                + 1j * rng.standard_normal(fields.shape, np.float32)))
    return BpmRequest(f"bpm_cell_{n}_{angles}", fields.astype(np.complex64), geom, illum, truth,
                      shapes)
