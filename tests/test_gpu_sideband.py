"""Device-side sideband search and illumination estimation against the oracle's find_sideband
(src/cpu/field.cpp:35-60) and the reference's own known answer (tests/test_field.cpp:39-50),
and the served loop's cold start from backgrounds alone."""

import numpy as np
import pytest

from test_gpu_parity import PipelineData

pytestmark = pytest.mark.gpu


def refined64(h, iy, ix):
    """Sub-bin position of the peak (iy, ix) from the double-precision spectrum of the
    modulated hologram (src/cpu/field.cpp:24-33 in FP64): the value both float paths round."""
    n = h.shape[-1]
    k = np.arange(n)
    sign = np.where((k[:, None] + k[None, :]) % 2 == 0, 1.0, -1.0)
    s = np.fft.fft2(h.astype(np.float64) * sign)
    c = n // 2
    y, x = iy + c, ix + c

    def refine(lo, mid, hi):
        den = 2.0 * mid - lo - hi
        cc = np.pi / n
        return float(np.clip(((lo - hi) / den).real * np.tan(cc) / cc, -0.5, 0.5))

    return (iy + refine(s[y - 1, x], s[y, x], s[y + 1, x]),
            ix + refine(s[y, x - 1], s[y, x], s[y, x + 1]))


def test_carrier_sideband_known_answer(device, qh, oracle):
    """tests/test_field.cpp:39-50: 128^2 background at 0.055 um, carrier 5 cycles/um: peak
    (35, 35), refined to 35.2 px within 0.05; the same values as the oracle."""
    g = qh.Geometry(0.532, 1.337, 1.2, 0.055)
    bg = qh.synthesize_holograms(np.ones((1, 128, 128), np.complex64), [[0, 0]], g,
                                 carrier=(5.0, 5.0))
    (s,) = qh.find_sidebands(bg)
    assert (s.iy, s.ix) == (35, 35)
    expected = 5.0 * 128 * 0.055
    assert abs(s.y - expected) < 0.05 and abs(s.x - expected) < 0.05
    iy, ix, y, x, mag = oracle.find_sideband(bg)
    assert abs(s.y - y) < 1e-4 and abs(s.x - x) < 1e-4
    assert s.magnitude == pytest.approx(mag, rel=1e-5)


@pytest.mark.parametrize("holo,angles", [(128, 9), (256, 25), (512, 49)])
def test_every_background_matches_oracle(device, qh, oracle, holo, angles):
    """Integer peaks equal the oracle's; the refined positions are as close to the FP64 value as
    the float oracle's are (within 2x, plus 1e-4 bin). Peaks next to the DC disc make the
    three-bin estimate ill-conditioned (a denominator of a few units against peaks of ~700),
    where both float paths move by up to 1e-3 bin."""
    d = PipelineData(qh, holo=holo, angles=angles)
    got = qh.find_sidebands(d.bg)
    for a, s in enumerate(got):
        iy, ix, y, x, mag = oracle.find_sideband(d.bg[a])
        assert (s.iy, s.ix) == (iy, ix), a
        y64, x64 = refined64(d.bg[a], iy, ix)
        assert abs(s.y - y64) <= 2 * abs(y - y64) + 1e-4, (a, s.y, y, y64)
        assert abs(s.x - x64) <= 2 * abs(x - x64) + 1e-4, (a, s.x, x, x64)
        assert s.magnitude == pytest.approx(mag, rel=1e-4)
    # A wider DC exclusion: the same integer peaks as the oracle with that radius.
    dc = float(holo) / 6.0
    wide = qh.find_sidebands(d.bg, dc_exclusion=dc)
    assert [(s.iy, s.ix) for s in wide] == [tuple(oracle.find_sideband(b, dc)[:2]) for b in d.bg]


def test_illuminations_from_backgrounds(device, qh, oracle):
    """examples/readme_figure.py:39-46 on the device: the carrier and every illumination
    frequency as the host estimate gives them, and within a fraction of a bin of the truth."""
    d = PipelineData(qh, holo=256, angles=25)
    carrier, est = qh.estimate_illuminations(d.bg, d.holo_geom)
    df = 1.0 / (256 * 0.1)
    c = oracle.find_sideband(d.bg[0])
    assert (carrier.iy, carrier.ix) == (c[0], c[1])
    pos = [oracle.find_sideband(b) for b in d.bg]
    exact = [refined64(b, p[0], p[1]) for b, p in zip(d.bg, pos)]
    host = np.array([[(p[2] - c[2]) * df, (p[3] - c[3]) * df] for p in pos])
    truth = np.array([[(e[0] - exact[0][0]) * df, (e[1] - exact[0][1]) * df] for e in exact])
    assert np.all(np.abs(est - truth) <= 2 * np.abs(host - truth) + 2e-4 * df)
    # Against the true illuminations where the sideband lies clear of the DC disc (n/8 bins) in
    # the searched half plane x > 0: elsewhere the search finds another peak or the conjugate
    # twin, on the host as on the device (the reference's algorithm).
    py, px = 36.0 + d.illum[:, 0] / df, 36.0 + d.illum[:, 1] / df
    clear = (np.hypot(py, px) > 256 / 8 + 3) & (px > 3)
    assert clear.sum() >= d.count // 2
    assert np.max(np.abs(est - d.illum)[clear]) < 0.5 * df


def test_server_cold_start_from_backgrounds(device, qh):
    """A geometry registered with backgrounds only (carrier and illuminations estimated on the
    device) reconstructs as a Pipeline built from the same estimates."""
    d = PipelineData(qh, holo=128, angles=9)
    carrier, est = qh.estimate_illuminations(d.bg, d.holo_geom)
    kw = dict(d.kw)
    kw.pop("center")
    server = qh.Server(max_slots=2)
    gid = server.add_geometry(d.count, d.size, None, None, d.geom, d.bg, **kw)
    ri, _, _ = server.wait(server.submit(gid, d.holo))
    p = qh.Pipeline(d.count, d.size, center=(float(carrier.iy), float(carrier.ix)),
                    illuminations=est, geometry=d.geom, **kw)
    p.set_background(d.bg)
    assert np.array_equal(ri, p.run(d.holo))
