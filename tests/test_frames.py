"""Frame output: the oracle's frame fields (downsample_rho_p,
dgsem_lobatto_2d.cu:826-844) against a brute-force numpy restatement, and the
PPM / plot_rt binary writers' layout (dgsem_lobatto_2d.cu:846-908,
plot_rt.py:10-12)."""
import numpy as np
import pytest

from fvstuff_amd import frames
from fvstuff_amd.capi import make_config
from oracle.oracle import Oracle


def brute_frame_fields(Q, K, N, g=1.4):
    NEL, N2 = K * K, N * N
    U = Q.reshape(4, NEL, N2)
    rho = np.empty(NEL); p = np.empty(NEL)
    for el in range(NEL):
        rs = ps = 0.0
        for n in range(N2):
            r = U[0, el, n] if U[0, el, n] >= 1e-14 else 1e-14
            pr = (g - 1.0) * (U[3, el, n] - 0.5 * (U[1, el, n] ** 2 + U[2, el, n] ** 2) / r)
            rs += r
            ps += pr if pr > 0.0 else 0.0
        rho[el] = rs / N2; p[el] = ps / N2
    return rho, p


def test_oracle_frame_fields_match_definition():
    K, P = 5, 3
    N = P + 1
    o = Oracle(make_config(K=K, p=P))
    rng = np.random.default_rng(7)
    n = K * K * N * N
    Q = np.concatenate([rng.uniform(0.5, 2.0, n), rng.normal(0, 1, n), rng.normal(0, 1, n), rng.uniform(2, 6, n)])
    Q[3] = 1e-20                     # density below the floor
    Q[3 * n + 5] = 0.0               # energy below the kinetic part: negative pressure
    rho, p = o.frame_fields(Q)
    rb, pb = brute_frame_fields(Q, K, N)
    assert np.array_equal(rho, rb)
    assert np.allclose(p, pb, rtol=1e-15, atol=0)
    assert np.all(p >= 0.0)


def test_image_range():
    lo, hi = frames.image_range(np.array([1.0, 2.0, 3.0]))
    assert lo == pytest.approx(1.0 - 0.1, rel=1e-6) and hi == pytest.approx(3.0 + 0.1, rel=1e-6)
    lo, hi = frames.image_range(np.full(4, 2.5))
    assert hi > lo and hi - lo < 1e-5


def test_colormap_endpoints():
    assert tuple(frames.PLASMA[0]) == (13, 8, 135) and tuple(frames.PLASMA[255]) == (240, 249, 33)
    assert tuple(frames.VIRIDIS[0]) == (68, 1, 84) and tuple(frames.VIRIDIS[255]) == (253, 231, 37)
    assert frames.PLASMA.shape == (256, 3) and frames.VIRIDIS.dtype == np.uint8


def test_ppm_pair_layout(tmp_path):
    K = 3
    # rho grows with the row j, p with the column i
    rho = np.repeat(np.arange(K, dtype=float), K)          # el = j*K + i
    p = np.tile(np.arange(K, dtype=float), K)
    fn = tmp_path / "f.ppm"
    frames.write_ppm_pair(str(fn), rho, (0.0, K - 1.0), p, (0.0, K - 1.0), K)
    w, h, pix = frames.read_ppm(str(fn))
    assert (w, h) == (2 * K, K)
    # top row of the file is the physical top row j = K-1: rho at the top of the map
    assert tuple(pix[0, 0]) == tuple(frames.PLASMA[255])
    assert tuple(pix[K - 1, 0]) == tuple(frames.PLASMA[0])
    # right panel: p at i = 0 (left edge of the panel) and i = K-1
    assert tuple(pix[0, K]) == tuple(frames.VIRIDIS[0])
    assert tuple(pix[0, 2 * K - 1]) == tuple(frames.VIRIDIS[255])
    # index mapping idx = int(255 t): the middle value maps to 127
    assert tuple(pix[1, 0]) == tuple(frames.PLASMA[127])


def test_sensor_ppm_and_nan(tmp_path):
    K = 2
    sen = np.array([0.0, np.nan, 0.5, 2.0])
    fn = tmp_path / "s.ppm"
    frames.write_ppm_sensor(str(fn), sen, K)
    w, h, pix = frames.read_ppm(str(fn))
    assert (w, h) == (K, K)
    # row 0 of the file is j = 1: elements 2 (0.5) and 3 (2.0 -> clipped to 1)
    assert tuple(pix[0, 0]) == tuple(frames.PLASMA[127]) and tuple(pix[0, 1]) == tuple(frames.PLASMA[255])
    # NaN and 0 map to index 0
    assert tuple(pix[1, 0]) == tuple(frames.PLASMA[0]) and tuple(pix[1, 1]) == tuple(frames.PLASMA[0])


def test_rt_bin_round_trip(tmp_path):
    K = 4
    rho = np.arange(K * K, dtype=float) * 0.5
    p = np.arange(K * K, dtype=float)[::-1].copy()
    fn = tmp_path / "rt_0000.bin"
    frames.write_rt_bin(str(fn), rho, p, K)
    assert fn.stat().st_size == 4 + 2 * 4 * K * K
    K2, r2, p2 = frames.read_rt_bin(str(fn))
    assert K2 == K
    assert np.array_equal(r2.reshape(-1), rho.astype(np.float32))
    assert np.array_equal(p2.reshape(-1), p.astype(np.float32))
