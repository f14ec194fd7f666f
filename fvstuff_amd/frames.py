"""Frame output of the reference's run (dgsem_lobatto_2d.cu:846-908, 984-1003)
and the binary frame format read by plot_rt.py (plot_rt.py:10-12).

The field data come from the device (`Solver.frame_fields`: per-element mean
rho and p, `Solver.sensor`: the shock sensor); this module only formats them:

- `image_range`: colour range of a field, padded by 5 % (:846-857);
- `write_ppm_pair`: binary PPM (P6) of width 2K and height K, rho on the left
  and p on the right, rows written top (j = K-1) to bottom (:860-888);
- `write_ppm_sensor`: single-panel PPM of the sensor on [0, 1] (:890-908);
- `write_rt_bin` / `read_rt_bin`: int32 K, then K*K float32 rho and K*K
  float32 p, row-major with j = 0 at the bottom (plot_rt.py:10-12);
- `FrameWriter`: the reference's WRITE_FRAME sequence (the colour ranges are
  fixed by the first frame and reused for every later one).

Colour maps: piecewise-linear approximations of matplotlib's `plasma` (rho,
sensor) and `viridis` (p) through nine anchor colours each.  The reference
embeds 256-entry tables of those maps; they are not copied here, so pixel
colours differ slightly while the format, layout and index mapping
(idx = int(255 t), t clipped to [0, 1]) are the reference's.
"""
import os
import struct

import numpy as np

_PLASMA = [(13, 8, 135), (75, 3, 161), (125, 3, 168), (168, 34, 150), (203, 70, 121),
           (229, 107, 93), (248, 148, 65), (253, 195, 40), (240, 249, 33)]
_VIRIDIS = [(68, 1, 84), (71, 44, 122), (59, 81, 139), (44, 113, 142), (33, 144, 141),
            (39, 173, 129), (92, 200, 99), (170, 220, 50), (253, 231, 37)]


def _table(anchors):
    a = np.asarray(anchors, dtype=np.float64)
    x = np.linspace(0.0, 1.0, len(a))
    t = np.arange(256) / 255.0
    return np.stack([np.interp(t, x, a[:, c]) for c in range(3)], axis=1).round().astype(np.uint8)


PLASMA = _table(_PLASMA)
VIRIDIS = _table(_VIRIDIS)


def image_range(img):
    """(vmin, vmax) of a field padded by 5 % of its span; a constant field
    gets +-1e-6 (:846-857)."""
    img = np.asarray(img, dtype=np.float32)
    lo, hi = float(img.min()), float(img.max())
    pad = np.float32(0.05) * np.float32(hi - lo + np.float32(1e-12))
    vmin, vmax = float(np.float32(lo) - pad), float(np.float32(hi) + pad)
    if not vmax > vmin:
        vmin, vmax = lo - 1e-6, hi + 1e-6
    return vmin, vmax


def _indices(img, vmin, vmax):
    """Colour index per pixel: t = (v - vmin)/(vmax - vmin + 1e-30) clipped to
    [0, 1] (NaN -> 0), idx = int(255 t)."""
    v = np.asarray(img, dtype=np.float32)
    t = (v - np.float32(vmin)) / (np.float32(vmax) - np.float32(vmin) + np.float32(1e-30))
    t = np.where(t > 0, t, 0).astype(np.float32)
    t = np.minimum(t, np.float32(1.0))
    return np.clip((t * np.float32(255.0)).astype(np.int32), 0, 255)


def _rows_top_down(field, K):
    """(K, K) image with row r holding physical row j = K-1-r."""
    return np.asarray(field).reshape(K, K)[::-1]


def write_ppm_pair(fname, rho, rho_range, p, p_range, K):
    """Two-panel P6 image, 2K x K: rho (plasma) | p (viridis)."""
    left = PLASMA[_indices(_rows_top_down(rho, K), *rho_range)]
    right = VIRIDIS[_indices(_rows_top_down(p, K), *p_range)]
    pix = np.concatenate([left, right], axis=1)
    with open(fname, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (2 * K, K))
        f.write(np.ascontiguousarray(pix, dtype=np.uint8).tobytes())


def write_ppm_sensor(fname, sensor, K, vmin=0.0, vmax=1.0):
    """Single-panel P6 image of the per-element sensor (plasma on [0, 1])."""
    pix = PLASMA[_indices(_rows_top_down(sensor, K), vmin, vmax)]
    with open(fname, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (K, K))
        f.write(np.ascontiguousarray(pix, dtype=np.uint8).tobytes())


def read_ppm(fname):
    """(width, height, pixels[h, w, 3]) of a P6 file written above."""
    with open(fname, "rb") as f:
        data = f.read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    w, h = (int(v) for v in parts[1].split())
    return w, h, np.frombuffer(parts[3], dtype=np.uint8).reshape(h, w, 3)


def write_rt_bin(fname, rho, p, K):
    """plot_rt.py frame: int32 K, K*K float32 rho, K*K float32 p (j = 0 first)."""
    with open(fname, "wb") as f:
        f.write(struct.pack("<i", int(K)))
        f.write(np.asarray(rho, dtype="<f4").reshape(-1).tobytes())
        f.write(np.asarray(p, dtype="<f4").reshape(-1).tobytes())


def read_rt_bin(fname):
    with open(fname, "rb") as f:
        K = struct.unpack("<i", f.read(4))[0]
        rho = np.frombuffer(f.read(4 * K * K), dtype="<f4").reshape(K, K)
        p = np.frombuffer(f.read(4 * K * K), dtype="<f4").reshape(K, K)
    return K, rho, p


class FrameWriter:
    """The reference's WRITE_FRAME (:984-1003) on top of a Solver: frame
    fields and sensor on the device, copied once per frame, written as
    <prefix>_NNNN.ppm (rho | p) and <sensor_prefix>_NNNN.ppm, plus the
    plot_rt binary <prefix>_NNNN.bin when `binary` is set.  The
    spectral-difference driver's WRITE_FRAME (fr_lobatto_2d.cu:904-919)
    writes the rho | p pair alone: prefix="fr_sod", sensor_prefix=None."""

    def __init__(self, solver, outdir=".", prefix="dgsem_sod", sensor_prefix="dgsem_sensor", binary=False):
        self.s = solver
        self.K = solver.cfg.K
        self.outdir = outdir
        self.prefix = prefix
        self.sensor_prefix = sensor_prefix
        self.binary = binary
        self.rho_range = None
        self.p_range = None

    def write(self, Q, index):
        import torch
        rho_d, p_d = self.s.frame_fields(Q)
        sen_d = self.s.sensor(Q) if self.sensor_prefix else None
        torch.cuda.synchronize()
        rho = rho_d.cpu().numpy(); p = p_d.cpu().numpy()
        sen = sen_d.cpu().numpy() if sen_d is not None else None
        if self.rho_range is None:
            self.rho_range = image_range(rho)
            self.p_range = image_range(p)
        base = os.path.join(self.outdir, f"{self.prefix}_{index:04d}")
        write_ppm_pair(base + ".ppm", rho, self.rho_range, p, self.p_range, self.K)
        if sen is not None:
            write_ppm_sensor(os.path.join(self.outdir, f"{self.sensor_prefix}_{index:04d}.ppm"), sen, self.K)
        if self.binary:
            write_rt_bin(base + ".bin", rho, p, self.K)
        return rho, p, sen
