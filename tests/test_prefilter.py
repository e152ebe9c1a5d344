"""The FP32 position mirror is a prefilter only: a pair is decided in FP32
when its FP32 squared distance lies outside cutoff^2 +- margin, where the
margin is derived from the largest box edge (4096 A) and the list radius.
This checks that bound empirically, with numpy's IEEE float32 arithmetic in
the order the kernels use, on pairs placed near the cutoff anywhere in the
largest box: the FP32 and FP64 squared distances never differ by the margin."""
import math

import numpy as np

MAX_BOX = 4096.0


def margin(rl):
    u = 2.0 ** -24
    delta = 2.5 * MAX_BOX * u
    return max(0.02, 2.0 * (2.0 * math.sqrt(3.0) * rl * delta + 3.0 * rl * rl * u))


def idnint(x):
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def test_fp32_squared_distance_within_margin():
    rng = np.random.default_rng(81)
    f = np.float32
    for rl in (5.2, 13.2, 23.2):
        box = np.array([MAX_BOX * 0.9999, MAX_BOX * 0.731, MAX_BOX * 0.9999])
        n = 400_000
        a = rng.random((n, 3)) * box
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        b = a + u * rl * (1.0 + 1e-3 * rng.uniform(-1, 1, n))[:, None]
        b[:, :2] %= box[:2]
        keep = (b[:, 2] >= 0) & (b[:, 2] < box[2])
        a, b = a[keep], b[keep]
        # FP64, the reference's minimum image (idnint in x, y)
        d = a - b
        d[:, :2] -= box[:2] * idnint(d[:, :2] / box[:2])
        d2 = (d * d).sum(1)
        # FP32 mirror: rounded coordinates, rounded box and inverse, rintf
        af, bf = a.astype(f), b.astype(f)
        bxf, byf = f(box[0]), f(box[1])
        ibx, iby = f(1.0 / box[0]), f(1.0 / box[1])
        vx = af[:, 0] - bf[:, 0]
        vy = af[:, 1] - bf[:, 1]
        vz = af[:, 2] - bf[:, 2]
        vx = vx - bxf * np.rint(vx * ibx)
        vy = vy - byf * np.rint(vy * iby)
        d2f = (vx * vx + vy * vy) + vz * vz
        err = np.abs(d2f.astype(np.float64) - d2).max()
        assert err < margin(rl) / 2, (rl, err, margin(rl))
