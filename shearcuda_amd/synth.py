"""Seeded synthetic inputs of the benchmark and test workloads (host, FP64).

The reference's demos use photographs and MATLAB's randn; these generators
replace them with a seeded numpy PCG64 stream so that the GPU path, the
oracle and the benchmark see identical inputs.
"""
import numpy as np


def ellipses(L, n=12, seed=0, axes=(10.0, 80.0), intensity=(0.0, 255.0)):
    """L x L image of n filled ellipses: centres uniform in the frame,
    semi-axes uniform in ``axes`` pixels, rotation uniform in [0, pi),
    intensities uniform in ``intensity``, summed where they overlap."""
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:L, 0:L].astype(np.float64)
    img = np.zeros((L, L))
    for _ in range(n):
        cy, cx = rng.uniform(0, L, size=2)
        a, b = rng.uniform(axes[0], axes[1], size=2)
        th = rng.uniform(0, np.pi)
        val = rng.uniform(intensity[0], intensity[1])
        dy, dx = r - cy, c - cx
        u = dx * np.cos(th) + dy * np.sin(th)
        v = -dx * np.sin(th) + dy * np.cos(th)
        img += val * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img


def noisy_ellipses(L, sigma=20.0, seed=0, n=12):
    """(clean, noisy) pair: ellipses plus i.i.d. N(0, sigma^2) noise."""
    clean = ellipses(L, n=n, seed=seed)
    rng = np.random.default_rng(seed + 1_000_003)
    return clean, clean + sigma * rng.standard_normal((L, L))


def white_noise(L, seed=12345):
    """Unit-variance white-noise image for the per-band noise scalars."""
    return np.random.default_rng(seed).standard_normal((L, L))


def psnr(ref, x, peak=255.0):
    mse = np.mean((np.asarray(ref, dtype=np.float64) - np.asarray(x, dtype=np.float64)) ** 2)
    return 10.0 * np.log10(peak * peak / mse)


def bezier_curves(L, n=20, seed=0, width=(1.0, 2.0), value=-60.0):
    """L x L image of n quadratic Bezier curves (control points uniform in the
    frame), each `width` pixels wide (uniform in the range), set to `value`."""
    rng = np.random.default_rng(seed)
    img = np.zeros((L, L))
    r, c = np.mgrid[0:L, 0:L].astype(np.float64)
    for _ in range(n):
        P = rng.uniform(0, L, size=(3, 2))
        w = rng.uniform(width[0], width[1])
        t = np.linspace(0.0, 1.0, 8 * L)[:, None]
        pts = (1 - t) ** 2 * P[0] + 2 * (1 - t) * t * P[1] + t ** 2 * P[2]
        mask = np.zeros((L, L), dtype=bool)
        rad = w / 2.0
        for py, px in pts[:: 2]:
            y0, y1 = int(max(0, np.floor(py - rad))), int(min(L, np.ceil(py + rad) + 1))
            x0, x1 = int(max(0, np.floor(px - rad))), int(min(L, np.ceil(px + rad) + 1))
            if y0 >= y1 or x0 >= x1:
                continue
            d2 = (r[y0:y1, x0:x1] - py) ** 2 + (c[y0:y1, x0:x1] - px) ** 2
            mask[y0:y1, x0:x1] |= d2 <= rad * rad
        img[mask] = value
    return img


def separation_input(L, seed=0, sigma=10.0):
    """(cartoon, curves, observed): 12 ellipses + 20 Bezier curves (-60) +
    N(0, sigma^2) noise."""
    cartoon = ellipses(L, n=12, seed=seed)
    curves = bezier_curves(L, n=20, seed=seed + 7)
    rng = np.random.default_rng(seed + 2_000_003)
    return cartoon, curves, cartoon + curves + sigma * rng.standard_normal((L, L))
