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


def _rotation(rng):
    """Uniformly random 3D rotation (QR of a Gaussian matrix, sign-fixed)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoids(n, count=8, seed=0, axes=(8.0, 40.0), intensity=(0.0, 255.0)):
    """n^3 volume of filled ellipsoids: centres uniform in the cube, semi-axes
    uniform in ``axes`` voxels, random orientation, intensities uniform in
    ``intensity``, summed where they overlap."""
    rng = np.random.default_rng(seed)
    g = np.arange(n, dtype=np.float64)
    vol = np.zeros((n, n, n))
    for _ in range(count):
        c = rng.uniform(0, n, size=3)
        a = rng.uniform(axes[0], axes[1], size=3)
        R = _rotation(rng)
        val = rng.uniform(intensity[0], intensity[1])
        d = [g - c[0], g - c[1], g - c[2]]
        # coordinates in the ellipsoid frame, u_k = sum_i R[i, k] d_i
        q = np.zeros((n, n, n))
        for k in range(3):
            u = (R[0, k] * d[0][:, None, None] + R[1, k] * d[1][None, :, None] + R[2, k] * d[2][None, None, :])
            q += (u / a[k]) ** 2
        vol += val * (q <= 1.0)
    return vol


def noisy_ellipsoids(n, sigma=25.0, seed=0, count=8):
    """(clean, noisy) volume pair: ellipsoids plus i.i.d. N(0, sigma^2) noise."""
    clean = ellipsoids(n, count=count, seed=seed)
    rng = np.random.default_rng(seed + 1_000_003)
    return clean, clean + sigma * rng.standard_normal((n, n, n))


def white_noise_volume(n, seed=12345):
    """Unit-variance white-noise volume for the per-band noise scalars."""
    return np.random.default_rng(seed).standard_normal((n, n, n))
