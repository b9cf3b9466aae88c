"""Independent pins of the 2D pseudo-polar wedge geometry of the Meyer
shearing filters (shearlet_toolbox_1/gen_x_y_cordinates.m:22-111,
rec_from_pol.m:16-35, shearing_filters_Myer.m:18-27, windowing.m:11-27).

The oracle (oracle/nsst.py) and the product (shearcuda_amd/filters.py) both
transcribe those files.  The checks here come from the geometry the files
implement instead, so a transcription mistake shared by both fails them:

* the 2n radial lines of the pseudo-polar grid of an n x n block (computed
  on an (n+1)-point grid whose centre is c = n/2 + 1): line r = 1..n joins
  (x = n+1-r, y = 1) to (x = r+1, y = n+1) and is sampled at y = 1..n; line
  n+r joins (x = 1, y = r+1) to (x = n+1, y = n+1-r) and is sampled at
  x = 1..n (the sample leaving the block is clamped to n).  Built here with
  exact rational arithmetic;
* the lines are in angular order: their orientation decreases strictly from
  135 to -45 degrees over r = 1..2n, read from the oracle's own tables;
* windowing.m places band k's Meyer window on consecutive lines, so each
  direction's frequency support (rec_from_pol before the inverse FFT) is one
  angular sector: 1 inside the flat part of the window, 0 outside its
  support, both bounded by the orientations of known lines;
* the spatial filter is the centred inverse FFT of that support, so its
  centred spectrum is the point-symmetric part of the support, times sqrt(n).

Each check is also run on the oracle with a planted mistake (the flipud of
gen_x_y_cordinates.m:74 removed; the x2/y2 roles swapped), and must fail.
"""
import inspect
import math
import textwrap
from fractions import Fraction

import numpy as np
import pytest

from oracle import nsst as o
from shearcuda_amd import filters as f

SCALES = [(32, 3), (32, 4), (16, 4), (16, 3), (8, 2)]


def _round_half_up(q):
    return int(math.floor(q + Fraction(1, 2)))


def closed_form_lines(n):
    """(x1, y1, x2, y2), 1-based, of the 2n pseudo-polar lines, from their
    end points (exact rationals; sample positions never fall within an ulp
    of a rounding boundary here, so this is what the FP64 MATLAB code
    computes for n a power of two)."""
    x1 = np.zeros((n, n), int)
    y1 = np.zeros((n, n), int)
    x2 = np.zeros((n, n), int)
    y2 = np.zeros((n, n), int)
    for r in range(1, n + 1):
        for s in range(1, n + 1):
            # line r: x moves from n+1-r to r+1 as y goes 1 -> n+1
            x1[r - 1, s - 1] = _round_half_up(Fraction(n + 1 - r) + Fraction((s - 1) * (2 * r - n), n))
            y1[r - 1, s - 1] = s
            # line n+r: y moves from r+1 to n+1-r as x goes 1 -> n+1
            x2[r - 1, s - 1] = s
            y2[r - 1, s - 1] = min(n, _round_half_up(Fraction(r + 1) + Fraction((s - 1) * (n - 2 * r), n)))
    return x1, y1, x2, y2


def line_orientations(x1, y1, x2, y2):
    """Orientation (degrees) of each of the 2n lines, from the first and last
    sample the tables hold."""
    ang = []
    for xs, ys in ((x1, y1), (x2, y2)):
        for r in range(xs.shape[0]):
            dx = float(xs[r, -1] - xs[r, 0])
            dy = float(ys[r, -1] - ys[r, 0])
            if dy < 0 or (dy == 0 and dx < 0):
                dx, dy = -dx, -dy
            ang.append(math.degrees(math.atan2(dy, dx)))
    ang = np.array(ang)
    # orientations are defined mod 180; unwrap along the line order
    return np.degrees(np.unwrap(np.radians(2 * ang)) / 2)


def check_lines_closed_form(grid, n):
    x1, y1, x2, y2, _ = grid
    for got, want in zip((x1, y1, x2, y2), closed_form_lines(n)):
        assert np.array_equal(got, want)


def check_lines_in_angular_order(grid, n):
    """The line orientations decrease strictly and cover half a turn."""
    th = line_orientations(*grid[:4])
    d = np.diff(th)
    assert (d < 0).all(), "pseudo-polar lines not in angular order"
    # first line ~135 deg, last ~ -45 deg (one sample short of a half turn)
    assert abs(th[0] - th[-1] - 180.0) < 360.0 / n + 1e-9


def _window_rays(n, D, k, eps=1e-20):
    """0-based line indices where band k of windowing(ones(2n), D) is
    positive, and where it equals 1."""
    wf = o.windowing(np.ones(2 * n), D)[:, k]
    return np.nonzero(wf > eps)[0], np.nonzero(wf == 1.0)[0]


def _ray_theta(n, r0):
    """Orientation (degrees, decreasing in r0) of 0-based line r0 from its
    closed-form direction vector."""
    r = r0 % (2 * n) + 1
    if r <= n:
        t = math.degrees(math.atan2(n, 2 * r - n))
    else:
        t = math.degrees(math.atan2(n - 2 * (r - n), n))
    # continue below -45 / above 135 for windows that wrap around the list
    return t - 180.0 * (r0 // (2 * n))


def _unwrap_cyclic(idx, m):
    """Sorted indices of a contiguous cyclic range of 0..m-1, the part after
    the wrap moved below 0."""
    idx = np.sort(idx)
    gap = np.nonzero(np.diff(idx) > 1)[0]
    if gap.size == 0:
        return idx
    assert gap.size == 1, "window lines not contiguous"
    return np.concatenate([idx[gap[0] + 1:] - m, idx[: gap[0] + 1]])


def check_direction_sectors(C, n, D):
    """C[k] = rec_from_pol output of direction k (n x n).  Away from the
    centre, C[k] is positive only inside the angular sector of the lines
    where window k is positive, and equals 1 only inside the sector of the
    lines where it is flat, and everywhere well inside that sector.  The
    margin is the angular size of one cell at the cell's radius (lines are
    rounded to the nearest cell; one end sample is clamped by a cell)."""
    c = n // 2      # 0-based centre index (n/2 + 1 in MATLAB)
    ys, xs = np.mgrid[0:n, 0:n]
    rho = np.hypot(xs - c, ys - c)
    phi = np.degrees(np.arctan2(ys - c, xs - c)) % 180.0
    far = rho >= 3.0
    margin = np.degrees(np.arcsin(np.minimum(1.0, 1.0 / np.maximum(rho, 1.0)))) + 1e-9

    def inside(a, b, m):
        # orientation within [b - m, a + m] (mod 180), a >= b
        mid, half = (a + b) / 2.0, (a - b) / 2.0
        return np.abs((phi - mid + 90.0) % 180.0 - 90.0) <= half + m

    for k in range(D):
        pos, flat = _window_rays(n, D, k)
        pos, flat = _unwrap_cyclic(pos, 2 * n), _unwrap_cyclic(flat, 2 * n)
        if flat.min() > pos.max():          # flat part lies past the wrap
            flat = flat - 2 * n
        hi, lo = _ray_theta(n, pos.min()), _ray_theta(n, pos.max())
        fhi, flo = _ray_theta(n, flat.min()), _ray_theta(n, flat.max())
        supp = far & (C[k] > 1e-12)
        one = far & (np.abs(C[k] - 1.0) < 1e-12)
        assert one.any()
        assert not (supp & ~inside(hi, lo, margin)).any(), ("direction", k, "positive outside its sector")
        assert not (one & ~inside(fhi, flo, margin)).any(), ("direction", k, "flat outside its flat sector")
        core = far & inside(fhi, flo, -margin)
        assert (C[k][core] > 1.0 - 1e-12).all(), ("direction", k, "not 1 inside its flat sector")


def _supports(gen, n, lev):
    x1, y1, x2, y2, F = gen(n)
    wf = o.windowing(np.ones(2 * n), 2 ** lev)
    return np.stack([o.rec_from_pol(np.outer(wf[:, k], np.ones(n)), n, x1, y1, x2, y2, F)
                     for k in range(2 ** lev)])


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_lines_equal_closed_form(n):
    check_lines_closed_form(o.gen_x_y_cordinates(n), n)
    check_lines_closed_form(f.pseudo_polar_grid(n), n)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 80])
def test_lines_in_angular_order(n):
    check_lines_in_angular_order(o.gen_x_y_cordinates(n), n)
    check_lines_in_angular_order(f.pseudo_polar_grid(n), n)


@pytest.mark.parametrize("n,lev", SCALES)
def test_direction_supports_are_angular_sectors(n, lev):
    C = _supports(o.gen_x_y_cordinates, n, lev)
    check_direction_sectors(C, n, 2 ** lev)
    assert np.allclose(C.sum(0), 1.0, atol=1e-14)


@pytest.mark.parametrize("n,lev", SCALES)
def test_spatial_filters_are_centred_inverse_ffts_of_the_sectors(n, lev):
    """fftshift(fft2(ifftshift(w_k))) * sqrt(n) is the point-symmetric part of
    the sector C_k (real part of the centred inverse FFT), for the oracle's
    and the product's filters."""
    C = _supports(o.gen_x_y_cordinates, n, lev)
    c = n // 2
    idx = (2 * c - np.arange(n)) % n                 # reflection through the centre
    sym = 0.5 * (C + C[:, idx][:, :, idx])
    for w in (o.shearing_filters_myer(n, lev), f.shearing_filters_myer(n, lev)):
        spec = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(w, axes=(1, 2))), axes=(1, 2)) * math.sqrt(n)
        assert np.abs(spec.imag).max() < 1e-12
        assert np.abs(spec.real - sym).max() < 1e-12


# ---- planted mistakes: each check above must reject them ---------------------

def _mutant(*edits):
    """gen_x_y_cordinates of the oracle with source edits (old, new) applied."""
    src = textwrap.dedent(inspect.getsource(o.gen_x_y_cordinates))
    for old, new in edits:
        assert old in src, old
        src = src.replace(old, new)
    ns = dict(o.__dict__)
    exec(compile(src, "<mutant>", "exec"), ns)
    return ns["gen_x_y_cordinates"]


MUTANTS = {
    "no_flipud": [("    x1n = np.flipud(x1n)\n", "")],
    # x2 and y2 exchanged; the end-sample clamp follows the exchanged table so
    # that every index stays inside the block and the mutant runs
    "x2_y2_swapped": [("    x2n = x2[1:n + 1, :n].copy()\n    y2n = y2[1:n + 1, :n].copy()\n",
                       "    x2n = y2[1:n + 1, :n].copy()\n    y2n = x2[1:n + 1, :n].copy()\n"),
                      ("    y2n[n - 1, 0] = n\n", "    x2n[n - 1, 0] = n\n")],
}


def _fails(fn, *args):
    try:
        fn(*args)
    except AssertionError:
        return True
    return False


@pytest.mark.filterwarnings("ignore::RuntimeWarning")
@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistakes_are_caught(name):
    gen = _mutant(*MUTANTS[name])
    n, lev = 32, 3
    assert _fails(check_lines_closed_form, gen(n), n)
    assert _fails(check_lines_in_angular_order, gen(n), n)
    assert _fails(lambda: check_direction_sectors(_supports(gen, n, lev), n, 2 ** lev))
    # the mutant runs, and the hit-count checks alone cannot see the flipud
    x1, y1, x2, y2, F = gen(n)
    if name == "no_flipud":
        assert F.sum() == 2 * n * n and F.min() >= 1
