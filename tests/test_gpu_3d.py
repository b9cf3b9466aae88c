"""GPU parity of the 3D shearlet volume path (shx3_* through the C ABI) against
the FP64 oracle (oracle/nsst3d.py), plus size-independent properties at the
128^3 workload size."""
import numpy as np
import pytest
import torch

from oracle import nsst3d as o3

pytestmark = pytest.mark.gpu

TOL = 1e-5  # max |gpu - oracle| / max |oracle| per band / subband


@pytest.fixture(scope="module")
def mods(gpu):
    from shearcuda_amd import api, shearlet3d
    return api, shearlet3d


_cache = {}


def _shear(mods, n, dirs=(8, 8, 4), mode=1):
    key = (n, dirs, mode)
    if key not in _cache:
        _cache[key] = mods[1].Shear3D(n, dirs, mode)
    return _cache[key]


def _oracle_filters(dirs):
    key = ("F",) + tuple(dirs)
    if key not in _cache:
        _cache[key] = o3.meyer_filters_3d(list(dirs))
    return _cache[key]


def _vol(n, seed, nvol=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nvol, n, n, n)) * 20.0


def _rel(g, o):
    return float(np.abs(g - o).max() / max(np.abs(o).max(), 1e-30))


@pytest.mark.parametrize("n,mode", [(32, 1), (64, 2), (64, 1)])
def test_pyr_dec_rec_match_oracle(mods, n, mode):
    print("case pyr n=%d mode=%g" % (n, mode), flush=True)
    sh = _shear(mods, n, (2, 2, 2), mode)
    plan = mods[1].Plan3D(sh, 1)
    x = _vol(n, 11)
    subs = plan.pyr_dec(torch.tensor(x, dtype=torch.float32, device="cuda"))
    want = o3.pyr_nd_dec(x[0], 3, mode)
    got = [t[0].cpu().numpy() for t in sh.split_subs(subs)]
    assert [g.shape for g in got] == [w.shape for w in want]
    for g, w in zip(got, want):
        assert _rel(g, w) < TOL
    # reconstruction of the oracle's subbands
    flat = torch.tensor(np.concatenate([w.ravel() for w in want])[None], dtype=torch.float32, device="cuda")
    rec = plan.pyr_rec(flat)[0].cpu().numpy()
    assert _rel(rec, o3.pyr_nd_rec(want, mode)) < TOL
    plan.close()


@pytest.fixture(scope="module")
def case32(mods):
    n, dirs = 32, (8, 8, 4)
    sh = _shear(mods, n, dirs)
    F = _oracle_filters(dirs)
    x = _vol(n, 21)
    subs_o = o3.pyr_nd_dec(x[0], 3, 1)
    bands_o = []
    for l in range(3):
        conv = o3._SameConv(subs_o[l], (24, 24, 24))
        bands_o.append([conv(f) for f in o3.band_list(F, l)])
    return sh, F, x, subs_o, bands_o


def test_analysis_matches_oracle(mods, case32):
    print("case analysis n=32", flush=True)
    sh, F, x, subs_o, bands_o = case32
    assert sh.fft_sides == [48, 48, 32]
    plan = mods[1].Plan3D(sh, 1, group=8)
    flat = torch.tensor(np.concatenate([s.ravel() for s in subs_o])[None], dtype=torch.float32, device="cuda")
    bands = plan.analysis(flat)
    worst = 0.0
    for l, lvl in enumerate(sh.split_bands(bands)):
        g = lvl[0].cpu().numpy()
        scale = max(np.abs(b).max() for b in bands_o[l])
        for b in range(g.shape[0]):
            worst = max(worst, np.abs(g[b] - bands_o[l][b]).max() / scale)
            assert np.abs(g[b] - bands_o[l][b]).max() < TOL * max(np.abs(bands_o[l][b]).max(), 1e-3 * scale)
    print("worst band error (rel. to level max) %.3e" % worst)
    plan.close()


def _thresholds(sh, bands_o, frac=0.8):
    """Per-band thresholds near a fraction of each band's std (plenty of kept
    and zeroed coefficients), each placed in the middle of the widest gap
    between neighbouring |coefficients| within 1000 ranks of that level, so
    that FP32 rounding of the coefficients cannot move one across it."""
    out = []
    for lvl in bands_o:
        for b in lvl:
            a = np.sort(np.abs(b).ravel())
            i0 = int(np.searchsorted(a, frac * np.std(b)))
            lo, hi = max(i0 - 1000, 0), min(i0 + 1000, a.size - 1)
            gaps = a[lo + 1:hi + 1] - a[lo:hi]
            i = lo + int(np.argmax(gaps))
            out.append(0.5 * (a[i] + a[i + 1]))
    return np.array(out)


def _min_gap(bands_o, th):
    """Smallest distance of any |coefficient| to its band's threshold,
    relative to the band's max."""
    r, k = np.inf, 0
    for lvl in bands_o:
        for b in lvl:
            r = min(r, np.abs(np.abs(b) - th[k]).min() / np.abs(b).max())
            k += 1
    return r


def test_dec_thresh_rec_matches_oracle(mods, case32):
    print("case dec_thresh_rec n=32", flush=True)
    sh, F, x, subs_o, bands_o = case32
    th = _thresholds(sh, bands_o)
    assert _min_gap(bands_o, th) > 2e-5
    plan = mods[1].Plan3D(sh, 1, group=5)
    flat = torch.tensor(np.concatenate([s.ravel() for s in subs_o])[None], dtype=torch.float32, device="cuda")
    th_t = torch.tensor(th[None], dtype=torch.float32, device="cuda")
    rec = plan.dec_thresh_rec(flat, th_t)
    got = [t[0].cpu().numpy() for t in sh.split_subs(rec)[:3]]
    for l in range(3):
        t_l = th[sh.band_first[l]:sh.band_first[l + 1]]
        want = sum(np.where(np.abs(b) <= t, 0.0, b) for b, t in zip(bands_o[l], t_l))
        assert _rel(got[l], want) < 2 * TOL
    plan.close()


def test_analysis_adjoint_identity_and_oracle(mods, case32):
    print("case adjoint n=32", flush=True)
    sh, F, x, subs_o, bands_o = case32
    plan = mods[1].Plan3D(sh, 1, group=7)
    rng = np.random.default_rng(5)
    s = torch.tensor(rng.standard_normal((1, sh.subs_elems)), dtype=torch.float32, device="cuda")
    y = torch.tensor(rng.standard_normal((1, sh.bands_elems)), dtype=torch.float32, device="cuda")
    As = plan.analysis(s)
    Aty = plan.analysis_adjoint(y)
    nlev = sum(v ** 3 for v in sh.sides[:3])
    lhs = float(torch.dot(As[0].double(), y[0].double()))
    rhs = float(torch.dot(s[0, :nlev].double(), Aty[0, :nlev].double()))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs), (lhs, rhs)
    # against the oracle transpose (directional part only)
    yb = [[b for b in lvl[0].cpu().numpy().astype(np.float64)] for lvl in sh.split_bands(y)]
    got = [t[0].cpu().numpy() for t in sh.split_subs(Aty)[:3]]
    for l in range(3):
        n = sh.sides[l]
        want = np.zeros((n, n, n))
        for b, f in zip(yb[l], o3.band_list(F, l)):
            G = np.zeros((n + 23,) * 3)
            G[12:12 + n, 12:12 + n, 12:12 + n] = b
            C = np.fft.irfftn(np.fft.rfftn(G) * np.conj(np.fft.rfftn(f, s=G.shape, axes=(0, 1, 2))), s=G.shape,
                              axes=(0, 1, 2))
            want += C[:n, :n, :n]
        assert _rel(got[l], want) < TOL
    plan.close()


def test_denoise_matches_oracle(mods, case32):
    print("case denoise n=32", flush=True)
    sh, F, x, subs_o, bands_o = case32
    th = _thresholds(sh, bands_o, 1.2)
    assert _min_gap(bands_o, th) > 2e-5
    plan = mods[1].Plan3D(sh, 1)
    xt = torch.tensor(x, dtype=torch.float32, device="cuda")
    out = plan.denoise(xt, torch.tensor(th[None], dtype=torch.float32, device="cuda"))[0].cpu().numpy()
    rec = []
    for l in range(3):
        t_l = th[sh.band_first[l]:sh.band_first[l + 1]]
        rec.append(sum(np.where(np.abs(b) <= t, 0.0, b) for b, t in zip(bands_o[l], t_l)))
    rec.append(subs_o[3])
    want = o3.pyr_nd_rec(rec, 1)
    err = np.sqrt(np.mean((out - want) ** 2)) / np.sqrt(np.mean(want ** 2))
    print("denoise rel rms %.3e rel max %.3e" % (err, _rel(out, want)))
    assert _rel(out, want) < 2 * TOL
    plan.close()


@pytest.mark.parametrize("n,cw", [(32, 32), (64, 64), (64, 32)])
def test_fused_plane_kernel_equals_separate_passes(mods, n, cw):
    """The fused plane kernel (rows + cols of a plane in one launch) against
    the separate row and column passes, on every band."""
    print("case fused plane n=%d cw=%d" % (n, cw), flush=True)
    sh = _shear(mods, n)
    plan = mods[1].Plan3D(sh, 2, group=6)
    x = torch.tensor(_vol(n, 61, nvol=2), dtype=torch.float32, device="cuda")
    subs = plan.pyr_dec(x)
    plan.set_schedule(plane_cols=-1)
    ref = plan.analysis(subs).clone()
    plan.set_schedule(plane_cols=cw)
    got = plan.analysis(subs)
    assert float((got - ref).abs().max() / ref.abs().max()) < 1e-6
    plan.close()


def test_batch_equals_single(mods, case32):
    print("case batch n=32", flush=True)
    sh = case32[0]
    x = _vol(32, 33, nvol=3)
    th = torch.full((3, sh.nbands), 2.0, device="cuda")
    planb = mods[1].Plan3D(sh, 3, group=8)
    outb = planb.denoise(torch.tensor(x, dtype=torch.float32, device="cuda"), th)
    plan1 = mods[1].Plan3D(sh, 1, group=8)
    for i in range(3):
        o1 = plan1.denoise(torch.tensor(x[i:i + 1], dtype=torch.float32, device="cuda"), th[i:i + 1].contiguous())
        assert torch.equal(o1[0], outb[i])
    planb.close()
    plan1.close()


def test_graph_replays_match_eager(mods, case32):
    print("case graph n=32", flush=True)
    api = mods[0]
    sh = case32[0]
    plan = mods[1].Plan3D(sh, 2)
    x = torch.tensor(_vol(32, 44, nvol=2), dtype=torch.float32, device="cuda")
    th = torch.full((2, sh.nbands), 3.0, device="cuda")
    want = plan.denoise(x, th).clone()
    out = torch.empty_like(x)
    g = api.capture(lambda: plan.denoise(x, th, out))
    for _ in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, want)
    g.close()
    plan.close()


def test_identity_and_lowpass_at_128(mods):
    """At the 128^3 workload size: with zero thresholds the cone filters sum
    to a unit impulse, so denoise is the identity; with infinite thresholds
    only the lowpass survives."""
    print("case identity n=128", flush=True)
    sh = _shear(mods, 128)
    assert sh.sides == [128, 128, 64, 32] and sh.fft_sides == [144, 144, 80]
    plan = mods[1].Plan3D(sh, 1)
    x = torch.tensor(_vol(128, 55), dtype=torch.float32, device="cuda")
    out = plan.denoise(x, torch.zeros(1, sh.nbands, device="cuda"))
    err = float((out - x).abs().max() / x.abs().max())
    print("identity rel err %.3e" % err)
    assert err < TOL
    out2 = plan.denoise(x, torch.full((1, sh.nbands), float("inf"), device="cuda"))
    subs = plan.pyr_dec(x)
    nlev = sum(v ** 3 for v in sh.sides[:3])
    subs[:, :nlev] = 0
    want = plan.pyr_rec(subs)
    assert float((out2 - want).abs().max() / want.abs().max()) < TOL
    plan.close()
