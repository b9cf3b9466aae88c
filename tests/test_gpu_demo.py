"""GPU checks of the demo_denoise entry points: the C++ demo program
(include/shearcuda_amd.hpp over the C ABI) against the Python mirror, the
directional halves (convolution / applyfilters) against the full transforms,
the noise generator and the image conversion."""
import os
import subprocess

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "shearcuda_amd", "demo", "demo_denoise")


@pytest.fixture(scope="module")
def api(gpu):
    from shearcuda_amd import api
    return api


def test_convolution_and_applyfilters_match_transforms(api):
    print("case convolution/applyfilters L=256", flush=True)
    from shearcuda_amd.synth import noisy_ellipses
    L = 256
    shear = api.make_dictionary(L)
    x = torch.tensor(noisy_ellipses(L, seed=3)[1], dtype=torch.float32, device="cuda")
    C = api.shear_trans_cuda(x, "maxflat", shear)
    y = api.atrousdec_cuda(x, "maxflat", shear.nscales)
    D = api.convolution_cuda(y, shear)
    for a, b in zip(C, D):
        assert torch.equal(a, b)
    z = api.applyfilters_cuda(C, shear)
    z[0] = C[0]
    rec = api.atrousrec_cuda(z, "maxflat")
    want = api.inverse_shear_cuda(C, "maxflat", shear)
    assert float((rec - want).abs().max() / want.abs().max()) < 1e-6
    shear.close()


def test_add_noise_statistics_and_determinism(api):
    print("case add_noise", flush=True)
    n = 1 << 22
    a = api.add_noise_cuda(torch.zeros(n, device="cuda"), 2.5, seed=7)
    b = api.add_noise_cuda(torch.zeros(n, device="cuda"), 2.5, seed=7)
    c = api.add_noise_cuda(torch.zeros(n, device="cuda"), 2.5, seed=8)
    assert torch.equal(a, b) and not torch.equal(a, c)
    v = a.double().cpu().numpy() / 2.5
    assert abs(v.mean()) < 5e-3 and abs(v.std() - 1) < 5e-3
    # fourth moment of a Gaussian is 3; tails present
    assert abs(np.mean(v ** 4) - 3) < 0.05 and np.abs(v).max() > 4.5


def test_convert_float_uint8(api):
    print("case convert", flush=True)
    from shearcuda_amd._lib import check, lib
    import ctypes
    f = torch.tensor([-3.0, 0.0, 0.99, 1.5, 254.9, 255.0, 300.0], device="cuda")
    u = torch.empty(7, dtype=torch.uint8, device="cuda")
    check(lib.shx_convert(ctypes.c_void_p(u.data_ptr()), 5, ctypes.c_void_p(f.data_ptr()), 0, 7, None))
    torch.cuda.synchronize()
    assert u.cpu().tolist() == [0, 0, 0, 1, 254, 255, 255]
    back = torch.empty(7, device="cuda")
    check(lib.shx_convert(ctypes.c_void_p(back.data_ptr()), 0, ctypes.c_void_p(u.data_ptr()), 5, 7, None))
    torch.cuda.synchronize()
    assert back.cpu().tolist() == [0.0, 0.0, 0.0, 1.0, 254.0, 255.0, 255.0]


@pytest.mark.skipif(not os.path.exists(DEMO), reason="demo binary not built")
def test_cpp_demo_matches_python_mirror(api, tmp_path):
    print("case C++ demo L=256 sigma=20", flush=True)
    from shearcuda_amd.synth import ellipses, psnr
    L, sigma = 256, 20.0
    img = np.clip(ellipses(L, seed=11), 0, 255).astype(np.uint8)
    (tmp_path / "img.u8").write_bytes(img.tobytes())
    api.save_shearing_filters(str(tmp_path / "f.bin"))
    out = subprocess.run([DEMO, str(tmp_path / "f.bin"), str(tmp_path / "img.u8"), str(L), str(sigma),
                          str(tmp_path / "out.f32"), str(tmp_path / "noisy.f32")],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    res = np.fromfile(tmp_path / "out.f32", dtype=np.float32).reshape(L, L)
    noisy = np.fromfile(tmp_path / "noisy.f32", dtype=np.float32).reshape(L, L)
    # the same noise through the Python mirror
    x = api.add_noise_cuda(torch.tensor(img, dtype=torch.float32, device="cuda"), sigma, seed=1234)
    assert np.array_equal(x.cpu().numpy(), noisy)
    shear = api.make_dictionary(L)
    w = api.add_noise_cuda(torch.zeros(L, L, device="cuda"), 1.0, seed=4321)
    sc = api.nsst_scalars_cuda(shear, w)
    want = api.denoise_cuda(x, shear, sigma, sc)[0].cpu().numpy()
    p_demo, p_py = psnr(img.astype(np.float64), res), psnr(img.astype(np.float64), want)
    lines = dict(l.split(" = ", 1) for l in out.stdout.splitlines() if " = " in l)
    assert abs(float(lines["Denoised image PSNR"].split()[0]) - p_demo) < 1e-3
    assert abs(p_demo - p_py) < 0.01 and p_demo > psnr(img.astype(np.float64), noisy) + 5
    assert np.sqrt(np.mean((res - want) ** 2)) / np.sqrt(np.mean(want ** 2)) < 1e-3
    shear.close()
