"""CPU checks of the C++ demo entry points (include/shearcuda_amd.hpp): the
SHFM0002 filter file written by the Python mirror is read back by the C++
ShearConfig (no GPU)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "shearcuda_amd", "demo", "demo_denoise")


@pytest.mark.skipif(not os.path.exists(DEMO), reason="demo binary not built (run __graft_entry__.build())")
@pytest.mark.parametrize("dtype", ["single", "double"])
def test_filter_file_round_trip(tmp_path, dtype):
    from shearcuda_amd import api
    from shearcuda_amd.filters import shearing_filters_myer
    path = str(tmp_path / "f.bin")
    api.save_shearing_filters(path, (32, 16), (3, 4), dtype)
    out = subprocess.run([DEMO, "--check-config", path, "1", "3"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "scales 2 type %d" % (0 if dtype == "single" else 2)
    for j, (n1, lev) in enumerate([(32, 3), (16, 4)]):
        w = shearing_filters_myer(n1, lev)
        f = lines[1 + j].split()
        assert (int(f[3]), int(f[5])) == (n1, 2 ** lev)
        v = float(f[7])
        want = float(np.float32(w[0, 1, 3])) if dtype == "single" else w[0, 1, 3]
        assert abs(v - want) <= 1e-9 * max(1.0, abs(want))
        assert abs(v - w[0, 3, 1]) > 1e-6  # the transposed element differs
    h = api.maxflat_filters()
    assert lines[-1] == "atrous " + " ".join(str(a.shape[0]) for a in h)


@pytest.mark.skipif(not os.path.exists(DEMO), reason="demo binary not built (run __graft_entry__.build())")
def test_bad_filter_file_is_rejected(tmp_path):
    path = tmp_path / "bad.bin"
    path.write_bytes(b"SHFM0001" + b"\0" * 16)
    out = subprocess.run([DEMO, "--check-config", str(path), "0", "0"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "Failed to read" in out.stderr
