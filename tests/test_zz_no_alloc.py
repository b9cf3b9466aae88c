"""The allocation-free pyramid decomposition (shx_atrousdec_noalloc, level
temporaries in a static device array) against shx_atrousdec with its own
allocation, through the C ABI: equal bit for bit on two calls in a row, the
second with a smaller image, and an image larger than the static scratch is
rejected before anything runs."""
import ctypes

import numpy as np
import pytest
import torch



def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.gpu
def test_noalloc_decomposition_equals_allocating_one(gpu):
    from shearcuda_amd._lib import check, lib, ptr_array
    for L, nlev, seed in ((512, 4, 1), (256, 3, 2)):
        print("case atrousdec_noalloc L=%d nlev=%d" % (L, nlev), flush=True)
        x = torch.from_numpy((np.random.default_rng(seed).standard_normal((L, L)) * 50).astype(np.float32)).cuda()
        ya = [torch.empty((L, L), device="cuda") for _ in range(nlev + 1)]
        yb = [torch.full((L, L), float("nan"), device="cuda") for _ in range(nlev + 1)]
        check(lib.shx_atrousdec(ctypes.c_void_p(x.data_ptr()), L, nlev, ptr_array([t.data_ptr() for t in ya]), None,
                                _stream()))
        check(lib.shx_atrousdec_noalloc(ctypes.c_void_p(x.data_ptr()), L, nlev,
                                        ptr_array([t.data_ptr() for t in yb]), _stream()))
        torch.cuda.synchronize()
        for j, (a, b) in enumerate(zip(ya, yb)):
            assert torch.equal(a, b), (L, j)


@pytest.mark.gpu
def test_noalloc_rejects_images_beyond_the_scratch(gpu):
    from shearcuda_amd._lib import lib, ptr_array
    L = 2048
    x = torch.zeros((L, L), device="cuda")
    y = [torch.empty((L, L), device="cuda") for _ in range(5)]
    rc = lib.shx_atrousdec_noalloc(ctypes.c_void_p(x.data_ptr()), L, 4, ptr_array([t.data_ptr() for t in y]),
                                   _stream())
    assert rc == -1 and b"static scratch" in lib.shx_last_error()


def test_noalloc_size_limit_is_checked_before_any_device_call():
    """The size check comes first, so it needs no GPU (the pointers are never
    dereferenced)."""
    from shearcuda_amd._lib import lib, ptr_array
    planes = ptr_array([16] * 5)
    assert lib.shx_atrousdec_noalloc(ctypes.c_void_p(16), 2048, 4, planes, None) == -1
    assert b"static scratch" in lib.shx_last_error()
