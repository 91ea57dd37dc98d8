"""The allocation-free mean (vf_mean_noalloc) against vf_mean through the C ABI: two calls in a row, the
second on a smaller input, with host and device outputs, and the size limit of its fixed scratch array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch

    assert torch.cuda.is_available()
    from voxelfem_amd import capi

    return capi.lib()


def _mean(lib, x):
    out = ctypes.c_double()
    rc = lib.vf_mean(x.numel(), ctypes.c_void_p(x.data_ptr()), 1, None, ctypes.byref(out))
    assert rc == 0, lib.vf_last_error()
    return out.value


def _mean_noalloc(lib, x, device_out):
    import torch

    if device_out:
        o = torch.full((1,), float("nan"), dtype=torch.float64, device=x.device)
        rc = lib.vf_mean_noalloc(x.numel(), ctypes.c_void_p(x.data_ptr()), None, ctypes.c_void_p(o.data_ptr()), 1)
        assert rc == 0, lib.vf_last_error()
        torch.cuda.synchronize()
        return float(o.item())
    out = ctypes.c_double(float("nan"))
    rc = lib.vf_mean_noalloc(x.numel(), ctypes.c_void_p(x.data_ptr()), None, ctypes.byref(out), 0)
    assert rc == 0, lib.vf_last_error()
    return out.value


@pytest.mark.parametrize("device_out", [False, True])
def test_mean_noalloc_matches_vf_mean(lib, device_out):
    import torch

    rng = np.random.default_rng(7)
    # the second call is smaller: its result must not see the first call's partial sums
    for n in (14_155_776, 1_000_003, 2049, 1):
        x = torch.from_numpy(rng.uniform(0.0, 1.0, n)).cuda()
        ref = _mean(lib, x)
        got = _mean_noalloc(lib, x, device_out)
        exact = float(np.mean(x.cpu().numpy(), dtype=np.float64))
        print(f"n={n} device_out={device_out}: vf_mean {ref!r} noalloc {got!r} numpy {exact!r}", flush=True)
        if device_out:
            assert got == pytest.approx(ref, rel=1e-15, abs=0)
        else:
            assert got == ref
        assert got == pytest.approx(exact, rel=1e-13)


def test_mean_noalloc_rejects_oversize(lib):
    import torch

    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    out = ctypes.c_double()
    rc = lib.vf_mean_noalloc(67_108_865, ctypes.c_void_p(x.data_ptr()), None, ctypes.byref(out), 0)
    assert rc != 0
    assert b"exceeds the maximum" in lib.vf_last_error()
