"""chip_bp_noalloc against chip_bp through the C ABI (run with -m gpu).

chip_bp_noalloc runs chip_bp's kernels with the part sums of a split
single-slice adjoint bp in the library's static device scratch. Two calls in a
row, the second on a smaller slice with the first call's partial sums still in
the scratch, must give chip_bp's bits; a skimage-mode call (no scratch) must
too, and a call whose parts do not fit the scratch is refused on the host
before anything is launched.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pycorrectedemissionct_amd import _native
    from pycorrectedemissionct_amd._projector_backends import HipPlan

    return torch, _native, HipPlan


def _case(torch, HipPlan, n, a, det, seed, maps=True):
    rng = np.random.default_rng(seed)
    ang = np.linspace(0, 2 * np.pi, a, endpoint=False).astype(np.float32)
    plan = HipPlan(n, ang, det_angles_rad=det)
    d = len(det)
    dev = torch.device("cuda", 0)
    sino = torch.from_numpy(rng.random((d, a, n), dtype=np.float32)).to(dev)
    m = torch.from_numpy(rng.uniform(0.2, 1.0, (a, d, n, n)).astype(np.float32)).to(dev) if maps else None
    return plan, sino, m


def _bp(torch, _native, name, plan, sino, maps, n, mode):
    out = torch.full((1, n, n), float("nan"), dtype=torch.float32, device=sino.device)
    rc = getattr(_native.lib(), name)(plan.handle, _native.ptr(sino), _native.ptr(maps), 0, _native.ptr(out), 1,
                                      mode, _native.stream())
    _native.check(rc, name)
    torch.cuda.synchronize()
    return out


def test_noalloc_matches_chip_bp_two_calls(env):
    torch, _native, HipPlan = env
    cases = [(256, 180, [np.pi / 2, -np.pi / 2], 31), (128, 96, [np.pi / 2], 32)]  # second call smaller
    for n, a, det, seed in cases:
        print(f"noalloc vs chip_bp: n={n} a={a} d={len(det)} seed={seed}", flush=True)
        plan, sino, maps = _case(torch, HipPlan, n, a, det, seed)
        ref = _bp(torch, _native, "chip_bp", plan, sino, maps, n, _native.CHIP_BP_ADJOINT)
        got = _bp(torch, _native, "chip_bp_noalloc", plan, sino, maps, n, _native.CHIP_BP_ADJOINT)
        assert torch.isfinite(got).all()
        diff = float((got - ref).abs().max())
        print(f"  max |noalloc - chip_bp| = {diff:.3e}")
        assert diff == 0.0
        plan.close()


def test_noalloc_skimage_mode_and_refusal(env):
    torch, _native, HipPlan = env
    plan, sino, _ = _case(torch, HipPlan, 96, 64, [np.pi / 2], 33, maps=False)
    ref = _bp(torch, _native, "chip_bp", plan, sino, None, 96, _native.CHIP_BP_SKIMAGE)
    got = _bp(torch, _native, "chip_bp_noalloc", plan, sino, None, 96, _native.CHIP_BP_SKIMAGE)
    assert float((got - ref).abs().max()) == 0.0
    plan.close()

    # 5 pinned parts of 1024^2 floats exceed the 4 Mi-float scratch: refused before any launch
    ang = np.linspace(0, np.pi, 96, endpoint=False).astype(np.float32)
    big = HipPlan(1024, ang)
    big.tune("bp_parts", 5)
    dev = torch.device("cuda", 0)
    sino = torch.zeros((1, 96, 1024), dtype=torch.float32, device=dev)
    out = torch.zeros((1, 1024, 1024), dtype=torch.float32, device=dev)
    rc = _native.lib().chip_bp_noalloc(big.handle, _native.ptr(sino), None, 0, _native.ptr(out), 1,
                                       _native.CHIP_BP_ADJOINT, _native.stream())
    msg = _native.lib().chip_last_error().decode()
    print(f"refusal: rc={rc} {msg}")
    assert rc == 1 and "static scratch" in msg
    big.close()
    assert ctypes.sizeof(ctypes.c_void_p) == 8
