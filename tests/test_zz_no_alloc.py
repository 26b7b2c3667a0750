"""GPU test of fmmfft_execute_static (through the C ABI): the transform on
the library's static device scratch equals fmmfft_execute's output bit for
bit, on two calls in a row (the second with a smaller input, so it runs on
what the first left in the scratch), and a plan whose workspace exceeds
FMMFFT_STATIC_SCRATCH_BYTES is rejected before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def test_static_scratch_equals_execute(fmm, orc, torch):
    cases = [(1 << 20, 16, 12, 32, fmm.C64), (1 << 15, 8, 16, 32, fmm.C128)]  # second: smaller
    for i, (n, p, terms, leaf, dt) in enumerate(cases):
        say(f"static scratch call {i}: n={n} p={p} terms={terms} leaf={leaf} dtype={dt}")
        _, wb = fmm.plan_sizes(n, p, terms, leaf, dt)
        assert wb <= fmm.STATIC_SCRATCH_BYTES
        plan = fmm.Plan(n, p, terms, leaf, dt)
        x = torch.from_numpy(orc.normal(n, 50 + i)).to(plan.torch_dtype).cuda()
        ref = plan.execute(x)
        got = plan.execute_static(x)
        inv_ref = plan.execute(x, direction=fmm.INVERSE)
        inv_got = plan.execute_static(x, direction=fmm.INVERSE)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), i
        assert torch.equal(inv_got, inv_ref), i
        # and the result is the transform (oracle FP64 FFT of the same input)
        X = orc.fft64(x.cpu().numpy().astype(np.complex128))
        err = orc.rel_l2(got.cpu().numpy().astype(np.complex128), X)
        say(f"  rel-L2 {err:.2e}")
        assert err < (1e-6 if dt == fmm.C64 else 1e-11)
        plan.close()


def test_static_scratch_rejects_large_plans(fmm, torch):
    n = 1 << 24
    _, wb = fmm.plan_sizes(n, 64, 16, 64, fmm.C64)
    assert wb > fmm.STATIC_SCRATCH_BYTES
    plan = fmm.Plan(n, 64, 16, 64, fmm.C64)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    with pytest.raises(fmm.FmmFftError) as e:
        plan.execute_static(x)
    assert e.value.status == fmm.ETOOBIG
    plan.close()
