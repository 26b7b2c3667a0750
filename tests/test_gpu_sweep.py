"""GPU parity over the parameter space the C ABI accepts: leaf sizes 4..128
(the vector leaf kernel below 16, the matrix-core one from 16), Chebyshev
orders 2..32 (odd ones included, padded to 16 or 32 inside the kernels),
P = 2..64, one- and two-pass local FFTs, both precisions.  Every case is
checked against the oracle's FMM-FFT of the same input (the same
approximation, so the bound does not depend on the order): complex128 to
1e-12 relative, complex64 to 5e-6 relative (FP32 arithmetic)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (N, P, terms, leaf): each leaf size, order and P appears several times;
# s = N / (P leaf) >= 8 leaves; leaf * P within the plan limits
CASES = [
    (1 << 10, 4, 5, 4),
    (1 << 11, 2, 16, 8),
    (1 << 12, 8, 2, 8),
    (1 << 13, 16, 13, 4),
    (1 << 12, 2, 3, 16),
    (1 << 14, 16, 32, 16),
    (1 << 13, 4, 17, 32),
    (1 << 15, 64, 8, 32),
    (1 << 15, 2, 12, 32),    # complex64: FP16 leaf with two sections per workgroup
    (1 << 16, 32, 24, 32),
    (1 << 14, 2, 31, 64),
    (1 << 16, 8, 16, 64),
    (1 << 17, 64, 11, 64),
    (1 << 15, 4, 20, 128),
    (1 << 17, 16, 15, 128),
    (1 << 18, 32, 12, 128),  # leaf * P = 4096
    (1 << 22, 2, 9, 64),     # three-pass local FFT (M = 2^21)
    (1 << 23, 2, 14, 16),    # three-pass, leaf 16 (M = 2^22)
]


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,p,terms,leaf", CASES)
@pytest.mark.parametrize("dtype", [1, 0])
def test_sweep_matches_oracle(fmm, orc, torch, n, p, terms, leaf, dtype):
    say(f"sweep n={n} p={p} terms={terms} leaf={leaf} dtype={dtype}")
    x = orc.normal(n, 1000 + terms + leaf)
    if dtype == 0:
        x = x.astype(np.complex64).astype(np.complex128)
    plan = fmm.Plan(n, p, terms, leaf, dtype)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(plan.torch_dtype).cuda()
    X = plan.execute(xd).cpu().numpy().astype(np.complex128)
    ref = orc.OraclePlan(n, p, terms, leaf).fmmfft(x)
    err = rel(X, ref)
    assert err < (1e-12 if dtype == 1 else 5e-6), err
    # and the inverse direction through the same plan
    Xi = plan.execute(xd, direction=-1).cpu().numpy().astype(np.complex128)
    refi = orc.OraclePlan(n, p, terms, leaf).fmmfft(x, direction=-1)
    assert rel(Xi, refi) < (1e-12 if dtype == 1 else 5e-6)
