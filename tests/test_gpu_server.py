"""GPU tests of the request server (fmmfft_server_*, through the C ABI):
every result of a mixed request stream equals the direct execute of the same
input (bit for bit: batched and single launches run the same arithmetic,
and a batch reads and writes every request's own buffers in place through
a device table of addresses),
requests that cannot fit the HBM budget fail cleanly, idle shapes are
evicted to admit new ones, and the statistics add up."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch

    assert torch.cuda.is_available()
    return torch


def shape_for(k, terms, c128):
    """Serving-form shape: N = 2^k, P = 2^min(6, k // 4), leaf 64 (32 when
    M < 1024)."""
    n = 1 << k
    p = 1 << min(6, k // 4)
    m = n // p
    leaf = 64 if m >= 1024 else 32
    return n, p, terms, leaf, 1 if c128 else 0


def make_requests(torch, fmm, count, seed, kmin=12, kmax=24):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        terms = int(rng.choice([12, 20]))
        c128 = rng.random() < 0.3
        n, p, t, leaf, dt = shape_for(k, terms, c128)
        g = torch.Generator(device="cuda").manual_seed(int(seed * 100003 + i))
        dtype = torch.complex128 if dt == 1 else torch.complex64
        x = torch.randn(n, dtype=dtype, device="cuda", generator=g)
        reqs.append(((n, p, t, leaf, dt), x, torch.empty_like(x)))
    # the inputs are complete on the device before they are submitted
    torch.cuda.synchronize()
    return reqs


def test_mixed_stream_matches_direct_execute(fmm, orc, torch):
    """64 requests, k in 12..24, both precisions, terms 12 and 20: every
    output equals the direct execute bit for bit, and one request per (k,
    dtype) matches the oracle's FP64 FFT."""
    say("server mixed stream")
    reqs = make_requests(torch, fmm, 64, 1)
    ks = {(r[0][0].bit_length() - 1, r[0][4]) for r in reqs}
    assert any(k > 20 for k, _ in ks) and {dt for _, dt in ks} == {0, 1}
    srv = fmm.Server(8 << 30, max_batch=16)
    tickets = [srv.submit(x, out, *shape) for shape, x, out in reqs]
    lats = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and all(v >= 0 for v in lats)
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["batches"] < len(reqs)  # same-shape requests were batched
    plans = {}
    for shape, x, out in reqs:
        if shape not in plans:
            plans[shape] = fmm.Plan(*shape)
        ref = plans[shape].execute(x)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), shape
        del ref
    # and against the oracle for one request of each (k, dtype): complex64
    # to rel-L2 <= 1e-6 of the FP64 FFT; complex128 to 1e-12 of the oracle's
    # FMM-FFT (the same approximation) and to the FP64 FFT within the
    # algorithm's own truncation error at that order (1e-11 for 20 terms;
    # 12 terms reach 1.2e-11 at M = 2^18 in the oracle itself, so 3e-11)
    seen = set()
    for shape, x, out in reqs:
        key = (shape[0], shape[4])
        if key in seen:
            continue
        seen.add(key)
        xh = x.cpu().numpy().astype(np.complex128)
        got = out.cpu().numpy().astype(np.complex128)
        X = orc.fft64(xh)
        err = orc.rel_l2(got, X)
        say(f"  N=2^{shape[0].bit_length() - 1} dtype={shape[4]} terms={shape[2]} rel-L2 {err:.2e}")
        if shape[4] == 1:
            fmm_ref = orc.OraclePlan(shape[0], shape[1], shape[2], shape[3]).fmmfft(xh)
            assert orc.rel_l2(got, fmm_ref) < 1e-12, shape
            assert err < (1e-11 if shape[2] >= 20 else 3e-11), (shape, err)
        else:
            assert err < 1e-6, (shape, err)


def test_alone_equals_in_the_mix(fmm, torch):
    say("server alone vs mix")
    reqs = make_requests(torch, fmm, 32, 2)
    srv = fmm.Server(8 << 30, max_batch=8)
    alone = []
    for shape, x, out in reqs[:8]:
        o = torch.empty_like(x)
        srv.wait(srv.submit(x, o, *shape))
        alone.append(o)
    tickets = [srv.submit(x, out, *shape) for shape, x, out in reqs]
    for t in tickets:
        srv.wait(t)
    srv.close()
    assert any(shape[0] > (1 << 20) for shape, _, _ in reqs[:8])
    for a, (_, _, out) in zip(alone, reqs[:8]):
        assert torch.equal(a, out)


def test_request_larger_than_budget_fails_cleanly(fmm, torch):
    say("server budget reject")
    srv = fmm.Server(1 << 20, max_batch=4)  # 1 MiB: no plan fits
    n, p, t, leaf, dt = shape_for(16, 12, False)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    out = torch.empty_like(x)
    tk = srv.submit(x, out, n, p, t, leaf, dt)
    with pytest.raises(fmm.FmmFftError) as e:
        srv.wait(tk)
    assert e.value.status == fmm.ENOMEM
    # the failed request's latency is reported (not negative)
    tk = srv.submit(x, out, n, p, t, leaf, dt)
    lat = ctypes.c_double(-1.0)
    assert fmm.lib().fmmfft_server_wait(srv._h, tk, ctypes.byref(lat)) == fmm.ENOMEM
    assert 0.0 <= lat.value < 60e3
    with pytest.raises(fmm.FmmFftError):
        srv.submit(x, out, n, 3, t, leaf, dt)  # invalid P
    srv.close()


def test_idle_shapes_are_evicted(fmm, torch):
    say("server eviction")
    a = shape_for(16, 12, False)
    b = shape_for(16, 20, True)
    _, wa = fmm.plan_sizes(*a)
    _, wb = fmm.plan_sizes(*b)
    # room for one shape (tables, two slots of workspace and address table)
    # at a time
    ta, _ = fmm.plan_sizes(*a)
    tb, _ = fmm.plan_sizes(*b)
    na = ta + 2 * (wa + 16)
    nb = tb + 2 * (wb + 16)
    budget = max(na, nb) + min(na, nb) // 2
    outs = []
    for i in range(6):
        shape = a if i % 2 == 0 else b
        dtype = torch.complex64 if shape[4] == 0 else torch.complex128
        x = torch.randn(shape[0], dtype=dtype, device="cuda")
        outs.append((shape, x, torch.empty_like(x)))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    srv = fmm.Server(budget, max_batch=1)
    low = free0
    for shape, x, o in outs:
        srv.wait(srv.submit(x, o, *shape))
        low = min(low, torch.cuda.mem_get_info()[0])
    st = srv.stats()
    srv.close()
    assert st["plans_evicted"] >= 1 and st["peak_bytes"] <= budget
    # what the device really lost while the server ran stays within the
    # budget (plus a little for the runtime's own stream, event and graph
    # objects): every allocation of the server is charged
    say(f"  device memory used {(free0 - low) / 2**20:.1f} MiB, budget {budget / 2**20:.1f} MiB")
    assert free0 - low <= budget + (32 << 20)
    for shape, x, o in outs:
        assert torch.equal(o, fmm.Plan(*shape).execute(x))


def test_submit_waits_for_the_callers_stream(fmm, torch):
    """The input is produced on a side stream behind a long queue of work and
    submitted without synchronizing: the server must wait for that stream."""
    say("server stream order")
    n, p, t, leaf, dt = shape_for(18, 12, False)
    side = torch.cuda.Stream()
    a = torch.randn(4096, 4096, device="cuda")
    src = torch.randn(n, dtype=torch.complex64, device="cuda")
    x = torch.zeros_like(src)
    out = torch.empty_like(src)
    torch.cuda.synchronize()
    srv = fmm.Server(4 << 30, max_batch=4)
    with torch.cuda.stream(side):
        for _ in range(20):
            a = a @ a * 1e-3  # keeps the side stream busy for a while
        x.copy_(src)
        tk = srv.submit(x, out, n, p, t, leaf, dt, stream=side)
    srv.wait(tk)
    srv.close()
    torch.cuda.synchronize()
    assert torch.equal(out, fmm.Plan(n, p, t, leaf, dt).execute(src))


def test_large_requests_batch_in_place(fmm, torch):
    """Requests above 2^20 points are batched too (2^21: four per batch),
    half of them in place (input buffer = output buffer): every result
    equals the direct out-of-place execute bit for bit, and the eight
    requests took fewer than eight launches."""
    say("server large batches in place")
    n, p, t, leaf, dt = shape_for(21, 20, False)
    xs = [torch.randn(n, dtype=torch.complex64, device="cuda") for _ in range(8)]
    refs = []
    plan = fmm.Plan(n, p, t, leaf, dt)
    for x in xs:
        refs.append(plan.execute(x))
    bufs = [x.clone() if i % 2 == 0 else torch.empty_like(x) for i, x in enumerate(xs)]
    torch.cuda.synchronize()
    srv = fmm.Server(16 << 30, max_batch=64)
    tickets = []
    for i, x in enumerate(xs):
        if i % 2 == 0:
            tickets.append(srv.submit(bufs[i], bufs[i], n, p, t, leaf, dt))  # in place
        else:
            tickets.append(srv.submit(x, bufs[i], n, p, t, leaf, dt))
    for tk in tickets:
        srv.wait(tk)
    st = srv.stats()
    srv.close()
    torch.cuda.synchronize()
    for i in range(8):
        assert torch.equal(bufs[i], refs[i]), i
    assert st["completed"] == 8 and st["batches"] < 8


@pytest.mark.parametrize("n,p,terms,leaf", [(1 << 22, 2, 16, 32),   # M = 2^21: three-pass local FFT
                                            (1 << 14, 2, 12, 16),   # P = 2
                                            (1 << 15, 4, 20, 32)])  # P = 4
def test_other_layouts_through_the_server(fmm, torch, n, p, terms, leaf):
    """Shapes the serving stream does not draw (the three-pass local FFT,
    P = 2 and 4), both precisions, three requests each in one batch: every
    result equals the direct execute bit for bit."""
    say(f"server layouts n={n} p={p}")
    srv = fmm.Server(16 << 30, max_batch=8)
    cases = []
    for dt in (0, 1):
        dtype = torch.complex128 if dt == 1 else torch.complex64
        g = torch.Generator(device="cuda").manual_seed(n + dt)
        xs = [torch.randn(n, dtype=dtype, device="cuda", generator=g) for _ in range(3)]
        outs = [torch.empty_like(x) for x in xs]
        cases.append((dt, xs, outs))
    torch.cuda.synchronize()
    tickets = [srv.submit(x, o, n, p, terms, leaf, dt) for dt, xs, outs in cases for x, o in zip(xs, outs)]
    for t in tickets:
        srv.wait(t)
    srv.close()
    torch.cuda.synchronize()
    for dt, xs, outs in cases:
        plan = fmm.Plan(n, p, terms, leaf, dt)
        for x, o in zip(xs, outs):
            assert torch.equal(o, plan.execute(x)), (n, p, dt)
        plan.close()
