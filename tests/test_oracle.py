"""CPU tests that pin the oracle (oracle/fmmfft_oracle.c) before it is trusted
as the checker of the HIP path.

Each operator table is checked against what the mathematics fixes -- the
far field of explicit source clusters evaluated directly -- not against a
retyped formula; the FMM against the brute-force cot sum and the
fractional-shift identity; the FFT against numpy golden vectors and a
long-double DFT; the whole path against the max errors the reference itself
recorded (plotting/raw_data/cluster_output_error.csv) and DFT invariants.
"""
import csv
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# ---------------------------------------------------------------------------
# geometry of the expansions (angle theta = pi * mu / M)
#   multipole node k of a box of half-width h: cot(theta_k) = c_k / (3 tan h)
#   local node k:                               tan(theta_k) = c_k tan h
# ---------------------------------------------------------------------------


def mp_nodes(cheb, h):
    return np.arctan2(3 * np.tan(h), cheb)


def loc_nodes(cheb, h):
    return np.arctan(cheb * np.tan(h))


def field(targets, sources, q):
    """sum_s q_s cot(t - s) for every target angle."""
    d = targets[:, None] - sources[None, :]
    return (np.cos(d) / np.sin(d)) @ q


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def plan16(orc):
    # N = 2^14, P = 4, M = 4096, terms 16, b = 32 -> s = 128, n = 7
    return orc.OraclePlan(1 << 14, 4, 16, 32)


def test_chebyshev_nodes_are_roots_of_T(plan16):
    c = plan16.cheb
    T = np.polynomial.chebyshev.Chebyshev.basis(plan16.terms)
    assert np.max(np.abs(T(c))) < 1e-13
    assert np.all(np.diff(c) > 0)


def test_s2m_is_far_field_at_multipole_nodes(plan16):
    pl = plan16
    a = math.pi / (2 * pl.s)  # leaf half-width
    th = (2 * np.arange(1, pl.b + 1) - pl.b - 1) / pl.b * a  # source angles
    nodes = mp_nodes(pl.cheb, a)
    d = nodes[None, :] - th[:, None]
    expect = np.cos(d) / np.sin(d)
    assert np.max(np.abs(pl.initch - expect) / np.abs(expect).max()) < 1e-13


@pytest.mark.parametrize("level", [2, 3, 5])
@pytest.mark.parametrize("which,side", [(0, -1), (1, +1)])
def test_m2m_translates_a_child_multipole(plan16, level, which, side):
    pl = plan16
    rng = np.random.default_rng(level * 7 + which)
    hc = math.pi / 2 ** (level + 2)  # child half-width, parent level `level`
    hp = 2 * hc
    src = rng.uniform(-hc, hc, 40)
    q = rng.standard_normal(40)
    child = field(mp_nodes(pl.cheb, hc), src, q)
    exact = field(mp_nodes(pl.cheb, hp), src + side * hc, q)
    mat = pl.ops[level - 2, which]
    assert rel(mat @ child, exact) < 1e-9


@pytest.mark.parametrize("level", [3, 4, 7])
@pytest.mark.parametrize("op,shft", [(2, -2), (3, 2), (4, -3), (5, 3)])
def test_flip_maps_multipole_to_local(plan16, level, op, shft):
    pl = plan16
    rng = np.random.default_rng(level * 13 + op)
    h = math.pi / 2 ** (level + 1)
    center = -shft * 2 * h  # source box centre in target coordinates
    src = rng.uniform(-h, h, 40)
    q = rng.standard_normal(40)
    mp = field(mp_nodes(pl.cheb, h), src, q)
    exact = field(loc_nodes(pl.cheb, h), src + center, q)
    mat = pl.ops[level - 3, op]
    assert rel(mat @ mp, exact) < 1e-9


def test_top_flip_maps_opposite_box(plan16):
    pl = plan16
    rng = np.random.default_rng(5)
    h = math.pi / 8
    src = rng.uniform(-h, h, 40)
    q = rng.standard_normal(40)
    mp = field(mp_nodes(pl.cheb, h), src, q)
    exact = field(loc_nodes(pl.cheb, h), src - 4 * h, q)
    assert rel(pl.topflp @ mp, exact) < 1e-9


@pytest.mark.parametrize("level", [3, 4, 7])
@pytest.mark.parametrize("op,side", [(6, +1), (7, -1)])
def test_l2l_shifts_local_to_child(plan16, level, op, side):
    pl = plan16
    rng = np.random.default_rng(level * 3 + op)
    hc = math.pi / 2 ** (level + 1)  # child at `level`
    hp = 2 * hc
    # far sources of the parent: [3 hp, pi - 3 hp] (the kernel is pi-periodic)
    src = rng.uniform(3.2 * hp, math.pi - 3.2 * hp, 40)
    q = rng.standard_normal(40)
    parent = field(loc_nodes(pl.cheb, hp), src, q)
    exact = field(loc_nodes(pl.cheb, hc) + side * hc, src, q)
    mat = pl.ops[level - 3, op]
    assert rel(mat @ parent, exact) < 1e-9


def test_l2t_rows_evaluate_local_expansion(plan16):
    pl = plan16
    rng = np.random.default_rng(11)
    a = math.pi / (2 * pl.s)
    src = rng.uniform(5 * a, 40 * a, 60) * rng.choice([-1, 1], 60)
    q = rng.standard_normal(60)
    loc = field(loc_nodes(pl.cheb, a), src, q)
    for sc in range(1, pl.P):
        j = np.arange(1, pl.b + 1)
        ffr = ((2 * j - 1) - 2 * sc / pl.P) / pl.b - 1
        exact = field(ffr * a, src, q)
        assert rel(pl.evalm[sc - 1] @ loc * pl.M, exact) < 1e-9
    for e in range(pl.P // 2):
        sc = pl.P // 2 + e
        ffr = ((2 * pl.b + 1) - 2 * sc / pl.P) / pl.b - 1
        exact = field(np.array([ffr * a]), src, q)
        assert rel(pl.evalmh[e] @ loc * pl.M, exact) < 1e-9


def test_combine_factors(plan16):
    pl = plan16
    w = np.exp(-2j * np.pi / pl.P)
    for j in range(pl.P):
        for sc in range(1, pl.P):
            expect = 1j * (1 - w ** sc) / 2 * w ** (sc * j)
            assert abs(pl.fo[j, sc - 1] - expect) < 1e-15


def test_brute_force_cot_sum_is_the_fractional_shift(orc):
    """psi_sc = ((C_sc x) * 2 / (1 - w_P^sc) - mean) / i with
    C_sc = F^-1 diag(w_N^{sc k}) F: pins sign and offset of the cot kernel."""
    M, P = 256, 8
    x = orc.normal(M * P, 3)
    xs = x.reshape(M, P)
    for sc in (1, 3, 4, 7):
        shifted = np.fft.ifft(np.fft.fft(xs[:, sc]) * np.exp(-2j * np.pi * sc * np.arange(M) / (M * P)))
        w = np.exp(-2j * np.pi * sc / P)
        psi = (shifted * 2 / (1 - w) - xs[:, sc].mean()) / 1j
        assert rel(orc.cot_direct(x, M, P, sc), psi) < 1e-13


@pytest.mark.parametrize("N,P,terms,b", [(1 << 13, 4, 16, 32), (1 << 13, 8, 12, 16), (1 << 12, 2, 20, 8)])
def test_fmm_section_sums_match_brute_force(orc, N, P, terms, b):
    pl = orc.OraclePlan(N, P, terms, b)
    x = orc.normal(N, 17)
    phi = pl.upward(x)
    psi = pl.downward(phi)
    ps = pl.evaluate(x, psi).reshape(pl.M, P)
    tol = {12: 1e-9, 16: 1e-12, 20: 1e-13}[terms]
    for sc in range(1, P):
        exact = orc.cot_direct(x, pl.M, P, sc)
        assert rel(ps[:, sc], exact) < tol, sc


def test_fft64_matches_numpy_golden(orc):
    g = np.load(os.path.join(GOLDEN, "fft_numpy.npz"))
    for n in (8, 16, 64, 256, 1024, 4096):
        X = orc.fft64(g[f"x{n}"])
        assert rel(X, g[f"X{n}"]) < 1e-15


def test_fft64_matches_long_double_dft(orc):
    x = orc.normal(512, 9)
    assert rel(orc.fft64(x), orc.dft_direct(x)) < 1e-15
    assert rel(orc.fft64(x, +1), orc.dft_direct(x, +1)) < 1e-15


def test_fmmfft_matches_fft_at_cpu_config(orc):
    # N=2^16, P=4, terms 16, b=64, real/imag uniform[-1,1], seed 1
    pl = orc.OraclePlan(1 << 16, 4, 16, 64)
    x = orc.uniform(1 << 16, 1)
    X = pl.fmmfft(x)
    Y = orc.fft64(x)
    assert orc.rel_l2(X, Y) < 1e-13
    assert orc.max_abs(X, Y) / np.abs(Y).max() < 1e-13


def _reference_rows(max_n):
    with open(os.path.join(GOLDEN, "reference_max_error.csv")) as f:
        rows = list(csv.DictReader(f))
    return [(int(r["N"]), int(r["P"]), int(r["B"]), int(r["T"]), r["max_rel_err_pct"])
            for r in rows if int(r["N"]) <= max_n]


@pytest.mark.parametrize("N,P,B,T,pct", _reference_rows(1 << 21))
def test_reproduces_reference_recorded_max_error(orc, N, P, B, T, pct):
    """The reference's fmmfft_error.c printed max_i |x_i - y_i|/|y_i| * 100 over
    rank 0's M outputs for the ramp x_n = n+1 (src/fmmfft_error.c:1097-1108,
    1267); the oracle must print the same '%lf'."""
    pl = orc.OraclePlan(N, P, T, B)
    x = orc.ramp(N)
    X = pl.fmmfft(x)
    Y = orc.fft64(x)
    M = N // P
    assert "%f" % orc.max_rel_pct(X[:M], Y[:M]) == pct


def test_reference_error_growth_at_larger_n(orc):
    for N, P, pct in [(1 << 22, 2, "0.000024"), (1 << 22, 8, "0.000007")]:
        pl = orc.OraclePlan(N, P, 15, 32)
        x = orc.ramp(N)
        X = pl.fmmfft(x)
        Y = orc.fft64(x)
        assert "%f" % orc.max_rel_pct(X[: N // P], Y[: N // P]) == pct


def test_dft_invariants(orc):
    N, P = 1 << 12, 4
    pl = orc.OraclePlan(N, P, 16, 16)
    a, b = orc.normal(N, 1), orc.normal(N, 2)
    Fa, Fb = pl.fmmfft(a), pl.fmmfft(b)
    # linearity
    assert rel(pl.fmmfft(2.5 * a - 1j * b), 2.5 * Fa - 1j * Fb) < 1e-13
    # delta at 0 -> all ones; delta at n0 -> pure phase ramp
    d = np.zeros(N, complex)
    d[0] = 1
    assert np.max(np.abs(pl.fmmfft(d) - 1)) < 1e-13
    d[:] = 0
    d[37] = 1
    k = np.arange(N)
    assert np.max(np.abs(pl.fmmfft(d) - np.exp(-2j * np.pi * 37 * k / N))) < 1e-12
    # pure tone -> delta of height N
    tone = np.exp(2j * np.pi * 101 * k / N)
    T = pl.fmmfft(tone)
    assert abs(T[101] - N) < 1e-9 and np.max(np.abs(np.delete(T, 101))) < 1e-9
    # Parseval
    assert abs(np.vdot(Fa, Fa).real / N - np.vdot(a, a).real) / np.vdot(a, a).real < 1e-13
    # circular shift -> phase ramp
    Fs = pl.fmmfft(np.roll(a, 5))
    assert rel(Fs, Fa * np.exp(-2j * np.pi * 5 * k / N)) < 1e-12
    # unnormalised inverse
    assert rel(pl.fmmfft(Fa, -1) / N, a) < 1e-13


def test_error_metrics_match_numpy(orc):
    """orc_rel_l2, orc_max_abs and orc_max_rel_pct against their numpy
    definitions (the error metrics every parity test reports), including a
    single planted outlier that only the max metrics see."""
    x = orc.normal(4096, 31)
    y = x + 1e-7 * orc.normal(4096, 32)
    y[123] += 0.5 - 0.25j
    d = x - y
    assert abs(orc.rel_l2(x, y) - np.linalg.norm(d) / np.linalg.norm(y)) < 1e-15
    assert orc.max_abs(x, y) == pytest.approx(np.max(np.abs(d)), rel=1e-15)
    assert orc.max_rel_pct(x, y) == pytest.approx(np.max(np.abs(d) / np.abs(y)) * 100, rel=1e-15)
    # the operands are not interchangeable: the reference value is the denominator
    assert orc.rel_l2(x, 2 * x) == pytest.approx(0.5, rel=1e-15)
    assert orc.rel_l2(x, x) == 0.0 and orc.max_abs(x, x) == 0.0
