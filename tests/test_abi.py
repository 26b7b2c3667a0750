"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/fmmfft.h declares, and validates plan parameters (no GPU needed:
fmmfft_plan_sizes makes no HIP call)."""
import ctypes
import os
import re

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "fmmfft.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(fmmfft_\w+)\(", text, re.M)))


def test_header_declares_the_entry_points(fmm):
    assert set(declared_symbols()) == set(fmm.EXPORTS)


def test_library_exports_every_declared_symbol(fmm):
    L = fmm.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_plan_sizes_accepts_the_benchmark_configs(fmm):
    for n, p, terms, leaf, dt in [
        (1 << 16, 4, 16, 64, fmm.C128),
        (1 << 24, 32, 20, 64, fmm.C128),
        (1 << 26, 64, 16, 64, fmm.C64),
        (1 << 20, 16, 12, 32, fmm.C64),
        (1 << 12, 8, 12, 32, fmm.C64),
    ]:
        tb, wb = fmm.plan_sizes(n, p, terms, leaf, dt)
        assert tb > 0 and wb > 0


def test_workspace_scales_with_precision(fmm):
    _, w64 = fmm.plan_sizes(1 << 20, 16, 12, 32, fmm.C64)
    _, w128 = fmm.plan_sizes(1 << 20, 16, 12, 32, fmm.C128)
    assert 1.8 < w128 / w64 < 2.1


@pytest.mark.parametrize(
    "n,p,terms,leaf,dtype,status",
    [
        (1 << 16, 3, 16, 64, 1, 1),      # P not a power of two
        (1 << 16, 1, 16, 64, 1, 1),      # P < 2
        (1 << 16, 128, 16, 4, 1, 1),     # P > 64
        (1 << 16 | 4, 4, 16, 64, 1, 1),  # N % P != 0 / leaf mismatch
        (1 << 16, 4, 1, 64, 1, 1),       # terms < 2
        (1 << 16, 4, 33, 64, 1, 1),      # terms > 32
        (1 << 16, 4, 16, 48, 1, 1),      # s = M / leaf not a power of two
        (1 << 10, 4, 16, 64, 1, 1),      # s < 8 (tree depth < 3)
        (1 << 16, 4, 16, 64, 7, 1),      # bad dtype
        (1 << 16, 4, 16, 6, 1, 1),       # M % leaf != 0
        (1 << 14, 4, 16, 2, 1, 4),       # leaf not a multiple of 4
        (1 << 31, 2, 16, 64, 1, 4),      # N > 2^30
        (3 << 20, 2, 16, 64, 1, 1),      # M = 3 * 2^19: s not a power of two
    ],
)
def test_plan_sizes_rejects_invalid(fmm, n, p, terms, leaf, dtype, status):
    with pytest.raises(fmm.FmmFftError) as e:
        fmm.plan_sizes(n, p, terms, leaf, dtype)
    assert e.value.status == status


def test_plan_sizes_accepts_the_recorded_large_shapes(fmm):
    """The reference's recorded runs go to M = 2^26 (N = 2^27, P = 2) and
    N = 2^29 (P = 4), b = 32, T = 15 (three-pass local FFTs here)."""
    for n, p in [(1 << 27, 2), (1 << 29, 4), (1 << 28, 2), (1 << 30, 64)]:
        tb, wb = fmm.plan_sizes(n, p, 15, 32, fmm.C128)
        assert tb > 0 and wb >= n * 16


def test_strerror(fmm):
    L = fmm.lib()
    for s in range(5):
        assert L.fmmfft_strerror(s)
    assert L.fmmfft_strerror(99) == b"unknown status"
