"""Host AddressSanitizer + LeakSanitizer + UndefinedBehaviorSanitizer run of
the CPU oracle (test infrastructure).  tests/sanitize_oracle.c drives every
exported oracle routine over small shapes; it is compiled here with ROCm's
clang and its compiler-rt sanitizers, so an out-of-bounds access, a leak or
an undefined operation in the checker fails this test instead of silently
bending the parity results.  (gcc 11's AddressSanitizer does not check
`double complex` loads, which is how the oracle reads its arrays, so clang is
the compiler here.)"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang not installed")
def test_oracle_under_address_leak_and_ub_sanitizers(tmp_path):
    exe = tmp_path / "sanitize_oracle"
    cmd = [CLANG, "-O1", "-g", "-std=c99", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(REPO, "oracle"), "-o", str(exe),
           os.path.join(REPO, "tests", "sanitize_oracle.c"),
           os.path.join(REPO, "oracle", "fmmfft_oracle.c"), "-lm", "-latomic"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    # the OpenMP runtime keeps its thread-pool descriptors until exit
    supp = tmp_path / "lsan.supp"
    supp.write_text("leak:libomp.so\n")
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:exitcode=23",
               LSAN_OPTIONS=f"suppressions={supp}",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitize ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
