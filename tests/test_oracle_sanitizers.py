"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (the
race/sanitizer coverage of the host-side restatement): oracle/sanitize_check.c
runs LJ with the CSR, half (newton on) and 2D lists, lj/cut/idial, SNAP at
twojmax 8 and 12, a dump round trip and the twojmax-12 Clebsch-Gordan table,
with two OpenMP threads.  A sanitizer report aborts the program; the test
needs a clean exit and every run reported ok."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path):
    cc = shutil.which(os.environ.get("CC", "gcc"))
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "sanitize_check")
    build = subprocess.run(
        [cc, "-O1", "-g", "-std=gnu11", "-fopenmp", "-ffp-contract=off",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
         "-I", ORACLE, os.path.join(ORACLE, "oracle.c"), os.path.join(ORACLE, "oracle_snap.c"),
         os.path.join(ORACLE, "sanitize_check.c"), "-lm", "-o", exe],
        capture_output=True, text=True)
    if build.returncode != 0 and "asan" in (build.stderr or "").lower():
        pytest.skip("the compiler has no sanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    res = subprocess.run([exe, str(tmp_path / "check.dump")], capture_output=True, text=True,
                         env=env, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, (res.stdout + res.stderr)[-4000:]
    assert "all runs clean" in res.stdout
    assert res.stdout.count(" ok\n") == 6
    assert "runtime error" not in res.stderr and "ERROR: AddressSanitizer" not in res.stderr
