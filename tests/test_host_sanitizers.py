"""The host code of libflgpu and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU (tests/sanitize: a host-sanitized
build, device code as usual): make_sparsity and the colouring, the
matrix-dump writer and reader (every truncation of a dump, and dumps with
corrupted headers, counts, row lengths and columns must fail cleanly or
read), and the oracle's assembly, solves and pressure path."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.mark.skipif(not os.path.exists(os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")) or shutil.which("make") is None,
                    reason="needs hipcc and make")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", SAN], check=True, capture_output=True, timeout=900)
    env = dict(os.environ, TMPDIR=str(tmp_path), ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([os.path.join(SAN, "_build", "host_check")], env=env, capture_output=True, text=True,
                         timeout=300)
    print(res.stdout[-2000:], res.stderr[-4000:])
    assert res.returncode == 0, res.stderr[-4000:]
    assert "host_check ok" in res.stdout
