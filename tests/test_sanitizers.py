"""The CPU side under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (oracle/_san/) and the host build of the kernels' clipping, cut
topology and adjoint code (tests/native/_san/) run the oracle's golden tests
and the kernel-logic tests in a child Python process with libasan preloaded.
Any sanitizer report ends that process with an error (UBSan without
recovery; ASan leak checks off: the interpreter keeps its allocations). The
largest cases (the RSA records, 1ctf and the 120-ball adjoint) are left out
to keep the run short; the same code paths run on the smaller fixtures."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _libasan():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if out.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_and_kernel_logic_under_sanitizers():
    asan = _libasan()
    if asan is None:
        pytest.skip("libasan not found")
    for d in ("oracle", os.path.join("tests", "native")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, d), "san"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        "ASAN_OPTIONS": "detect_leaks=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "VLTR_ORACLE_LIB": os.path.join(ROOT, "oracle", "_san", "liboracle_vltr.so"),
        "VLTR_CLIP_HOST_LIB": os.path.join(ROOT, "tests", "native", "_san", "libclip_host.so"),
    })
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
           os.path.join(ROOT, "tests", "test_oracle_goldens.py"), os.path.join(ROOT, "tests", "test_host_kernel_logic.py"),
           "-k", "not rsa and not 1ctf and not 120"]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, cwd=ROOT, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
