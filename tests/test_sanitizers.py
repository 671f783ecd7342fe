"""CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

* The oracle (oracle/ckks_oracle.c, `make -C oracle sanitize`) runs its own
  unit tests and the CPU bootstrap tests in a child Python process with
  libasan preloaded and FIDES_ORACLE_SANITIZE=1.
* The host C++ library's sources (fideslib.cpp, serialize.cpp,
  bootstrap.cpp) are built into tests/cpp/api_check_san
  (`make -C tests/cpp san`), which runs the CPU serialization mode:
  round trips, a file of another type, a truncated file.

Any sanitizer report fails the child (UBSan with -fno-sanitize-recover,
ASan by default); leak checking is off (the Python interpreter keeps its
allocations until exit).  CPU only: nothing here touches a GPU.
"""
import os
import subprocess
import sys

import pytest

import test_serialize as TS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV_SAN = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _reports(text):
    return [ln for ln in text.splitlines() if "ERROR: AddressSanitizer" in ln or "runtime error:" in ln]


@pytest.fixture(scope="module")
def asan():
    if os.environ.get("LD_PRELOAD"):
        pytest.skip("a library is already preloaded here; the child would have to replace it")
    lib = _runtime("libasan.so")
    if lib is None:
        pytest.skip("gcc has no AddressSanitizer runtime")
    return lib


# a planted mistake on the CPU: an NTT of N = 4096 over an 8-element array
PLANTED = """
import numpy as np
from oracle import oracle as O
a = np.zeros(8, np.uint64)
psi = np.ones(4096, np.uint64)
O.lib().or_ntt(a.ctypes.data_as(O.u64p), 12, 12289, psi.ctypes.data_as(O.u64p))
print("no report")
"""


def test_sanitizer_reports_a_planted_overflow(asan):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=300)
    env = dict(os.environ, **ENV_SAN, FIDES_ORACLE_SANITIZE="1", LD_PRELOAD=asan)
    r = subprocess.run([sys.executable, "-c", PLANTED], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "no report" not in r.stdout
    assert "AddressSanitizer: heap-buffer-overflow" in r.stderr, r.stderr[-2000:]


def test_oracle_under_asan_ubsan(asan):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=300)
    env = dict(os.environ, **ENV_SAN, FIDES_ORACLE_SANITIZE="1", LD_PRELOAD=asan, OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        "tests/test_oracle.py", "tests/test_bootstrap_host.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert not _reports(out), "\n".join(_reports(out)[:20])
    assert r.returncode == 0, out[-4000:]
    assert " passed" in r.stdout


def test_host_serialization_under_asan_ubsan(asan, tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "san"], check=True, timeout=600)
    TS._write_inputs(str(tmp_path), 5)
    env = dict(os.environ, **ENV_SAN)
    r = subprocess.run([os.path.join(ROOT, "tests", "cpp", "api_check_san"), str(tmp_path), "serial-host"],
                       env=env, capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert not _reports(out), "\n".join(_reports(out)[:20])
    assert r.returncode == 0, out[-4000:]
    assert "serial-host: ok" in r.stdout
