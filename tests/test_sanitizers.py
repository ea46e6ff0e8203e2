"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's C restatement (its own CPU tests run against the sanitizer build)
and the engine's host-only C++ (cross-section tables, eduPIC's run report,
the ABI layout) called on ordinary and edge-case inputs.  gcc's sanitizer
runtimes are preloaded into a child Python; any report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


ASAN, UBSAN = _runtime("libasan.so"), _runtime("libubsan.so")
pytestmark = pytest.mark.skipif(not (ASAN and UBSAN), reason="gcc's sanitizer runtimes are not installed")


def _env(**extra):
    env = dict(os.environ)
    env.update({"LD_PRELOAD": f"{ASAN}:{UBSAN}", "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "PYTHONPATH": ROOT}, **extra)
    return env


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    return os.path.join(ORACLE, "_san")


def _clean(r):
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]


def test_oracle_tests_under_sanitizers(built):
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "tests/test_oracle_rng.py", "tests/test_oracle_physics.py", "tests/test_oracle_eepf.py",
                        "tests/test_oracle_adjoint.py"],
                       cwd=ROOT, env=_env(PIC_ORACLE_LIB=os.path.join(built, "libpic_oracle.so")),
                       capture_output=True, text=True, timeout=900)
    _clean(r)
    assert " passed" in r.stdout


HOST_SCRIPT = r'''
import ctypes as C, math, sys
import numpy as np
from cudapic_amd import _abi
lib = C.CDLL(sys.argv[1])
lib.pic_cross_sections.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
cs = np.zeros(5 * 1_000_000, np.float32); se = np.zeros(1_000_000, np.float32); si = np.zeros_like(se)
for n_gas in (2.0694e21, 1e19):
    assert lib.pic_cross_sections(n_gas, cs.ctypes.data, se.ctypes.data, si.ctypes.data) == 0
    assert np.isfinite(cs).all() and np.isfinite(se).all() and np.isfinite(si).all()
assert lib.pic_cross_sections(0.0, cs.ctypes.data, se.ctypes.data, si.ctypes.data) == _abi.PIC_ERR_ARG
assert lib.pic_cross_sections(1e21, None, None, None) == _abi.PIC_ERR_ARG
from cudapic_amd.params import params
p = params()
lib.pic_cross_sections(p.gas_density, cs.ctypes.data, se.ctypes.data, si.ctypes.data)
lib.pic_report.argtypes = [C.POINTER(_abi.PicParams), C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                           C.c_ulonglong, C.c_ulonglong, C.c_uint32, C.c_uint32, C.POINTER(_abi.PicReport)]
rep = _abi.PicReport()
for dens, mean_e, n_e in ((7.4e15, 3.2, 100000), (0.0, float("nan"), 0), (7.4e15, float("nan"), 5), (1e30, 1e9, 1)):
    assert lib.pic_report(C.byref(p), se.ctypes.data, si.ctypes.data, dens, mean_e, 7.4e-6, 10, 10, n_e, 1,
                          C.byref(rep)) == 0
out = (C.c_uint64 * 6)()
assert lib.pic_abi_layout(out) == 0 and out[0] == C.sizeof(_abi.PicParams)
print("host ok")
'''


def test_host_code_under_sanitizers(built):
    r = subprocess.run([sys.executable, "-c", HOST_SCRIPT, os.path.join(built, "libpic_host.so")], cwd=ROOT,
                       env=_env(), capture_output=True, text=True, timeout=300)
    _clean(r)
    assert "host ok" in r.stdout
