"""Host sanitizers: the CPU suite with AddressSanitizer and
UndefinedBehaviorSanitizer builds of the oracle (oracle/oracle.c, gcmc.c) and
of libpmhip.so's host code (the C ABI's validation and set-up paths in
pmhip.hip, the pocket blocking in pm_pockets.cpp), loaded through PMHIP_LIB /
PM_ORACLE_LIB into a Python process that preloads the LLVM sanitizer runtime.

A canary library with a known heap overflow must be reported first, so a
clean run means the instrumentation is live.  Builds come from
tests/sanitize/Makefile (about a minute the first time, cached after)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_DIR = os.path.join(ROOT, "tests", "sanitize")
BUILD = os.path.join(SAN_DIR, "_build")
CLANG = "/opt/rocm/llvm/bin/clang"


def _runtime():
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_env():
    if not (os.path.exists(CLANG) and shutil.which("make") and os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("LLVM clang / hipcc not available")
    rt = _runtime()
    if rt is None:
        pytest.skip("no LLVM AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-C", SAN_DIR, "-j4"], check=True, timeout=900)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "PMHIP_LIB": os.path.join(BUILD, "libpmhip.so"),
        "PM_ORACLE_LIB": os.path.join(BUILD, "liboracle.so"),
    })
    return env


def test_canary_is_reported(san_env):
    code = f"import ctypes; ctypes.CDLL({os.path.join(BUILD, 'canary.so')!r}).canary_overflow(4)"
    r = subprocess.run([sys.executable, "-c", code], env=san_env, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]


def test_cpu_suite_clean_under_sanitizers(san_env):
    # the instrumented libraries are the ones the suite loads
    probe = ("import sys; sys.path.insert(0, %r); import __graft_entry__ as g; pm = g.load_package(); "
             "from oracle import oracle; oracle.lib(); pm._lib.load(); "
             "maps = open('/proc/self/maps').read(); "
             "print(%r in maps, %r in maps)" % (ROOT, san_env["PMHIP_LIB"], san_env["PM_ORACLE_LIB"]))
    r = subprocess.run([sys.executable, "-c", probe], env=san_env, capture_output=True, text=True, timeout=300,
                       cwd=ROOT)
    assert r.stdout.split() == ["True", "True"], r.stdout + r.stderr[-2000:]
    cmd = [sys.executable, "-m", "pytest", "tests", "-m", "not gpu", "-x", "-q", "-p", "no:cacheprovider",
           "--deselect", "tests/test_sanitizers.py::test_canary_is_reported",
           "--deselect", "tests/test_sanitizers.py::test_cpu_suite_clean_under_sanitizers"]
    r = subprocess.run(cmd, env=san_env, capture_output=True, text=True, timeout=1200, cwd=ROOT)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
