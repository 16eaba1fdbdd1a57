"""Host code under sanitizers: the whole library built from source with
AddressSanitizer + UndefinedBehaviorSanitizer, and with ThreadSanitizer, host
passes only (tests/cpp/Makefile), driving the multi-threaded parts --
structure parsing (16 threads, damaged inputs), the TSV writer, the generator
and the serving loop's queue, admission and tickets over a host executor with
four client threads (tests/cpp/sanitize_host.cpp).  The GPU test runs the same
driver's device section (default context from several threads, a device
server with concurrent submit / release / wait) under AddressSanitizer."""
import os
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "bin")
REPORTS = ("ERROR: AddressSanitizer", "WARNING: ThreadSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def build():
    r = subprocess.run(["make", "-s", "-C", CPP, "-j8"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def run(name, *args, env=None):
    exe = os.path.join(BIN, name)
    r = subprocess.run([exe, f"--golden={GOLDEN}", *args], capture_output=True, text=True, timeout=600,
                       cwd=ROOT, env={**os.environ, **(env or {})})
    out = r.stdout + r.stderr
    print(out[-2000:])
    return r.returncode, out


@pytest.mark.parametrize("name", ["sanitize_asan", "sanitize_tsan"])
def test_host_code_under_sanitizers(name):
    build()
    rc, out = run(name)
    assert rc == 0, out[-4000:]
    assert "all sections passed" in out
    assert not any(s in out for s in REPORTS), out[-4000:]
    for sec in ("ingest", "tsv", "generate", "queue"):
        assert f"ok {sec}" in out


@pytest.mark.gpu
def test_device_server_under_asan():
    exe = os.path.join(BIN, "sanitize_asan")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/bin/sanitize_asan is built by __graft_entry__.build() on the build host")
    # the HIP runtime keeps its allocations until exit: no leak report for it
    rc, out = run("sanitize_asan", "--gpu", env={"ASAN_OPTIONS": "detect_leaks=0"})
    assert rc == 0, out[-4000:]
    assert "ok gpu" in out and "all sections passed" in out
    assert not any(s in out for s in REPORTS), out[-4000:]
