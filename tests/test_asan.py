"""AddressSanitizer on the host code of the C ABI (tests/asan/capi_stress:
the library's sources with -fsanitize=address on the host side only, and a
driver that calls the C ABI from several threads). Without a device it
covers argument resolution and the creation error paths; on the GPU it also
runs concurrent process calls on one handle, a rolling clip, the serving
loop with client threads, the no-allocation entry point, WNNM and the fault
hooks. A heap error aborts the driver with a report, which fails the test."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "asan", "capi_stress")
SOURCES = [os.path.join(HERE, "asan", "capi_stress.cpp")] + [
    os.path.join(HERE, "..", "vapoursynth-nssfactory_amd", "csrc", f)
    for f in os.listdir(os.path.join(HERE, "..", "vapoursynth-nssfactory_amd", "csrc"))]


def _stale():
    if not os.path.exists(EXE):
        return True
    t = os.path.getmtime(EXE)
    return any(os.path.getmtime(s) > t for s in SOURCES)


def _run(env_extra):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", **env_extra)
    return subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)


def test_asan_driver_without_device():
    if _stale():
        subprocess.check_call([sys.executable, os.path.join(HERE, "asan", "build_asan.py")],
                              stdout=subprocess.DEVNULL)
    r = _run({})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_asan_driver_on_the_device():
    assert os.path.exists(EXE), "tests/asan/capi_stress is not built (python tests/asan/build_asan.py)"
    r = _run({})
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr
    assert "device paths done; 0 failure(s)" in r.stdout
