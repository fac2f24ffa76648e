"""The libvc1 adapter (libvc1_amd/integration/hip_compute.h) compiled as C++
against include/vc1me.h and run the way the encoder's compute slot runs it
(tests/adapter/adapter_test.cpp): the scorer against the CPU law, the
whole-frame analysis and its seeded form against the C ABI, one context
shared by two host threads, and the fault-injection contract (VC1ME_FAIL_AT:
the scorer reports the failure, the caller re-scores on the CPU, results do
not change)."""
import os
import subprocess

import pytest

from libvc1_amd import build

pytestmark = pytest.mark.gpu


def _run(mode, env=None):
    if not os.path.exists(build.ADAPTER_BIN):
        pytest.fail(f"{build.ADAPTER_BIN} is not built (python -m libvc1_amd.build)")
    print(f"[case] adapter {mode}", flush=True)
    e = dict(os.environ)
    e.pop("VC1ME_FAIL_AT", None)
    e.update(env or {})
    r = subprocess.run([build.ADAPTER_BIN, mode], capture_output=True, text=True, timeout=100, env=e)
    print(r.stdout, r.stderr, flush=True)
    return r


def test_adapter_full():
    r = _run("full")
    assert r.returncode == 0 and "ADAPTER OK" in r.stdout, r.stdout + r.stderr


def test_adapter_injected_failure_rescored_on_cpu():
    r = _run("inject", {"VC1ME_FAIL_AT": "1"})
    assert r.returncode == 0 and "ADAPTER OK" in r.stdout, r.stdout + r.stderr
