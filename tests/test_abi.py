"""The C-ABI library: it loads, exports every symbol include/vc1me.h declares,
and refuses to run without a device (no silent CPU fallback)."""
import ctypes as C
import os
import re

import pytest

from libvc1_amd import abi, backend

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "vc1me.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(vc1me_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = backend.load_library()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert set(names) == set(backend.EXPORTS)


def test_abi_version_and_defaults():
    lib = backend.load_library()
    assert lib.vc1me_abi_version() == 1
    p = abi.Params()
    lib.vc1me_params_default(C.byref(p), 1920, 1080)
    ref = abi.default_params(1920, 1080)
    for name, _ in abi.Params._fields_:
        assert getattr(p, name) == getattr(ref, name), name


def test_struct_layouts_match_header():
    # sizes of the POD structs as the C compiler lays them out
    assert C.sizeof(abi.Params) == 16 * 4 + 2 * 8
    assert C.sizeof(abi.Candidate) == 8
    assert C.sizeof(abi.PResult) == 9 * 8 + 2 * 8 + 5 * 8 + 2 * 4 + 8


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a device is present")
def test_create_fails_loudly_without_device():
    with pytest.raises(backend.BackendError) as e:
        backend.Backend(0)
    assert e.value.code == abi.VC1ME_EDEVICE


def test_adapter_compiles_against_the_header():
    """The libvc1 adapter (libvc1_amd/integration/hip_compute.h) and its test
    program compile as C++17 against include/vc1me.h alone."""
    import shutil
    import subprocess
    from libvc1_amd import build
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'libvc1_amd', 'integration')}",
                        f"-I{os.path.join(ROOT, 'tests', 'adapter')}", build.ADAPTER_SRC],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_relaxation_worklist_entries_round_trip(tmp_path):
    """The relaxation's worklist entry (job in the top 8 bits, macroblock in
    the low 24, stored in an int32 list) unpacks to the same job and
    macroblock for every job a step may hold (up to 254: the top bit set from
    job 128 on) and every macroblock index below 2^24; the XCD bands that
    order the whole-frame kernels' work and segment the lists tile the grid,
    and every item maps to the band that holds it (host build of the device
    helpers)."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("no host C++ compiler or HIP headers")
    src = tmp_path / "entry.cpp"
    src.write_text(r'''
#include "vc1me_internal.h"
#include <cstdio>
int main() {
    int bad = 0;
    for (int job : {0, 1, 127, 128, 200, 254})
        for (int mb : {0, 1, 8159, 32399, (1 << 24) - 1}) {
            const int32_t stored = static_cast<int32_t>(vc1me::relax_entry(job, mb));
            const uint32_t back = static_cast<uint32_t>(stored);
            if (vc1me::relax_entry_job(back) != job || vc1me::relax_entry_pos(back) != mb) ++bad;
        }
    // XCD bands of a grid: every item lies in its own band, and the bands
    // tile [0, total) in order
    for (int total : {1, 5, 7, 8, 9, 15, 16, 17, 130560, 255 * 8160 + 3, 255 * 32400})
        for (int x = 0, next = 0; x < vc1me::kXcds; ++x) {
            const int start = vc1me::xcd_band_start(total, x), size = vc1me::xcd_band_size(total, x);
            if (start != next) ++bad;
            next = start + size;
            if (x == vc1me::kXcds - 1 && next != total) ++bad;
            for (int item : {start, start + size / 2, start + size - 1})
                if (size > 0 && vc1me::xcd_band_of(total, item) != x) ++bad;
        }
    for (int total : {1, 7, 8, 9, 23, 1000})
        for (int item = 0; item < total; ++item) {
            const int x = vc1me::xcd_band_of(total, item);
            if (x < 0 || x >= vc1me::kXcds || item < vc1me::xcd_band_start(total, x) ||
                item >= vc1me::xcd_band_start(total, x) + vc1me::xcd_band_size(total, x))
                ++bad;
        }
    std::printf("%d\n", bad);
    return bad != 0;
}
''')
    exe = tmp_path / "entry"
    r = subprocess.run([cxx, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        f"-I{os.path.join(ROOT, 'libvc1_amd', 'csrc')}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr
