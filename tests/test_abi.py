"""CPU-side checks of the product boundary: the gfx950 library loads, exports every
symbol include/saqgpu.h declares, validates arguments before touching the GPU,
and its host plan matches the oracle's restatement of SaqDataMaker."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import saq_amd
from oracle_lib import oracle
from saq_amd.plan import make_plan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "saqgpu.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(saqgpu_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = saq_amd.lib()
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n


def test_create_refuses_more_than_64_code_groups():
    """The fast scan flushes segment sums at 64-bit group masks: more than 4096 quantised dims
    are refused (SAQGPU_EUNSUPPORTED) before any device work."""
    L = saq_amd.lib()
    h = C.c_void_p()
    sizes = np.array([10], np.int64)
    for dims, bits in (([4160], [1]), ([2048, 2048, 64], [2, 3, 1]), ([8192], [4])):
        d = np.array(dims, np.int32)
        b = np.array(bits, np.int32)
        desc = saq_amd._Desc(10, int(d.sum()), 1, len(dims), 0, d.ctypes.data, b.ctypes.data, None, None, None,
                             sizes.ctypes.data, 0)
        assert L.saqgpu_index_create(C.byref(desc), C.byref(h)) == 5
        assert b"4096" in L.saqgpu_last_error()


def test_bad_arguments_fail_without_a_device():
    L = saq_amd.lib()
    h = C.c_void_p()
    assert L.saqgpu_index_create(None, C.byref(h)) == 1
    assert b"null" in L.saqgpu_last_error()
    dims = np.array([100], np.int32)
    bits = np.array([4], np.int32)
    sizes = np.array([10], np.int64)
    desc = saq_amd._Desc(10, 100, 1, 1, 0, dims.ctypes.data, bits.ctypes.data, None, None, None,
                         sizes.ctypes.data, 0)
    assert L.saqgpu_index_create(C.byref(desc), C.byref(h)) == 1  # segment dim not a multiple of 64
    bits[0] = 14
    dims[0] = 128
    assert L.saqgpu_index_create(C.byref(desc), C.byref(h)) == 1  # more than 13 bits
    nb = C.c_int64(0)
    assert L.saqgpu_single_layout(None, C.byref(nb), None) == 1
    assert L.saqgpu_single_quantize(None, None, 0, None, None, 0, 0, None, None) == 1
    assert L.saqgpu_single_estimate(None, None, 0, None, 0, 0, 4.0, 0, None, None, None, 0, None) == 1


@pytest.mark.parametrize("d,avg", [(128, 4.0), (256, 1.5), (960, 4.0), (960, 1.5), (1024, 2.0), (1536, 1.5)])
def test_plan_matches_oracle(d, avg):
    g = np.random.default_rng(d)
    var = (g.uniform(0.2, 2.0, d) / (1 + np.arange(d) / 32)).astype(np.float32)
    got = make_plan(var, d, avg)
    D_pad = (d + 63) // 64 * 64
    vp = np.zeros(D_pad, np.float32)
    vp[:d] = var
    dims = np.zeros(64, np.int32)
    bits = np.zeros(64, np.int32)
    n = oracle().orc_plan_dp(vp, D_pad, avg, dims, bits)
    assert got == [(int(dims[i]), int(bits[i])) for i in range(n)]
    assert sum(x for x, _ in got) == D_pad


def test_plan_equal_segmentation_and_disabled():
    var = np.ones(300, np.float32)
    assert make_plan(var, 300, 4.0, seg_eqseg=2) == [(192, 4), (128, 4)]
    assert make_plan(var, 300, 3.0, enable_segmentation=False) == [(320, 3)]


def test_host_tool_links_against_the_library():
    """saq_amd/host/ivf_run (the saqlib::IVF host API over the C ABI) is built and resolves
    libsaqgpu.so; without arguments it prints its usage and touches no device."""
    import os
    import subprocess
    tool = os.path.join(os.path.dirname(saq_amd.LIB_PATH), "host", "ivf_run")
    assert os.path.exists(tool)
    ldd = subprocess.run(["ldd", tool], capture_output=True, text=True).stdout
    assert "libsaqgpu.so" in ldd and "not found" not in ldd.split("libsaqgpu.so", 1)[1].split("\n", 1)[0]
    r = subprocess.run([tool], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
