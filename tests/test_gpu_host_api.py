"""The C++ host API (saq_amd/host/saqlib_gpu.hpp, the saqlib::IVF surface over the C ABI)
end to end: saq_amd/host/ivf_run constructs an index with saqlib_gpu::IVF::construct and
searches it with search()/search_batch(); IDs and distances must equal the oracle's search
over the oracle-encoded index of the same data, bit for bit.  IVF::save / IVF::load: the
file it writes must follow the reference's format (tests/index_file.py) byte for byte against
the oracle-encoded arrays, and a file written to that format from the oracle's arrays must
load and search like the oracle."""
import os
import subprocess

import ctypes as C

import numpy as np
import pytest

import index_file as IF
import saq_fixtures as F
from oracle_lib import Sizes, oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "saq_amd", "host", "ivf_run")


def write_input(path, case, topk, nprobe, avg_bits):
    with open(path, "wb") as f:
        n, d = case.base.shape
        np.array([n, d, case.K, case.queries.shape[0], topk, nprobe, len(case.plan), case.rd_lmt],
                 np.int64).tofile(f)
        np.array(case.plan, np.int64).tofile(f)
        np.array([case.eps, avg_bits], np.float32).tofile(f)
        for a in (case.base, case.centroids, case.variance[:d], case.queries):
            np.ascontiguousarray(a, np.float32).tofile(f)
        has_rot = case.rotations[0] is not None
        np.array([1 if has_rot else 0], np.int32).tofile(f)
        if has_rot:
            for r in case.rotations:
                np.ascontiguousarray(r, np.float32).tofile(f)
        np.ascontiguousarray(case.cluster_ids, np.uint32).tofile(f)


@pytest.mark.parametrize("kw", [dict(n=2500, d=256, K=10, avg_bits=4.0, seed=41, tau=32),
                                dict(n=1800, d=200, K=7, plan=[(128, 6), (128, 0)], seed=43),
                                dict(n=1500, d=128, K=6, plan=[(128, 4)], seed=45, rotate=False)])
def test_ivf_run_matches_oracle(tmp_path, kw):
    avg_bits = kw.get("avg_bits", 4.0)
    case = F.make_case(nq=16, **kw)
    topk, nprobe = 10, 4
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(inp, case, topk, nprobe, avg_bits)
    print(f"ivf_run case {kw}", flush=True)
    ids, dists, fails = run_tool(inp, out, case, topk)
    probes = F.orc_probes(case, nprobe)
    ids_o, d_o, _ = F.orc_search(case, probes, topk)
    assert np.array_equal(ids, ids_o)
    assert np.array_equal(dists.view(np.uint32), d_o.view(np.uint32))
    assert fails == 0


def run_tool(inp, out, case, topk, *extra):
    r = subprocess.run([TOOL, inp, out, *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nq = case.queries.shape[0]
    raw = open(out, "rb").read()
    ids = np.frombuffer(raw, np.uint32, nq * topk).reshape(nq, topk)
    dists = np.frombuffer(raw, np.float32, nq * topk, nq * topk * 4).reshape(nq, topk)
    fails = np.frombuffer(raw, np.int64, 1, nq * topk * 8 + nq * 24)[0]
    return ids, dists, fails


def cluster_sizes_of(plan, n, compact):
    s = Sizes()
    oracle().orc_cluster_sizes(C.byref(F.make_plan([d for d, _ in plan], [b for _, b in plan], 1, compact)), n,
                               C.byref(s))
    return s.short_factor_floats, s.short_code_bytes, s.long_factor_floats


FILE_CASES = [dict(n=1800, d=256, K=7, plan=[(128, 5), (64, 9), (64, 0)], seed=81, tau=32),
              dict(n=1500, d=128, K=6, plan=[(128, 4)], seed=83),
              dict(n=1300, d=200, K=5, plan=[(128, 3), (128, 1)], seed=85, rotate=False)]


@pytest.mark.parametrize("kw", FILE_CASES)
def test_saved_index_follows_the_reference_format(tmp_path, kw):
    case = F.make_case(nq=8, **kw)
    inp, out, path = str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(tmp_path / "x.index")
    write_input(inp, case, 10, 3, 4.0)
    print(f"ivf_run --save {kw}", flush=True)
    ids_a, d_a, _ = run_tool(inp, out, case, 10, "--save", path)
    f = IF.read_index(path, cluster_sizes_of)
    assert (f["num_data"], f["num_dim"], f["K"]) == (case.base.shape[0], case.num_dim, case.K)
    assert np.array_equal(f["centroids"].view(np.uint32), case.centroids.view(np.uint32))
    assert f["plan"] == [tuple(p) for p in case.plan]
    assert f["saq_num_dim"] == case.num_dim and f["cfg"]["single"]["use_fastscan"] == 1
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance[:case.num_dim]
    assert np.array_equal(f["variance"], vpad)
    for s, P in enumerate(case.rotations):
        if P is None:
            assert f["rotations"][s] is None
        else:
            assert np.array_equal(f["rotations"][s].view(np.uint32), np.asarray(P, np.float32).view(np.uint32))
    assert np.array_equal(f["sizes"], case.sizes())
    for c in range(case.K):
        got, ref = f["clusters"][c], case.clusters[c]
        n = int(ref["num_vec"])
        assert np.array_equal(got["ids"], ref["ids"][:n])
        for k in ("short_factors", "short_code", "long_factors", "centroid_rot"):
            assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), f"cluster {c} {k}"
        lb = len(got["long_code"])
        assert np.array_equal(got["long_code"], ref["long_code"][:lb]), f"cluster {c} long_code"
    # and the file loads back to the same search results
    ids_b, d_b, _ = run_tool(inp, out, case, 10, "--load", path)
    assert np.array_equal(ids_a, ids_b) and np.array_equal(d_a.view(np.uint32), d_b.view(np.uint32))


@pytest.mark.parametrize("kw", FILE_CASES)
def test_reference_format_file_loads_and_searches(tmp_path, kw):
    case = F.make_case(nq=8, **kw)
    inp, out, path = str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(tmp_path / "ref.index")
    write_input(inp, case, 10, 3, 4.0)
    vpad = np.zeros(case.D_pad, np.float32)
    vpad[:case.num_dim] = case.variance[:case.num_dim]
    rots = None if case.rotations[0] is None else case.rotations
    cfg = {"avg_bits": 4.0, "enable_segmentation": 1, "use_compact_layout": 0,
           "single": {"random_rotation": 1 if rots is not None else 0, "use_fastscan": 1, "caq_adj_rd_lmt": 6,
                      "caq_adj_eps": 1e-8}}
    IF.write_index(path, case.num_dim, case.plan, case.centroids, vpad, rots, case.clusters, cfg)
    print(f"ivf_run --load {kw}", flush=True)
    ids, dists, _ = run_tool(inp, out, case, 10, "--load", path)
    ids_o, d_o, _ = F.orc_search(case, F.orc_probes(case, 3), 10)
    assert np.array_equal(ids, ids_o)
    assert np.array_equal(dists.view(np.uint32), d_o.view(np.uint32))
