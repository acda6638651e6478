"""The C++ host API (saq_amd/host/saqlib_gpu.hpp, the saqlib::IVF surface over the C ABI)
end to end: saq_amd/host/ivf_run constructs an index with saqlib_gpu::IVF::construct and
searches it with search()/search_batch(); IDs and distances must equal the oracle's search
over the oracle-encoded index of the same data, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import saq_fixtures as F

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
    r = subprocess.run([TOOL, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nq = case.queries.shape[0]
    raw = open(out, "rb").read()
    ids = np.frombuffer(raw, np.uint32, nq * topk).reshape(nq, topk)
    dists = np.frombuffer(raw, np.float32, nq * topk, nq * topk * 4).reshape(nq, topk)
    fails = np.frombuffer(raw, np.int64, 1, nq * topk * 8 + nq * 24)[0]
    probes = F.orc_probes(case, nprobe)
    ids_o, d_o, _ = F.orc_search(case, probes, topk)
    assert np.array_equal(ids, ids_o)
    assert np.array_equal(dists.view(np.uint32), d_o.view(np.uint32))
    assert fails == 0
