"""saqgpu_probe_noalloc (the entry point that allocates nothing) against saqgpu_probe.

Both give the reference's coarse-quantiser lists (FlatInitializer::centroids_distances +
partial_sort, index/initializer.hpp:52-76); the no-alloc form keeps its (query, centroid)
distances in the library's fixed per-GPU scratch array.  Two calls in a row, the second with
fewer queries, must each equal saqgpu_probe's lists exactly; sizes beyond the scratch are
refused before any device work.  Device buffers come from the library's own HIP runtime
(hip_mem), as in the other device-pointer tests."""
import numpy as np
import pytest

import saq_amd
import saq_fixtures as F
from hip_mem import DevArray, hip

pytestmark = pytest.mark.gpu


def test_noalloc_probe_equals_probe():
    case = F.make_case(n=6000, d=200, K=300, nq=48, avg_bits=4.0, seed=17, tau=32, encode=False)
    ix = F.gpu_index(case)  # the coarse quantiser needs only the centroids
    L = saq_amd.lib()
    for dist_type in (0, 1):
        for nq, nprobe in ((48, 37), (11, 9)):
            want = ix.probe(case.queries[:nq], nprobe, dist_type=dist_type)
            q = DevArray(case.queries[:nq])
            out = DevArray(np.full((nq, nprobe), 0xFFFFFFFF, np.uint32))
            saq_amd._check(L.saqgpu_probe_noalloc(ix.h, q.data_ptr(), nq, nprobe, dist_type, out.data_ptr(), None))
            assert hip().hipDeviceSynchronize() == 0
            got = out.numpy()
            q.free()
            out.free()
            assert np.array_equal(want, got), f"dist_type {dist_type} nq {nq}: lists differ"
    # more queries than the scratch holds: refused with a status before anything is launched
    big = saq_amd.SAQGPU_NOALLOC_SCRATCH_FLOATS // case.K + 1
    q = DevArray(case.queries[:1])
    out = DevArray(shape=(1, 4), dtype=np.uint32)
    assert L.saqgpu_probe_noalloc(ix.h, q.data_ptr(), big, 4, 0, out.data_ptr(), None) == 1
    assert b"scratch" in L.saqgpu_last_error()
    q.free()
    out.free()
    ix.close()
