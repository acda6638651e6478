"""Parity of the gfx950 CAQ encoder (saqgpu_encode / saqgpu_encode_residuals) against
the oracle: every byte of the reference-layout cluster arrays identical (short and
long codes, block factors, ExFactor, rotated centroids), then searches over the
GPU-built index identical to searches over the oracle-built one."""
import ctypes as C

import numpy as np
import pytest

import saq_fixtures as F
from oracle_lib import CaqCfg, oracle

pytestmark = pytest.mark.gpu

KEYS = ("short_factors", "short_code", "long_code", "long_factors", "centroid_rot")


def compare_clusters(case, ix, keys=KEYS):
    for c in range(case.K):
        got = ix.get_cluster(c)
        ref = case.clusters[c]
        for k in keys:
            a, b = got[k], ref[k][:got[k].size]
            if k in ("short_factors", "long_factors", "centroid_rot"):
                a, b = a.view(np.uint32), b.view(np.uint32)
            bad = np.nonzero(a != b)[0]
            assert len(bad) == 0, f"cluster {c} {k}: {len(bad)} mismatches, first at {bad[:5]}"
        assert np.array_equal(got["ids"], ref["ids"][:ref["num_vec"]])


@pytest.mark.parametrize("plan", [[(128, 4)], [(64, 1), (64, 5), (128, 3)], [(64, 9), (64, 13), (64, 2), (64, 0)],
                                  [(128, 8), (128, 6)], [(64, 7), (64, 10), (64, 12), (64, 11)]])
def test_full_encode_matches_oracle(plan):
    d = sum(x for x, _ in plan)
    case = F.make_case(n=2500, d=d - 24 if d > 128 else d, K=9, nq=12, plan=plan, seed=31, tau=32)
    ix = F.gpu_index(case)
    # wipe, then encode on the GPU
    fresh = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                               case.variance, case.rotations))
    nfail = fresh.encode(case.base, case.cluster_ids, rd_lmt=case.rd_lmt, eps=case.eps)
    assert nfail == 0
    compare_clusters(case, fresh)
    probes = F.orc_probes(case, 4)
    a, da, _ = ix.search(case.queries, 10, 4, probes=probes)
    b, db, _ = fresh.search(case.queries, 10, 4, probes=probes)
    assert np.array_equal(a, b) and np.array_equal(da.view(np.uint32), db.view(np.uint32))
    ix.close()
    fresh.close()


@pytest.mark.parametrize("rd_lmt,eps", [(0, 1e-8), (1, 1e-8), (6, 0.0), (20, 1e-3)])
def test_adjustment_settings(rd_lmt, eps):
    case = F.make_case(n=1200, d=128, K=5, nq=4, plan=[(64, 4), (64, 3)], seed=33, rd_lmt=rd_lmt, eps=eps)
    fresh = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                               case.variance, case.rotations))
    assert fresh.encode(case.base, case.cluster_ids, rd_lmt=rd_lmt, eps=eps) == 0
    compare_clusters(case, fresh)
    fresh.close()


def test_no_rotation_encode():
    case = F.make_case(n=1500, d=200, K=6, nq=4, plan=[(128, 4), (128, 2)], seed=35, rotate=False)
    fresh = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                               case.variance, case.rotations))
    fresh.encode(case.base, case.cluster_ids)
    compare_clusters(case, fresh)
    fresh.close()


def test_encode_from_residuals_matches_oracle():
    """Codes and factors from given rotated residuals, independent of the rotation order."""
    plan = [(128, 5), (64, 9), (64, 1)]
    case = F.make_case(n=900, d=256, K=4, nq=2, plan=plan, seed=37, encode=False)
    g = np.random.default_rng(5)
    O = oracle()
    cfg = CaqCfg(6, 1e-8, 0)
    base_case = F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                       case.variance, [None] * len(plan))
    F.encode_oracle(base_case)  # identity rotation: residual = x - c exactly
    ix = F.gpu_index(base_case)
    # reset codes: upload zeros but keep centroids
    for c in range(case.K):
        clu = base_case.clusters[c]
        z = {k: np.zeros_like(v) if isinstance(v, np.ndarray) else v for k, v in clu.items()}
        z["centroid_rot"] = clu["centroid_rot"]
        z["ids"] = clu["ids"]
        ix.set_cluster(c, z)
    for c in range(case.K):
        members = np.nonzero(case.cluster_ids == c)[0]
        off = 0
        for s, (ds, b) in enumerate(plan):
            cp = max(0, min(ds, case.num_dim - off))
            X = np.zeros((len(members), ds), np.float32)
            X[:, :cp] = case.base[members, off:off + cp]
            cen = np.zeros(ds, np.float32)
            cen[:cp] = case.centroids[c, off:off + cp]
            res = (X - cen[None, :]).astype(np.float32)
            ix.encode_residuals(c, s, res, cen)
            off += ds
    compare_clusters(base_case, ix, keys=("short_factors", "short_code", "long_code", "long_factors"))
    ix.close()


def test_degenerate_vectors():
    """Zero residuals (delta = 0) and constant vectors encode like the reference."""
    plan = [(64, 4)]
    n = 96
    ids = np.zeros(n, np.uint32)
    case = F.make_case(n=n, d=64, K=1, nq=2, plan=plan, seed=39, cluster_ids=ids, encode=False)
    case.base[:10] = case.centroids[0]  # residual exactly zero
    case.base[10:20] = case.centroids[0] + 1.0
    F.encode_oracle(case)
    fresh = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                               case.variance, case.rotations))
    fresh.encode(case.base, case.cluster_ids)
    compare_clusters(case, fresh)
    fresh.close()


def test_encode_rejects_bad_cluster_ids():
    """IVF::construct's bucketing, done on the GPU, still refuses a cluster id past the last list
    and ids whose lists' sizes differ from the index's; valid ids from host or device memory build
    the same index."""
    from hip_mem import DevArray
    from saq_amd import SaqGpuError
    case = F.make_case(n=700, d=64, K=5, nq=2, plan=[(64, 4)], seed=43)
    mk = lambda: F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids,  # noqa: E731
                                    case.cluster_ids, case.variance, case.rotations))
    ix = mk()
    bad = case.cluster_ids.copy()
    bad[123] = 5
    with pytest.raises(SaqGpuError, match="out of range"):
        ix.encode(case.base, bad)
    moved = case.cluster_ids.copy()
    moved[7] = (moved[7] + 1) % 5
    with pytest.raises(SaqGpuError, match="disagree"):
        ix.encode(case.base, moved)
    ix.encode(case.base, case.cluster_ids)
    compare_clusters(case, ix)
    ix.close()
    ix = mk()
    xb, cb = DevArray(case.base), DevArray(case.cluster_ids)
    ix.encode(xb, cb)
    compare_clusters(case, ix)
    xb.free()
    cb.free()
    ix.close()


@pytest.mark.parametrize("plan,qb", [([(64, 4), (64, 3)], 7), ([(128, 1)], 4), ([(64, 5), (64, 0)], 9)])
def test_ori_qb_encode_matches_oracle(plan, qb):
    """caq_ori_qB: codes found at qb bits and shifted down (downUpSample) -- every byte equals the
    oracle's with the same setting, and (above 1 bit, where the shifted sign codes can coincide
    with the plain ones) differs from the plain encode."""
    d = sum(x for x, _ in plan)
    case = F.make_case(n=900, d=d, K=5, nq=2, plan=plan, seed=45, encode=False)
    F.encode_oracle(case, ori_qb=qb)
    ix = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                            case.variance, case.rotations))
    ix.encode(case.base, case.cluster_ids, ori_qb=qb)
    compare_clusters(case, ix)
    if max(b for _, b in plan) > 1:
        ix.encode(case.base, case.cluster_ids)
        plain = ix.get_cluster(0)
        assert not np.array_equal(plain["long_factors"], case.clusters[0]["long_factors"])
    ix.close()


def test_ori_qb_must_exceed_every_segment():
    from saq_amd import SaqGpuError
    case = F.make_case(n=300, d=128, K=3, nq=2, plan=[(64, 4), (64, 2)], seed=47)
    ix = F.gpu_index(case)
    for qb in (4, 3, 17):
        with pytest.raises(SaqGpuError, match="caq_ori_qB"):
            ix.encode(case.base, case.cluster_ids, ori_qb=qb)
    ix.close()


def test_one_centroid_encode_matches_oracle():
    """use_1_centroid: every list quantised against data.colwise().mean() (computed in the
    reference's order by saqgpu_column_mean_var) -- bytes equal the oracle's, every list's rotated
    centroid is the shared one, and search still probes by the lists' own centroids."""
    import saq_amd
    case = F.make_case(n=1500, d=200, K=6, nq=8, plan=[(128, 4), (128, 2)], seed=49, encode=False)
    mean, _ = saq_amd.column_mean_var(case.base)
    F.encode_oracle(case, one_centroid=np.concatenate([mean, np.zeros(case.D_pad - case.num_dim, np.float32)]))
    ix = F.gpu_index(F.Case(case.num_dim, case.plan, case.base, case.queries, case.centroids, case.cluster_ids,
                            case.variance, case.rotations))
    ix.encode(case.base, case.cluster_ids, centroid=mean)
    compare_clusters(case, ix)
    crot = [ix.get_cluster(c)["centroid_rot"] for c in range(case.K)]
    assert all(np.array_equal(crot[0], x) for x in crot[1:])
    assert np.array_equal(ix.probe(case.queries, 3), F.orc_probes(case, 3))
    ix.close()
