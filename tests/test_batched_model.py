"""The batched scan's numpy model (tests/batched_model.py) against the accurate SAQ estimate
computed in float64 with the unquantised residual query (caq_estimator.hpp:190-217:
|o|^2 + |r|^2 - 2 sum_s rescale_s ip_s, ip_s = sqd_s sum r code + (-1 + sqd_s / 2) sum r).

The model differs from that only by the query's quantisation, whose error has a closed-form
bound per segment: |r - s1 v1| <= s1 / 2 (int8), |r - s1 v1 - s2 v2| <= s2 / 2 (int8x2), and
|r / amax - bf16(r / amax)| <= 2^-9 |r / amax| (bf16, round to nearest), each times sqd sum |code|
(weighted by |r| for bf16) and 2 |rescale|.  A wrong scale, sign, shift correction or constant
term exceeds these bounds.  CPU only.
"""
import numpy as np
import pytest

import batched_model as BM
import saq_fixtures as F


def exact_distances(case, tables, qrot, qi, c, qs):
    ids, segs = tables[c]
    crot = case.clusters[c]["centroid_rot"].astype(np.float64)
    offs = np.cumsum([0] + [d for d, _ in case.plan])
    r_all = qrot[qi].astype(np.float64) - crot
    dist = np.zeros(len(ids)) + np.sum(r_all[:offs[-1]] ** 2)
    bound = np.zeros(len(ids))
    for s, (ds, b) in enumerate(case.plan):
        codes, norm, resc = segs[s]
        dist += norm.astype(np.float64) ** 2
        if b == 0:
            continue
        r = r_all[offs[s]:offs[s] + ds]
        sqd = 2.0 / (1 << b)
        ip = sqd * (codes @ r) + (-1 + sqd / 2) * r.sum()
        dist -= 2 * resc.astype(np.float64) * ip
        amax = np.abs(r.astype(np.float32)).max()
        ac = np.abs(codes).astype(np.float64)
        per = {1: (amax / 127) / 2 * ac.sum(1), 2: (amax / 127 / 254) / 2 * ac.sum(1),
               3: 2.0 ** -9 * (ac @ np.abs(r))}
        bound += 2 * np.abs(resc) * sqd * per[qs]
    return ids, dist, bound


@pytest.mark.parametrize("plan", [[(128, 8)], [(64, 9), (32, 3), (32, 0)], [(128, 4)]])
@pytest.mark.parametrize("qs", [1, 2, 3])
def test_model_within_query_quantisation_bound(plan, qs):
    d = sum(x for x, _ in plan)
    case = F.make_case(n=600, d=d, K=4, nq=5, plan=plan, seed=81)
    tabs = BM.vector_tables(case)
    qr = BM.rotated_queries(case)
    probes = F.orc_probes(case, 2)
    _, alld = BM.model_search(case, tabs, qr, probes, 10, qs=qs)
    checked = 0
    for qi in range(qr.shape[0]):
        for c in probes[qi]:
            ids, ref, bound = exact_distances(case, tabs, qr, qi, int(c), qs)
            for u, want, bd in zip(ids.tolist(), ref.tolist(), bound.tolist()):
                got = alld[qi][u]
                slack = 2e-6 * max(1.0, abs(want))  # float32 epilogue
                assert abs(got - want) <= bd + slack, f"q{qi} list {c} id {u}: {got} vs {want} (bound {bd})"
                checked += 1
    assert checked > 100


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -(1.0 + 2.0 ** -8 + 2.0 ** -12),
                  0.0, 2.0 ** -130], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * 2.0 ** -7, 1.0, -(1.0 + 2.0 ** -7), 0.0, 2.0 ** -130], np.float32)
    assert np.array_equal(BM.bf16_rne(x), want)
