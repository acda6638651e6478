"""Benchmark of the SAQ IVF scan on one MI355X: the batched int8-MFMA scan against the plain
(reference-exact) GPU scan, at equal recall.

Workloads (synthetic, seeded; PCA-rotated with an FP64 covariance; k-means lists, seed 0):

zipf10m-d1024 (default: the configuration the headline metric and its targets are quoted on)
    N = 10,000,000, D = 1024, a 2,048-component mixture with Zipf(1.1) component weights
    (centres ~ N(0, 4I), per-dim std 1/sqrt(1 + i/32)), nlist = 16384, B = 4 (--bits),
    10,000 queries whose nprobe is drawn per query from U[32, 256] (seed 3), k = 10.  The
    recall@10 both scans reach is in the JSON line: with at most 256 of 16384 uneven lists
    probed, probing, not the codes, caps it well below 0.95 on this generator.
mix1m-d960
    N = 1,000,000 base vectors, D = 960 fp32, from a 256-component Gaussian mixture
    (std decay 1/sqrt(1 + i/32)), nlist = 4096, 1,000 held-out queries, k = 10.  On this
    mixture the estimator's quantisation error, not probing, limits recall (B = 4 levels off
    near 0.81 recall@10 for any nprobe), so the bits per dimension are the smallest of 4..8
    whose recall@10 against exact FP64 ground truth reaches 0.95 at nprobe 64, and nprobe is
    then the smallest value of a sweep that keeps it.
serve5m-d1536
    N = 5,000,000, D = 1536, a 512-component mixture (std decay 1/sqrt(1 + i/64)),
    nlist = 8192, B = 1.5, a pool of 4,096 queries; besides the batch measurements, the
    serving loop at offered loads of 2k..50k queries/s.

Every run ends with the serving loop (Poisson arrivals of 1..64-query requests, k in
{10, 100}, nprobe in [16, 128], dynamically batched onto the batched scan) at the offered
loads 2k..50k queries/s; its p50/p99 request latency per load is in config.serving.  It runs
on the serving configuration's own index (serve5m-d1536, built after the other workload's
measurements), or on the workload's index when that is serve5m-d1536 itself.

A step = one search of the whole query batch (GPU coarse quantiser + query rotation +
scan), inputs resident in HBM.  The plain scan is timed first; the batched scan is then
run at the smallest nprobe (and the cheaper of its int8 / int8x2 precisions) whose
recall@10 is at least the plain scan's (with per-query nprobe: the same nprobe, the
cheaper precision that reaches it), with the query tile its autotuner picks.  `value` is
the batched scan's throughput; the plain scan's is in config.plain.  The JSON line also
carries each scan kernel's roofline (the batched tile kernel: int8 MFMA ops, 2 * sum D_s per
(query, probed vector) pair, vs 5.0 POP/s, with its HBM rate beside it; the plain scan:
algorithmic bytes / HIP-event kernel time vs 8 TB/s), the GPU encoder's throughput, the
single-vector records' quantiser and estimator rates (config.single_vector), and a CPU
baseline timed on every host core on a bounded sample of the same queries, whose IDs are
also compared with the plain GPU scan's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS @ recall@10=0.95 batched vs plain GPU; encode Mvec/s; p99 ms"
HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X

WORKLOADS = {
    "mix1m-d960": dict(n=1_000_000, dim=960, comps=256, tau=32, zipf=0.0, nlist=4096, nq=1000, bits=0.0,
                       npq_range=None, steps=20, kmeans_iters=8, cpu_queries=0),
    "zipf10m-d1024": dict(n=10_000_000, dim=1024, comps=2048, tau=32, zipf=1.1, nlist=16384, nq=10000, bits=4.0,
                          npq_range=(32, 256), steps=10, kmeans_iters=6, cpu_queries=64),
    "serve5m-d1536": dict(n=5_000_000, dim=1536, comps=512, tau=64, zipf=0.0, nlist=8192, nq=4096, bits=1.5,
                          npq_range=(16, 128), steps=3, kmeans_iters=6, cpu_queries=16,
                          serve_loads=(2000, 5000, 10000, 20000, 50000)),
}


def serve(torch, ix, queries, loads, seconds, seed, precision, exact_probe, log_fn):
    """The serving loop: Poisson request arrivals at each offered load (queries/s), each request
    1..64 queries with k in {10, 100} and one nprobe in [16, 128]; whatever has arrived when the
    GPU frees up is served as one batched search (k = the batch's largest, nprobe per query,
    at most 4096 queries).  Latency = completion - arrival of each request.  Returns one row per
    load."""
    rng = np.random.default_rng(seed)
    dev = queries.device
    nq_pool = queries.shape[0]
    # warm-up (untimed): the largest batch the loop can form sizes the index's workspaces, so no
    # timed request waits on their growth
    nw = min(4096, nq_pool)
    ix.search_batched(queries[:nw].contiguous(), 100, 128,
                      nprobe_per_query=torch.full((nw,), 128, dtype=torch.int32, device=dev), precision=precision,
                      out_ids=torch.empty(nw, 100, dtype=torch.int32, device=dev), want_dists=False,
                      exact_probe=exact_probe)
    torch.cuda.synchronize()
    rows = []
    for load in loads:
        mean_q = 32.5
        rate = load / mean_q  # requests/s
        nreq = max(1, int(rate * seconds))
        arr = np.cumsum(rng.exponential(1.0 / rate, nreq))
        nqs = rng.integers(1, 65, nreq)
        ks = np.where(rng.random(nreq) < 0.5, 10, 100)
        nps = rng.integers(16, 129, nreq)
        starts = rng.integers(0, nq_pool - 64, nreq)
        lat = np.zeros(nreq)
        batches = []
        i = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while i < nreq:
            now = time.perf_counter() - t0
            if arr[i] > now:
                time.sleep(min(arr[i] - now, 0.0005))
                continue
            j, tot = i, 0
            while j < nreq and arr[j] <= now and tot + nqs[j] <= 4096:
                tot += int(nqs[j])
                j += 1
            idx = np.concatenate([np.arange(starts[r], starts[r] + nqs[r]) for r in range(i, j)])
            npq = np.concatenate([np.full(nqs[r], nps[r], np.int32) for r in range(i, j)])
            kb = int(ks[i:j].max())
            qb = queries[torch.from_numpy(idx).to(dev)]
            npq_d = torch.from_numpy(npq).to(dev)
            ids = torch.empty(tot, kb, dtype=torch.int32, device=dev)
            ix.search_batched(qb, kb, int(npq.max()), nprobe_per_query=npq_d, precision=precision, out_ids=ids,
                              want_dists=False, exact_probe=exact_probe)
            torch.cuda.synchronize()
            done = time.perf_counter() - t0
            lat[i:j] = done - arr[i:j]
            batches.append(tot)
            i = j
        wall = time.perf_counter() - t0
        served = float(nqs.sum())
        row = {"offered_qps": load, "achieved_qps": round(served / wall, 1), "requests": nreq,
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "mean_batch_queries": round(float(np.mean(batches)), 1)}
        log_fn(f"serving {row}")
        rows.append(row)
    return rows


def serving_config(torch, saq_amd, make_plan, sdata, dev, seconds, precision, log_fn, workload="serve5m-d1536"):
    """The serving configuration on an index of its own: the workload's base (5M x 1536, a
    512-component mixture, nlist 8192, B = 1.5) built on the GPU, then the serving loop at each of
    the workload's offered loads.  Returns {"workload": ..., "rows": [...]}."""
    W = WORKLOADS[workload]
    n, D, nlist = W["n"], W["dim"], W["nlist"]
    log_fn(f"serving workload {workload}: N={n} D={D} nlist={nlist} B={W['bits']}")
    t0 = time.time()
    base, queries = make_data(torch, n, D, W["comps"], W["tau"], W["zipf"], W["nq"], 1, dev)
    mean, comps, var = pca(torch, base)
    base = rotate_inplace(torch, base, mean, comps)
    queries = ((queries - mean) @ comps).contiguous()
    cent, cid = kmeans(torch, base, nlist, W["kmeans_iters"], 0)
    sizes = torch.bincount(cid, minlength=nlist).cpu().numpy().astype(np.int64)
    plan = make_plan(var, D, W["bits"])
    rots = [sdata.random_rotation(d, 1000 + i) for i, (d, _) in enumerate(plan)]
    ix = saq_amd.Index(D, plan, cent.cpu().numpy(), sizes, var, rotations=rots)
    ix.encode(base, cid.to(torch.int32).contiguous())
    del base, cid
    ix.prepare_batched()
    torch.cuda.synchronize()
    log_fn(f"serving index built in {time.time() - t0:.1f}s, plan {plan}")
    rows = serve(torch, ix, queries, W["serve_loads"], seconds, 5, precision, False, log_fn)
    ix.close()
    del queries
    torch.cuda.empty_cache()
    return {"workload": f"{workload}: N={n} D={D} nlist={nlist} B={W['bits']} plan {plan}; requests of 1..64 "
                        f"queries, k in {{10, 100}}, nprobe in [16, 128], Poisson arrivals, {seconds:g} s per load",
            "rows": rows}


def read_traffic(name, n, bits, nprobe, **extra):
    """HBM bytes per launch of a kernel from its committed FETCH_SIZE/WRITE_SIZE summary, when
    that summary was taken on this workload (extra: further keys that must match)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        t = json.load(open(path))
        if t.get("nprobe") == nprobe and t.get("bits") == bits and t.get("n") == n and \
                all(t.get(k) == v for k, v in extra.items() if k != "query_tile"):
            if "by_query_tile" in t:
                return t["by_query_tile"].get(str(extra.get("query_tile")))
            return t.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(torch, n, d, comps, tau, zipf, nq, seed, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    centers = torch.randn(comps, d, generator=g, device=dev) * 2.0
    std = 1.0 / torch.sqrt(1.0 + torch.arange(d, device=dev, dtype=torch.float32) / tau)
    if zipf > 0:
        w = 1.0 / torch.arange(1, comps + 1, device=dev, dtype=torch.float64) ** zipf
    else:
        w = torch.ones(comps, device=dev, dtype=torch.float64)
    w = (w / w.sum()).float()

    def draw(m):
        if zipf > 0:
            comp = torch.multinomial(w, m, replacement=True, generator=g)
        else:
            comp = torch.randint(0, comps, (m,), generator=g, device=dev)
        return centers[comp] + torch.randn(m, d, generator=g, device=dev) * std

    base = torch.empty(n, d, device=dev)
    step = 1 << 18
    for s in range(0, n, step):
        e = min(n, s + step)
        base[s:e] = draw(e - s)
    queries = draw(nq)
    return base, queries


def pca(torch, base):
    """Mean and covariance accumulated in FP64, chunk by chunk; eigenvectors by decreasing
    variance."""
    n, d = base.shape
    step = 1 << 17
    mean = torch.zeros(d, dtype=torch.float64, device=base.device)
    for s in range(0, n, step):
        mean += base[s:s + step].double().sum(0)
    mean /= n
    cov = torch.zeros(d, d, dtype=torch.float64, device=base.device)
    for s in range(0, n, step):
        xc = base[s:s + step].double() - mean
        cov += xc.T @ xc
    cov /= n
    w, v = torch.linalg.eigh(cov.cpu())
    order = torch.argsort(w, descending=True)
    comps = v[:, order].float().to(base.device)
    var = w[order].clamp(min=0).float().numpy()
    return mean.float(), comps, var


def rotate_inplace(torch, x, mean, comps):
    step = 1 << 18
    for s in range(0, x.shape[0], step):
        x[s:s + step] = (x[s:s + step] - mean) @ comps
    return x


def kmeans(torch, x, k, iters, seed):
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    cent = x[torch.randperm(x.shape[0], generator=g, device=x.device)[:k]].clone()
    step = 1 << 18
    for _ in range(iters):
        ids = assign(torch, x, cent)
        sums = torch.zeros(k, x.shape[1], dtype=torch.float64, device=x.device)
        for s in range(0, x.shape[0], step):
            sums.index_add_(0, ids[s:s + step], x[s:s + step].double())
        cnt = torch.bincount(ids, minlength=k)
        nz = cnt > 0
        cent[nz] = (sums[nz] / cnt[nz, None]).float()
    return cent, assign(torch, x, cent)


def assign(torch, x, cent):
    cn = (cent * cent).sum(1)
    out = torch.empty(x.shape[0], dtype=torch.int64, device=x.device)
    step = max(1, min(1 << 16, (1 << 30) // cent.shape[0]))
    for s in range(0, x.shape[0], step):
        d = cn[None, :] - 2.0 * (x[s:s + step] @ cent.T)
        out[s:s + step] = d.argmin(1)
    return out


def exact_topk(torch, base, queries, k, cand=64):
    """Exact FP64 top-k: FP32 distances pick `cand` candidates per query, FP64 re-ranks them."""
    bn = (base * base).sum(1)
    out = []
    qstep = max(1, min(256, (1 << 31) // base.shape[0]))
    for s in range(0, queries.shape[0], qstep):
        q = queries[s:s + qstep]
        d = bn[None, :] - 2.0 * (q @ base.T)
        c = d.topk(cand, dim=1, largest=False).indices
        xb = base[c].double()  # qstep x cand x D
        dd = ((xb - q.double()[:, None, :]) ** 2).sum(2)
        o = dd.argsort(dim=1)[:, :k]
        out.append(torch.gather(c, 1, o))
    return torch.cat(out).cpu().numpy()


def recall(found, gt, k):
    hit = 0
    for f, g in zip(found, gt):
        hit += len(set(f[:k].tolist()) & set(g[:k].tolist()))
    return hit / (len(gt) * k)


def single_records(torch, ix, base, queries, dev, nrec, nqs, reps=3, cpu=None):
    """The single-vector path (SAQuantizerSingle / SaqSingleEstimator) on nrec base rows and nqs
    queries, device buffers throughout: quantiser throughput, and the estimator's pair rate and
    its record bandwidth (each record is read from HBM once; the tile's other queries hit L2).
    cpu = (plan, rotations, variance): also time the oracle's estimator (one thread) on a bounded
    sample of the same records and queries and compare its estimates with the GPU's."""
    nbytes = ix.single_layout()[0]
    stride = (nbytes + 63) // 64 * 64
    data = base[:nrec].contiguous()
    recs = torch.zeros(nrec * stride, dtype=torch.uint8, device=dev)
    q = queries[:nqs].contiguous()
    outs = tuple(torch.empty(nqs, nrec, dtype=torch.float32, device=dev) for _ in range(3))
    ix.single_quantize(data, stride=stride, out=recs)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.single_quantize(data, stride=stride, out=recs)
    torch.cuda.synchronize()
    t_q = time.perf_counter() - t0
    ix.single_estimate(q, recs, nrec, stride, out=outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ix.single_estimate(q, recs, nrec, stride, out=outs)
    torch.cuda.synchronize()
    t_e = (time.perf_counter() - t0) / reps
    res = {"records": nrec, "record_bytes": nbytes, "queries": nqs,
           "quantize_mvec_per_s": round(nrec / t_q / 1e6, 3), "estimate_ms": round(t_e * 1e3, 3),
           "estimate_pairs_per_s": round(nqs * nrec / t_e, 1),
           "estimate_record_gbs": round(nrec * nbytes / t_e / 1e9, 1)}
    if cpu is not None:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import ctypes as C
        from oracle_lib import make_plan, oracle
        plan, rots, variance = cpu
        O = oracle()
        ns = min(nrec, 20000)
        host = recs[:ns * stride].cpu().numpy()
        qs = q.cpu().numpy()
        p = make_plan([x for x, _ in plan], [b for _, b in plan], qs.shape[1])
        rp = (C.c_void_p * len(plan))(*[r.ctypes.data for r in rots])
        vpad = np.zeros(sum(x for x, _ in plan), np.float32)
        vpad[:len(variance)] = variance
        got = [o[:, :ns].cpu().numpy() for o in outs]
        wv, wf, wa = (np.zeros(ns, np.float32) for _ in range(3))
        pairs, dt, mism = 0, 0.0, 0
        while dt < 3.0 and pairs < nqs * ns:
            qi = (pairs // ns) % nqs
            t0 = time.perf_counter()
            O.orc_single_estimate(C.byref(p), vpad, rp, host, ns, stride, np.ascontiguousarray(qs[qi]), 4.0, 0,
                                  wv, wf, wa)
            dt += time.perf_counter() - t0
            pairs += ns
            mism += int(np.sum((wa != got[2][qi]) | (wf != got[1][qi]) | (wv != got[0][qi])))
        res["cpu_baseline"] = {"pairs_per_s": round(pairs / dt, 1), "cores": 1, "kind": "port",
                               "sample": f"{pairs} (query, record) pairs: {ns} records x {pairs // ns} queries, "
                                         f"{dt:.1f}s", "estimates_differing_from_gpu": mism}
    log(f"single-vector records: {res}")
    del recs, outs
    return res


def precision_report(torch, saq_amd, make_plan, sdata, base, queries, cent, sizes, var, cid32, npq_np, topk, gt,
                     dev, stream, steps, log_fn, nq_exact=1000, nq_ref=16):
    """The same base at B=8 (10M x 1024: codes of 8 bits, stored shifted to signed int8), searched by
    the batched scan with int8, int8x2 and bf16 queries.  Per precision: the relative error
    |d - d*| / d* of the returned top-k distances against (a) the FP64 exact distance |q - o|^2 of
    the same vector (nq_exact queries) and (b) the reference's accurate estimate of the same vector
    (compAccurateDist, fp32 with the unquantised query, saqgpu_estimate; nq_ref queries at k=256),
    which isolates the query quantisation; recall@k and the step time on the whole query batch."""
    D = base.shape[1]
    plan = make_plan(var, D, 8.0)
    rots = [sdata.random_rotation(d, 2000 + i) for i, (d, _) in enumerate(plan)]
    ix = saq_amd.Index(D, plan, cent.cpu().numpy(), sizes, var, rotations=rots)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.encode(base, cid32)
    torch.cuda.synchronize()
    t_enc = time.perf_counter() - t0
    ix.prepare_batched()
    nq = queries.shape[0]
    npmax = int(npq_np.max()) if npq_np is not None else 64
    npq_all = (torch.from_numpy(npq_np).to(dev) if npq_np is not None
               else torch.full((nq,), npmax, dtype=torch.int32, device=dev))
    probes = ix.probe(queries, npmax)  # one probe list for every precision
    ids = torch.zeros(nq, topk, dtype=torch.int32, device=dev)
    dd = torch.zeros(nq, topk, dtype=torch.float32, device=dev)
    qe = queries[:nq_exact].double()
    out = {"bits": 8.0, "plan": plan, "encode_mvec_per_s": round(base.shape[0] / t_enc / 1e6, 3),
           "nprobe": "per query as the headline" if npq_np is not None else npmax,
           "vs_fp64_exact": {}, "vs_reference_estimate": {}, "recall_at_10": {}, "ms_per_step": {}}
    names = {saq_amd.PREC_INT8: "int8", saq_amd.PREC_INT8X2: "int8x2", saq_amd.PREC_BF16: "bf16"}
    qn = queries[:nq_ref].cpu().numpy()
    pr_np = probes.cpu().numpy().view(np.uint32) if hasattr(probes, "cpu") else np.asarray(probes).view(np.uint32)
    ref_est = []
    for i in range(nq_ref):
        npi = int(npq_np[i]) if npq_np is not None else npmax
        e_ids, e_acc, _, _ = ix.estimate(qn[i], npi, probes=pr_np[i, :npi])
        ref_est.append(dict(zip(e_ids.tolist(), e_acc.tolist())))
    for prec, name in names.items():
        def go():
            ix.search_batched(queries, topk, npmax, probes=probes, nprobe_per_query=npq_all, precision=prec,
                              out_ids=ids, out_dists=dd, stream=stream)
        go()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            go()
        torch.cuda.synchronize()
        out["ms_per_step"][name] = round((time.perf_counter() - t0) / steps * 1e3, 3)
        got_i = ids.cpu().numpy()
        out["recall_at_10"][name] = round(recall(got_i, gt, topk), 4)
        # (a) against FP64 exact distances
        sel = ids[:nq_exact].long()
        exact = ((base[sel].double() - qe[:, None, :]) ** 2).sum(2)
        rel = ((dd[:nq_exact].double() - exact).abs() / exact.clamp_min(1e-30)).cpu().numpy().ravel()
        out["vs_fp64_exact"][name] = {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
                                      "pairs": int(rel.size)}
        # (b) against the reference's accurate estimate, top-256 of nq_ref queries
        kk = 256
        bi = torch.zeros(nq_ref, kk, dtype=torch.int32, device=dev)
        bdv = torch.zeros(nq_ref, kk, dtype=torch.float32, device=dev)
        ix.search_batched(queries[:nq_ref], kk, npmax, probes=probes[:nq_ref], nprobe_per_query=npq_all[:nq_ref],
                          precision=prec, out_ids=bi, out_dists=bdv, stream=stream)
        torch.cuda.synchronize()
        rr, missing = [], 0
        for i, (row_i, row_d) in enumerate(zip(bi.cpu().numpy().view(np.uint32), bdv.cpu().numpy())):
            for u, dv in zip(row_i.tolist(), row_d.tolist()):
                if u not in ref_est[i]:  # not one of the probed vectors
                    missing += 1
                    continue
                want = ref_est[i][u]
                rr.append(abs(dv - want) / max(abs(want), 1e-30))
        rr = np.asarray(rr)
        out["vs_reference_estimate"][name] = {"max_rel": float(rr.max()), "mean_rel": float(rr.mean()),
                                              "pairs": int(rr.size), "ids_not_probed": missing}
        log_fn(f"precision B=8 {name}: recall {out['recall_at_10'][name]}, step {out['ms_per_step'][name]} ms, "
               f"vs FP64 exact max {rel.max():.3e} mean {rel.mean():.3e}, vs reference estimate max "
               f"{rr.max():.3e} mean {rr.mean():.3e}")
    ix.close()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ix, plan, queries_np, variance, rots, probes, npq, topk, m, gpu_ids, threads, min_seconds=10.0):
    """The CPU restatement of the reference search on this host's cores: the oracle compiled with
    its fast scan on AVX-512 (oracle/libcpubase.so; the scalar oracle where the host has no
    AVX-512), queries taken one at a time per thread and each IVF::search call timed as
    test_qps does (src/test_qps.cpp:77-84, 115-123).  Only the lists the sample probes are
    downloaded from the GPU index.  Returns (test_qps QPS, wall-clock QPS, seconds, passes,
    queries whose IDs differ from the GPU's, library kind)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    from oracle_lib import Cluster, cpubase, make_plan, oracle
    L = cpubase()
    kind = "AVX-512 fast scan (oracle/libcpubase.so)"
    if L is None:
        L, kind = oracle(), "scalar (oracle/liboracle.so; host without AVX-512)"
    K = ix.num_cen
    clus = (Cluster * K)()
    keep = []
    need = set()
    for i in range(probes.shape[0]):
        need.update(probes[i, :(npq[i] if npq is not None else probes.shape[1])].tolist())
    for c in need:
        a = ix.get_cluster(c)
        keep.append(a)
        clus[c] = Cluster(int(ix.cluster_sizes[c]), a["short_factors"].ctypes.data if a["short_factors"].size else None,
                          a["short_code"].ctypes.data if a["short_code"].size else None,
                          a["long_code"].ctypes.data if a["long_code"].size else None,
                          a["long_factors"].ctypes.data if a["long_factors"].size else None,
                          a["ids"].ctypes.data if a["ids"].size else None, a["centroid_rot"].ctypes.data)
    p = make_plan([x for x, _ in plan], [b for _, b in plan], queries_np.shape[1])
    rp = (C.c_void_p * len(plan))(*[r.ctypes.data for r in rots])
    vpad = np.zeros(sum(x for x, _ in plan), np.float32)
    vpad[:len(variance)] = variance
    nq = queries_np.shape[0]
    ids = np.zeros((nq, topk), np.uint32)
    lat = np.zeros(nq, np.float64)
    queries_np = np.ascontiguousarray(queries_np)
    npq_c = np.ascontiguousarray(npq, np.int32) if npq is not None else None
    passes, dt, lat_sum, mism = 0, 0.0, 0.0, 0
    while dt < min_seconds and passes < 200:  # whole passes, to a bounded sample of ~10 s
        t0 = time.perf_counter()
        L.orc_search_batch_timed(C.byref(p), vpad, rp, clus, np.ascontiguousarray(probes, np.uint32), probes.shape[1],
                                 npq_c.ctypes.data if npq_c is not None else None, queries_np, nq, topk, m, threads,
                                 ids, None, lat)
        dt += time.perf_counter() - t0
        lat_sum += float(lat.sum())
        passes += 1
        if passes == 1:
            mism = int(np.sum((ids != gpu_ids).any(1)))
    qps_testqps = nq * passes * 1e3 / lat_sum * threads
    return qps_testqps, nq * passes / dt, dt, passes, mism, kind


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="zipf10m-d1024")
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0 = the workload's default)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--nq", type=int, default=0)
    ap.add_argument("--bits", type=float, default=-1.0,
                    help="SAQ bits per dimension (0 = the smallest of 4..8 reaching the target recall)")
    ap.add_argument("--topk", type=int, default=10)
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--nprobe", type=int, default=0, help="fixed nprobe (0 = smallest reaching the target)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="duration of each serving-loop load")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving loop (kernel profiling runs)")
    ap.add_argument("--no-single", action="store_true", help="skip the single-vector record measurement")
    ap.add_argument("--traffic-run", action="store_true",
                    help="only the launches the committed HBM-traffic summaries are measured on")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the B=8 index and its int8 / int8x2 / bf16 distance errors")
    ap.add_argument("--exact-probe", action="store_true",
                    help="batched scan with the reference-order coarse quantiser instead of the fp32-MFMA one")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (the known-size kernel that scales "
                         "FETCH_SIZE/WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    W = dict(WORKLOADS[args.workload])
    n = args.n or W["n"]
    nq = args.nq or W["nq"]
    steps = args.steps or W["steps"]
    bits_arg = W["bits"] if args.bits < 0 else args.bits

    import torch
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": nb, "launches": 5}))
        return
    import saq_amd
    from saq_amd.plan import make_plan
    from saq_amd import data as sdata

    dev = torch.device("cuda:0")
    t_all = time.time()
    D, nlist = W["dim"], W["nlist"]
    log(f"workload {args.workload}: N={n} D={D} nlist={nlist} queries={nq}")
    base, queries = make_data(torch, n, D, W["comps"], W["tau"], W["zipf"], nq, 1, dev)
    mean, comps, var = pca(torch, base)
    base = rotate_inplace(torch, base, mean, comps)
    queries = ((queries - mean) @ comps).contiguous()
    log("k-means")
    cent, cid = kmeans(torch, base, nlist, W["kmeans_iters"], 0)
    sizes = torch.bincount(cid, minlength=nlist).cpu().numpy().astype(np.int64)
    log(f"list sizes: mean {sizes.mean():.0f} max {sizes.max()} empty {(sizes == 0).sum()}")
    log("ground truth")
    gt = exact_topk(torch, base, queries, args.topk)
    cid32 = cid.to(torch.int32).contiguous()
    del cid
    out_ids = torch.zeros(nq, args.topk, dtype=torch.int32, device=dev)
    out_d = torch.zeros(nq, args.topk, dtype=torch.float32, device=dev)
    met = torch.zeros(nq, 3, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    npq_np = npq = None
    if W["npq_range"]:
        lo, hi = W["npq_range"]
        npq_np = np.random.default_rng(3).integers(lo, hi + 1, nq).astype(np.int32)
        npq = torch.from_numpy(npq_np).to(dev)

    def build(bits):
        plan = make_plan(var, D, bits)
        rots = [sdata.random_rotation(d, 1000 + i) for i, (d, _) in enumerate(plan)]
        ix = saq_amd.Index(D, plan, cent.cpu().numpy(), sizes, var, rotations=rots)
        ix.encode(base, cid32)
        return ix, plan, rots

    def search(ix, nprobe):
        ix.search(queries, args.topk, nprobe, out_ids=out_ids, out_dists=out_d, metrics=met, stream=stream,
                  nprobe_per_query=npq)

    def recall_at(ix, nprobe):
        search(ix, nprobe)
        torch.cuda.synchronize()
        return recall(out_ids.cpu().numpy(), gt, args.topk)

    # ---- bits: quantisation, not probing, limits recall on these mixtures
    bits_tried = {}
    probe_check = W["npq_range"][1] if W["npq_range"] else 64
    candidates = [bits_arg] if bits_arg else [4.0, 5.0, 6.0, 7.0, 8.0]
    ix = plan = rots = None
    for bits in candidates:
        if ix is not None:
            ix.close()
        ix, plan, rots = build(bits)
        r = recall_at(ix, probe_check)
        bits_tried[bits] = round(r, 4)
        log(f"B={bits} plan {plan}: recall@{args.topk} at nprobe {'U' if npq is not None else probe_check} = {r:.4f}")
        if r >= args.target_recall:
            break
    log(f"plan {plan}; index {ix.device_bytes() / 1e9:.2f} GB")
    if args.traffic_run:
        # the launches the HBM-traffic summaries under profiles/ count, in this order: one plain
        # search (sliced scan) at the workload's nprobe, then one batched int8 search per query
        # tile 16, 32, 64, 128 (one bscan_kernel each)
        np_t = int(W["npq_range"][1]) if W["npq_range"] else (args.nprobe or 64)
        search(ix, np_t)
        ix.prepare_batched()
        tb_i = torch.zeros(nq, args.topk, dtype=torch.int32, device=dev)
        tb_d = torch.zeros(nq, args.topk, dtype=torch.float32, device=dev)
        for qt in (16, 32, 64, 128):
            ix.search_batched(queries, args.topk, np_t, precision=saq_amd.PREC_INT8, query_tile=qt, out_ids=tb_i,
                              out_dists=tb_d, stream=stream, nprobe_per_query=npq)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_run": args.workload, "n": n, "bits": bits, "nprobe": np_t,
                          "query_tiles": [16, 32, 64, 128]}))
        return
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nfail = ix.encode(base, cid32)
    torch.cuda.synchronize()
    t_enc = time.perf_counter() - t0
    log(f"encode {n / t_enc / 1e6:.2f} Mvec/s ({t_enc * 1e3:.0f} ms), check failures {nfail}")
    single = None if args.no_single else single_records(torch, ix, base, queries, dev, min(n, 1 << 20), 16,
                                                         cpu=None if args.no_cpu else (plan, rots, var))
    # IVF::estimate (every vector of the probed lists, accurate / fast / variance estimates): one
    # query per call, host buffers, as the reference's test_relative_error drives it
    qn = queries[:8].cpu().numpy()
    ix.estimate(qn[0], 64)
    t0 = time.perf_counter()
    nv = sum(len(ix.estimate(q, 64)[0]) for q in qn)
    t_est = (time.perf_counter() - t0) / len(qn)
    est = {"nprobe": 64, "vectors_per_query": nv // len(qn), "ms_per_query": round(t_est * 1e3, 3),
           "vectors_per_s": round(nv / len(qn) / t_est, 1)}
    log(f"estimate: {est}")

    # ---- plain nprobe
    if npq is not None:
        nprobe = int(W["npq_range"][1])
        sweep = [nprobe]
        rec = bits_tried[bits]
    else:
        nprobe, rec = args.nprobe, None
        sweep = [args.nprobe] if args.nprobe else [4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128]
        for npb in sweep:
            rec = recall_at(ix, npb)
            log(f"nprobe {npb}: recall@{args.topk} {rec:.4f}")
            nprobe = npb
            if rec >= args.target_recall:
                break

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ix.timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms, launches = ix.timing(False)
        return dt, kms / max(1, launches)

    # ---- plain scan (bit-exact with the reference)
    dt, kern_ms = timed(lambda: search(ix, nprobe))
    qps = nq * steps / dt
    ids_gpu = out_ids.cpu().numpy().astype(np.uint32)
    m = met.cpu().numpy().astype(np.int64)

    # algorithmic bytes of one scan launch: every probed block record (short codes of all quantised
    # segments + per-segment factors), the refined extra-bit codes, their ExFactor/id, and the
    # probed rotated centroids
    block_stride = len(plan) * 32 * 4 + sum(d // 8 * 32 for d, b in plan if b > 0)
    blocks = m[:, 2].sum() // 32
    per_refine = sum(d * (b - 1) // 8 for d, b in plan if b > 1)
    refines = m[:, 1].sum() // (per_refine * 8) if per_refine else 0  # full-refine equivalents
    d_pad = sum(d for d, _ in plan)
    probed = int(npq_np.sum()) if npq_np is not None else nq * nprobe
    alg_bytes = blocks * block_stride + m[:, 1].sum() // 8 + refines * (8 * len(plan) + 4) + probed * d_pad * 4
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    traffic = read_traffic("scan_traffic.json", n, bits, nprobe)
    plain = {"qps": round(qps, 1), "recall_at_10": round(rec, 4), "nprobe": nprobe if npq is None else "U" + str(
        W["npq_range"]), "ms_per_step": round(dt / steps * 1e3, 4),
             "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                          "kernel": "sliced scan: list_scan_kernel + refine_kernel + replay_kernel per slice",
                          "alg_bytes_per_launch": int(alg_bytes), "kernel_ms": round(kern_ms, 4)}}
    log(f"plain: {qps:.0f} QPS, scan kernel {kern_ms:.3f} ms, {achieved:.0f} GB/s")
    # the two forms of the exact coarse quantiser (GEMM filter, Eigen order for every pair) on the
    # full batch: their lists must be equal
    p_np = W["npq_range"][1] if W["npq_range"] else nprobe
    prev = os.environ.pop("SAQGPU_PROBE", None)
    os.environ["SAQGPU_PROBE"] = "gemm"
    pr_gemm = ix.probe(queries, p_np).cpu().numpy()
    os.environ["SAQGPU_PROBE"] = "tile"
    pr_tile = ix.probe(queries, p_np).cpu().numpy()
    del os.environ["SAQGPU_PROBE"]
    if prev is not None:
        os.environ["SAQGPU_PROBE"] = prev
    plain["probe_lists_differing_queries"] = int((pr_gemm != pr_tile).any(1).sum())
    log(f"exact coarse quantiser, GEMM filter vs every pair: {plain['probe_lists_differing_queries']} queries differ")

    # ---- batched int8-MFMA scan at a recall no lower than the plain scan's
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ix.prepare_batched()
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    bids = torch.zeros(nq, args.topk, dtype=torch.int32, device=dev)
    bd = torch.zeros(nq, args.topk, dtype=torch.float32, device=dev)

    def bsearch(npb, prec, qt=0):
        ix.search_batched(queries, args.topk, npb, precision=prec, query_tile=qt, out_ids=bids, out_dists=bd,
                          stream=stream, exact_probe=args.exact_probe, nprobe_per_query=npq)

    def brecall(npb, prec):
        bsearch(npb, prec)
        torch.cuda.synchronize()
        return recall(bids.cpu().numpy(), gt, args.topk)

    choices, best_seen = [], None
    for prec in (saq_amd.PREC_INT8, saq_amd.PREC_INT8X2):
        for npb in [x for x in sweep if x <= max(nprobe * 2, nprobe)]:
            r = brecall(npb, prec)
            log(f"batched prec={prec} nprobe {npb}: recall@{args.topk} {r:.4f}")
            if best_seen is None or r > best_seen[0]:
                best_seen = (r, prec, npb)
            if r >= rec:
                qt = ix.autotune_batched(queries, args.topk, npb, precision=prec, stream=stream,
                                         exact_probe=args.exact_probe, nprobe_per_query=npq)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(3):
                    bsearch(npb, prec)
                torch.cuda.synchronize()
                choices.append(((time.perf_counter() - t0) / 3, prec, npb, qt, r))
                break
    below = not choices
    if below:  # report the most accurate configuration, flagged
        r, prec, npb = best_seen
        qt = ix.autotune_batched(queries, args.topk, npb, precision=prec, stream=stream,
                                 exact_probe=args.exact_probe, nprobe_per_query=npq)
        choices.append((0.0, prec, npb, qt, r))
    _, bprec, bnp, bqt, brec = min(choices)
    bdt, bkern_ms = timed(lambda: bsearch(bnp, bprec, bqt))
    bqps = nq * steps / bdt
    # algorithmic work of one tile-kernel launch: per work tile (a list and up to bqt of its
    # queries) one pass over the list's 16-vector records; each pair's query operand row (int8:
    # one byte per quantised dimension; int8x2 and BF16 two) and constants read once; the per-pair
    # candidate lists; MFMA ops = 2 * K per (query, vector) pair and term
    pr = ix.probe(queries.cpu().numpy(), bnp).astype(np.int64)  # exact lists; the fast ones differ only at ties
    if npq_np is not None:
        pr = np.concatenate([pr[i, :npq_np[i]] for i in range(nq)])
    cnt = np.bincount(pr.ravel(), minlength=nlist)
    nmb = (sizes + 15) // 16
    kq_bytes = sum((512 if b <= 4 else (1024 if b <= 8 else 2048)) * (d // 64) for d, b in plan if b > 0)
    nqs = sum(1 for _, b in plan if b > 0)
    rstride = (kq_bytes + 64 * (2 + nqs) + 255) // 256 * 256
    tiles = (cnt + bqt - 1) // bqt
    a_row = sum(d for d, b in plan if b > 0) * (1 if bprec == saq_amd.PREC_INT8 else 2)
    b_bytes = int((tiles * nmb).sum() * rstride + pr.size * (a_row + (3 * nqs + 1) * 4 + args.topk * 8))
    # the int8 MFMA work the estimator needs: 2 * sum D_s (quantised segments) per (query, probed
    # vector) pair; issued ops also count int8x2's second term and the 9+-bit codes' second plane
    pairs = float((cnt * sizes).sum())
    k_alg = sum(d for d, b in plan if b > 0)
    b_ops = pairs * 2 * k_alg
    k_eff = sum(d * (2 if b > 8 else 1) for d, b in plan if b > 0)
    b_ops_issued = pairs * 2 * k_eff * (2 if bprec == saq_amd.PREC_INT8X2 else 1)
    b_gbs = b_bytes / (bkern_ms / 1e3) / 1e9
    b_tops = b_ops / (bkern_ms / 1e3) / 1e12
    # HBM rate in the reference's own layout: per scanned vector its codes (sum D_s b_s / 8), |o| and
    # ExFactor per segment and its id, once per work tile over its list
    ref_vec = sum(d * b // 8 for d, b in plan if b > 0) + 12 * len(plan) + 4
    b_ref_bytes = float((tiles * sizes).sum()) * ref_vec
    b_ref_gbs = b_ref_bytes / (bkern_ms / 1e3) / 1e9
    prec_name = "int8x2" if bprec == saq_amd.PREC_INT8X2 else "int8"
    batched = {"qps": round(bqps, 1), "recall_at_10": round(brec, 4),
               "nprobe": bnp if npq is None else "U" + str(W["npq_range"]), "precision": prec_name,
               "query_tile": bqt, "coarse_quantiser": "exact (reference order)" if args.exact_probe else "rocBLAS bf16 GEMM of hi/lo bf16 parts (hi.hi + hi.lo + lo.hi, fp32 accumulation)",
               "recall_below_plain": below,
               "ms_per_step": round(bdt / steps * 1e3, 4), "layout_build_ms": round(t_prep * 1e3, 1),
               "kernel_ms": round(bkern_ms, 4), "layout_bytes_per_launch": b_bytes,
               "layout_hbm_gbs": round(b_gbs, 1), "int8_mfma_tops": round(b_tops, 2),
               "int8_mfma_frac": round(b_tops / INT8_PEAK_TOPS, 5),
               "int8_mfma_issued_tops": round(b_ops_issued / (bkern_ms / 1e3) / 1e12, 2),
               "pairs_per_launch": pairs}
    log(f"batched: {bqps:.0f} QPS ({bqps / qps:.2f}x plain), tile kernel {bkern_ms:.3f} ms, {b_gbs:.0f} GB/s, "
        f"{b_tops:.1f} TOPS")
    btraffic = read_traffic("bscan_traffic.json", n, bits, bnp, precision=prec_name, query_tile=bqt)

    precision = None if args.no_precision else precision_report(
        torch, saq_amd, make_plan, sdata, base, queries, cent, sizes, var, cid32, npq_np, args.topk, gt, dev,
        stream, steps, log)

    cpu = None
    if not args.no_cpu:
        # the host cores this process may use: the box's CPU share per GPU (OMP_NUM_THREADS, set
        # there to it), else every core
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0) or 0) or os.cpu_count() or 1
        sample = np.arange(W["cpu_queries"] or nq)
        qs_np = queries.cpu().numpy()[sample]
        probes = ix.probe(qs_np, nprobe)
        try:
            cq, cwall, cdt, passes, mism, kind = cpu_baseline(ix, plan, qs_np, var, rots, probes,
                                                              npq_np[sample] if npq_np is not None else None,
                                                              args.topk, 4.0, ids_gpu[sample], threads)
            cpu = {"value": round(cq, 1), "unit": "queries/s", "cores": threads, "kind": "port",
                   "cpu_model": cpu_model(), "wall_clock_qps": round(cwall, 1),
                   "sample": f"{passes} passes over {len(sample)} of the {nq} queries with the GPU's probe lists, "
                             f"{cdt:.1f}s; {kind}, one query per thread at a time on {threads} threads; value = "
                             f"test_qps's NQ * 1e3 / sum(latency ms) * threads, wall_clock_qps = NQ / elapsed; "
                             f"queries whose IDs differ from the plain GPU scan: {mism}"}
        except Exception as e:  # the baseline is reported, never the measured path
            log(f"cpu baseline failed: {e}")

    # ---- serving loop: latency under Poisson arrivals (dynamic batching over the batched scan), on
    # the workload's own index when it defines loads, else on the serving configuration's index
    serving = None
    if not args.no_serve:
        if W.get("serve_loads"):
            serving = {"workload": args.workload,
                       "rows": serve(torch, ix, queries, W["serve_loads"], args.serve_seconds, 5, bprec,
                                     args.exact_probe, log)}
        else:
            ix.close()
            ix = None
            del base
            torch.cuda.empty_cache()
            serving = serving_config(torch, saq_amd, make_plan, sdata, dev, args.serve_seconds, saq_amd.PREC_INT8,
                                     log)

    out = {
        "metric": METRIC,
        "value": round(bqps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(bdt / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8-mfma/i32/f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: batched SAQ IVF scan vs plain GPU scan, N={n} D={D} "
                               f"nlist={nlist} B={bits} nq={nq} k={args.topk}",
                   "bits": bits, "plan": plan, "recall_at_10_by_bits": bits_tried,
                   "batched": batched, "plain": plain, "batched_vs_plain": round(bqps / qps, 3),
                   "recall_note": (f"recall@10 reached: plain {plain['recall_at_10']}, batched {brec:.4f}; both scans "
                                   f"run at the same probe lists" + (
                                       " (nprobe drawn from U" + str(W["npq_range"]) + "); the 0.95 of the metric "
                                       "is not reachable there: the configuration probes at most 256 of "
                                       f"{nlist} Zipf-sized lists" if W["npq_range"] else "")),
                   "serving": serving, "p99_ms": serving["rows"][-1]["p99_ms"] if serving else None,
                   "encode_mvec_per_s": round(n / t_enc / 1e6, 3), "single_vector": single,
                   "precision_b8": precision,
                   "estimate": est},
        "roofline": {"bound": "mfma", "achieved": round(b_tops, 2), "peak": INT8_PEAK_TOPS, "unit": "TOPS",
                     "frac": round(b_tops / INT8_PEAK_TOPS, 5), "traffic": btraffic,
                     "kernel": "bscan_kernel", "alg_ops_per_launch": b_ops,
                     "alg_ops": "2 * sum_{b_s>0} D_s per (query, probed vector) pair",
                     "kernel_ms": round(bkern_ms, 4),
                     "hbm": {"achieved": round(b_ref_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(b_ref_gbs / HBM_PEAK_GBS, 4), "ref_layout_bytes_per_launch": b_ref_bytes,
                             "bytes": "per work tile, its list's vectors in the reference layout "
                                      "(codes, |o| and ExFactor per segment, id)"}},
        "cpu_baseline": cpu,
    }
    log(f"total {time.time() - t_all:.0f}s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
