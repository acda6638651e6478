"""The reference's index file format, restated for the tests (test infrastructure).

IVF::save / IVF::load (saqlib/index/ivf.hpp:197-264) write, in order:
  size_t num_data, num_dim, num_cen
  float centroids[num_cen][num_dim]                         FlatInitializer::save (initializer.hpp:78-84)
  SaqData (saq_data.hpp:29-39):
    QuantizeConfig, raw (32 bytes, layout below)
    size_t num_dim; size_t n_var; float variance[n_var]     (utils/IO.hpp:31-35; padded dimension)
    size_t num_seg; per segment (quantizer_data.hpp:26-35):
      size_t dim_pad; size_t bits; QuantSingleConfig, raw (20 bytes); char has_rotator;
      float P[dim_pad][dim_pad] when has_rotator           (rotator.hpp:55-63, row-major)
  size_t cluster_sizes[num_cen]
  per cluster, SaqCluData::save (cluster_data.hpp:327-336):
    short factors, short codes, longb_code_bytes_ * num_vec long-code bytes, ExFactor[num_vec][num_seg],
    PID ids[num_vec], then every segment's rotated centroid
The struct layouts are pinned against the reference's compiled config.h (oracle/_ref).
"""
import struct

import numpy as np

CFG_BYTES, SINGLE_CFG_BYTES, SINGLE_OFF = 32, 20, 12
CFG_FIELDS = {"avg_bits": (0, "<f4"), "seg_eqseg": (4, "<i4"), "enable_segmentation": (8, "u1"),
              "use_compact_layout": (9, "u1")}
SINGLE_FIELDS = {"quant_type": (0, "<i4"), "random_rotation": (4, "u1"), "use_fastscan": (5, "u1"),
                 "caq_adj_rd_lmt": (8, "<i4"), "caq_adj_eps": (12, "<f4"), "caq_ori_qB": (16, "<i4")}


def pack_fields(nbytes, fields, values):
    b = bytearray(nbytes)
    for k, (off, dt) in fields.items():
        v = np.array([values.get(k, 0)], dtype=dt).tobytes()
        b[off:off + len(v)] = v
    return bytes(b)


def unpack_fields(b, fields):
    return {k: np.frombuffer(b, dtype=dt, count=1, offset=off)[0].item() for k, (off, dt) in fields.items()}


def file_long_bytes(plan, n, compact):
    """SaqCluData::save writes longb_code_bytes_ * num_vec_ long-code bytes: the padded per-vector
    stride in the per-vector layout, the unpadded segment sum in the compact one (cluster_data.hpp:220-229)."""
    compact = compact or len(plan) == 1
    per = 0
    for d, b in plan:
        lb = d * (b - 1) // 8 if b > 1 else 0
        per += lb if compact else (lb + 15) // 16 * 16
    return per * n


def write_index(path, num_dim, plan, centroids, variance_pad, rotations, clusters, cfg):
    K = centroids.shape[0]
    sizes = [int(c["num_vec"]) for c in clusters]
    with open(path, "wb") as f:
        f.write(struct.pack("<3Q", sum(sizes), num_dim, K))
        f.write(np.ascontiguousarray(centroids, np.float32).tobytes())
        single = pack_fields(SINGLE_CFG_BYTES, SINGLE_FIELDS, cfg.get("single", {}))
        top = bytearray(pack_fields(CFG_BYTES, CFG_FIELDS, cfg))
        top[SINGLE_OFF:SINGLE_OFF + SINGLE_CFG_BYTES] = single
        f.write(bytes(top))
        f.write(struct.pack("<Q", num_dim))
        f.write(struct.pack("<Q", len(variance_pad)))
        f.write(np.ascontiguousarray(variance_pad, np.float32).tobytes())
        f.write(struct.pack("<Q", len(plan)))
        for s, (d, b) in enumerate(plan):
            f.write(struct.pack("<2Q", d, b))
            f.write(single)
            P = rotations[s] if rotations is not None else None
            f.write(bytes([1 if P is not None else 0]))
            if P is not None:
                f.write(np.ascontiguousarray(P, np.float32).tobytes())
        f.write(np.array(sizes, np.uint64).tobytes())
        compact = bool(cfg.get("use_compact_layout", 0))
        for c, clu in enumerate(clusters):
            n = sizes[c]
            f.write(clu["short_factors"].astype(np.float32).tobytes())
            f.write(clu["short_code"].astype(np.uint8).tobytes())
            f.write(clu["long_code"].astype(np.uint8).tobytes()[:file_long_bytes(plan, n, compact)])
            f.write(clu["long_factors"].astype(np.float32).tobytes())
            f.write(clu["ids"][:n].astype(np.uint32).tobytes())
            f.write(clu["centroid_rot"].astype(np.float32).tobytes())


def read_index(path, cluster_sizes_of):
    """cluster_sizes_of(plan, n, compact) -> (short_factor_floats, short_code_bytes, long_factor_floats)."""
    data = open(path, "rb").read()
    pos = 0

    def take(nbytes):
        nonlocal pos
        b = data[pos:pos + nbytes]
        assert len(b) == nbytes, "file truncated"
        pos += nbytes
        return b

    num_data, num_dim, K = struct.unpack("<3Q", take(24))
    out = {"num_data": num_data, "num_dim": num_dim, "K": K}
    out["centroids"] = np.frombuffer(take(4 * K * num_dim), np.float32).reshape(K, num_dim)
    cfgb = take(CFG_BYTES)
    out["cfg"] = unpack_fields(cfgb, CFG_FIELDS)
    out["cfg"]["single"] = unpack_fields(cfgb[SINGLE_OFF:SINGLE_OFF + SINGLE_CFG_BYTES], SINGLE_FIELDS)
    out["saq_num_dim"], = struct.unpack("<Q", take(8))
    nvar, = struct.unpack("<Q", take(8))
    out["variance"] = np.frombuffer(take(4 * nvar), np.float32)
    nseg, = struct.unpack("<Q", take(8))
    plan, rots = [], []
    for _ in range(nseg):
        d, b = struct.unpack("<2Q", take(16))
        out.setdefault("single_cfgs", []).append(unpack_fields(take(SINGLE_CFG_BYTES), SINGLE_FIELDS))
        flag = take(1)[0]
        plan.append((d, b))
        rots.append(np.frombuffer(take(4 * d * d), np.float32).reshape(d, d) if flag else None)
    out["plan"], out["rotations"] = plan, rots
    sizes = np.frombuffer(take(8 * K), np.uint64).astype(np.int64)
    out["sizes"] = sizes
    compact = bool(out["cfg"]["use_compact_layout"])
    d_pad = sum(d for d, _ in plan)
    out["clusters"] = []
    for c in range(K):
        n = int(sizes[c])
        sff, scb, lff = cluster_sizes_of(plan, n, compact)
        clu = {"short_factors": np.frombuffer(take(4 * sff), np.float32),
               "short_code": np.frombuffer(take(scb), np.uint8),
               "long_code": np.frombuffer(take(file_long_bytes(plan, n, compact)), np.uint8),
               "long_factors": np.frombuffer(take(4 * lff), np.float32),
               "ids": np.frombuffer(take(4 * n), np.uint32),
               "centroid_rot": np.frombuffer(take(4 * d_pad), np.float32)}
        out["clusters"].append(clu)
    assert pos == len(data), "trailing bytes"
    return out
