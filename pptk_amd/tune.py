"""Autotuner for the batched path: bucket (leaf) size and query block width.

pptk fixes its bucket size by hand (BuildParams::num_leaf_points, default 10;
pptk/kdtree/src/kdtree.h:29-42). Here the bucket size trades tree depth (and
build launches) against points scanned per leaf, and the block width of the
fused query kernel trades occupancy against the divergence of neighbouring
lanes; neither changes a result. The tuner times one batch plan per
(leaf, block) pair on a sample of the workload and keeps the fastest.

Results are cached per (device, dim, dtype, k, normals_k, size profile) in a
JSON file ($PPTK_TUNE_CACHE, default ~/.cache/pptk_amd/tune.json).
"""
import json
import os

import numpy as np

LEAVES = (8, 12, 16, 24, 32, 48, 64, 96, 128, 160, 192, 224, 256)
BLOCKS = (64, 128, 256)


def _cache_path():
    return os.environ.get("PPTK_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pptk_amd",
                                       "tune.json"))


def profile_key(sizes, dim, dtype, k, normals_k, leaves=LEAVES, blocks=BLOCKS):
    """Key of a workload: device, shape parameters, the size distribution
    in octaves (clouds per power-of-two size class) and the candidate set."""
    import torch
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    sz = np.asarray(sizes, np.int64)
    hist = np.bincount(np.log2(np.maximum(sz, 1)).astype(int), minlength=32)
    frac = np.round(hist / max(len(sz), 1), 2)
    return "%s|dim%d|%s|k%d|nk%d|%s|L%s|B%s" % (
        dev, dim, np.dtype(dtype).name, k, normals_k, ",".join("%g" % f for f in frac),
        ",".join(map(str, leaves)), ",".join(map(str, blocks)))


def load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass


def tune_batch(pts_d, sizes, *, dim=3, dtype=np.float32, k=16, normals_k=16, dmax=np.inf,
               leaves=LEAVES, blocks=BLOCKS, reps=2, stream=None, log=None, use_cache=True):
    """Times every (leaf, block) pair on `pts_d` (a torch device tensor of the
    concatenated clouds with the given sizes) and returns (best, table); each
    table row is {"leaf", "block", "ms"} with the best of `reps` graph
    replays."""
    import torch

    from . import capi

    key = profile_key(sizes, dim, dtype, k, normals_k, leaves, blocks)
    if use_cache:
        hit = load_cache().get(key)
        if hit is not None:
            return hit["best"], hit["table"]
    n = int(np.sum(sizes))
    dev = pts_d.device
    st = stream if stream is not None else torch.cuda.Stream(device=dev)
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    idx = torch.empty((max(n, 1), max(k, 1)), dtype=torch.int32, device=dev)
    nrm = torch.empty((max(n, 1), 3), dtype=tdt, device=dev)
    torch.cuda.synchronize()
    sh = st.cuda_stream
    table = []
    for leaf in leaves:
        for qb in blocks:
            plan = capi.Batch(sizes, dim=dim, dtype=dtype, k=k, normals_k=normals_k, dmax=dmax,
                              leaf=leaf, use_graph=True, query_block=qb)
            args = (pts_d.data_ptr(), idx.data_ptr() if k else None, None,
                    nrm.data_ptr() if normals_k else None, None)
            plan.run(*args, stream=sh)  # capture + first replay
            best = None
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(st)
                plan.run(*args, stream=sh)
                e1.record(st)
                st.synchronize()
                t = e0.elapsed_time(e1)
                best = t if best is None else min(best, t)
            plan.free()
            table.append({"leaf": leaf, "block": qb, "ms": best})
            if log:
                log("tune: leaf %2d block %3d  %.3f ms" % (leaf, qb, best))
    best = min(table, key=lambda r: r["ms"])
    if use_cache:
        cache = load_cache()
        cache[key] = {"best": best, "table": table}
        save_cache(cache)
    return best, table
