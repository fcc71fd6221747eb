"""Benchmark: k-d tree build + kNN + r-near + normals on the 10M-point lidar scene.

One step = one pass of the hot path over the synthetic 10M-point float32
lidar scene (sensor at the origin, ground + 40 walls + 200 cylinders, density
~ 1/range^2, 1 cm noise; pptk_amd.scenes.lidar_scene, seed 1), inputs already
resident in HBM:
  1. build the tree from the device point array;
  2. kNN k=16 self query for every point (idx + counts to HBM);
  3. r-near dmax=0.1 m for 1M query points (sampled without replacement,
     seed 2, queried as points; CSR to HBM);
  4. PCA normals k=16, r=inf for every point on the resident tree.
value = (kNN queries + normals) per second of step time.

Per-phase durations come from HIP events recorded on the stream the library
launches on. The CPU baseline runs pptk's own k-d tree (compiled from the
reference sources, oracle/_ref) or, when that is absent, the oracle's
restatement, on a bounded sample, on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("kNN queries/s + normals/s at 10M pts vs plain GPU & CPU; "
          "served p50/p99 latency")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n, k, m_rnear, rnear_total):
    s = 4  # float32 coordinates
    return {
        # read points, write reordered coordinates + permutation
        "build": n * (3 * s + 3 * s + 4),
        # read query coordinates, write k int32 indices + 1 int32 count
        "knn": n * (3 * s + 4 * k + 4),
        # read query, write its int32 count and int64 offset
        "rnear_count": m_rnear * (3 * s + 4 + 8),
        # read query and its offsets, write the int32 neighbour ids
        "rnear_fill": m_rnear * (3 * s + 16) + 4 * rnear_total,
        # read point, write the unit normal
        "normals": n * (3 * s + 3 * s),
    }


# The kernel that dominates each phase (names as rocprofv3 reports them).
PHASE_KERNEL = {
    "knn": "knn_kernel<float, 3, 16>",
    "rnear_count": "rnear_kernel<float, 3, false>",
    "rnear_fill": "rnear_kernel<float, 3, true>",
    "normals": "normals_knn_kernel<float, 16, 1, 1>",
}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json (rocprofv3
    FETCH_SIZE and WRITE_SIZE runs scaled by a plain-copy calibration)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)["kernels"][kernel]
        return t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(pts, qpts, k, dmax, gpu_knn_rows=None):
    """Time pptk's k-d tree (or the oracle port) on a bounded sample, and
    compare its kNN rows with the GPU's rows for the same queries."""
    from oracle import oracle as O
    cores = min(16, os.cpu_count() or 1)
    n = len(pts)
    kind = "reference" if O.ref_available() else "port"
    t0 = time.perf_counter()
    tree = O.RefTree(pts) if kind == "reference" else O.OracleTree(pts)
    t_build = time.perf_counter() - t0
    mq = min(n, 1_000_000)
    t0 = time.perf_counter()
    if kind == "reference":
        cpu_rows, _ = tree.knn(None, k, nthreads=cores, m_self=mq)
    else:
        cpu_rows, _ = tree.knn(np.arange(mq), k, nthreads=cores)
    t_knn = time.perf_counter() - t0
    same = None
    if gpu_knn_rows is not None:
        same = int(np.sum(np.all(gpu_knn_rows[:mq] == cpu_rows, axis=1)))
    mr = min(len(qpts), 200_000)
    t0 = time.perf_counter()
    tree.rnear(qpts[:mr].astype(pts.dtype), dmax, nthreads=cores)
    t_rn = time.perf_counter() - t0
    del tree
    port = O.OracleTree(pts)
    mn = min(n, 300_000)
    t0 = time.perf_counter()
    port.normals(k, np.inf, rows=np.arange(mn), nthreads=cores)
    t_nrm = time.perf_counter() - t0
    del port
    knn_rate = mq / t_knn
    nrm_rate = mn / t_nrm
    rn_rate = mr / t_rn
    step = t_build + n / knn_rate + len(qpts) / rn_rate + n / nrm_rate
    return {
        "value": 2 * n / step,
        "unit": "(kNN queries + normals)/s",
        "cores": cores,
        "kind": kind,
        "sample": ("full %d-point tree build (%s, serial: %.2fs); kNN k=%d self queries on "
                   "%d points; r-near dmax=%g on %d query points; normals k=%d on %d points "
                   "(oracle port, FP64 Jacobi); per-phase rates scaled to the full workload"
                   % (n, "pptk KdTree" if kind == "reference" else "oracle port", t_build, k,
                      mq, dmax, mr, k, mn)),
        "build_s": t_build,
        "knn_qps": knn_rate,
        "normals_ps": nrm_rate,
        "rnear_qps": rn_rate,
        "knn_rows_identical_to_cpu": same,
        "knn_rows_compared": mq,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--dmax", type=float, default=0.1)
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported (queries and trees "
              "shard by replication / by cloud, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from pptk_amd import capi, scenes

    n, k, dmax = args.n, args.k, float(np.float32(args.dmax))
    log("bench: generating lidar scene n=%d" % n)
    pts = scenes.lidar_scene(n, seed=1)
    rng = np.random.default_rng(2)
    qsel = rng.choice(n, size=min(args.queries, n), replace=False)
    qpts = np.ascontiguousarray(pts[qsel])
    m = len(qpts)

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    sh = ctypes.c_void_p(stream.cuda_stream)
    pts_d = torch.from_numpy(pts).to(dev)
    q_d = torch.from_numpy(qpts).to(dev)
    idx_d = torch.empty((n, k), dtype=torch.int32, device=dev)
    cnt_d = torch.empty(n, dtype=torch.int32, device=dev)
    nrm_d = torch.empty((n, 3), dtype=torch.float32, device=dev)
    off_d = torch.empty(m + 1, dtype=torch.int64, device=dev)
    ridx_d = None
    torch.cuda.synchronize()
    L = capi.lib()
    qall, _ = capi.make_queries(None, mem=capi.MEM_DEVICE)
    qall.kind = capi.Q_ALL
    qpt, _ = capi.make_queries(dev_ptr=q_d.data_ptr(), m=m, kind=capi.Q_POINTS,
                               dtype=np.float32)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    phases = ["build", "knn", "rnear_count", "rnear_fill", "normals"]
    state = {"rtotal": 0, "nfb": 0}

    def step(record):
        nonlocal ridx_d
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(phases) + 1)]
        with torch.cuda.stream(stream):
            ev[0].record(stream)
            tree = capi.Tree.build(dev_ptr=pts_d.data_ptr(), n=n, dim=3, dtype=np.float32,
                                   leaf=args.leaf, stream=sh)
            ev[1].record(stream)
            capi.check(L.pptk_knn(ctypes.c_void_p(tree.handle), ctypes.byref(qall), k,
                                  float("inf"), ctypes.c_void_p(idx_d.data_ptr()),
                                  ctypes.c_void_p(cnt_d.data_ptr()), None, capi.MEM_DEVICE, sh))
            ev[2].record(stream)
            total = ctypes.c_int64()
            capi.check(L.pptk_rnear_count(ctypes.c_void_p(tree.handle), ctypes.byref(qpt), dmax,
                                          ctypes.c_void_p(off_d.data_ptr()), ctypes.byref(total),
                                          capi.MEM_DEVICE, sh))
            ev[3].record(stream)
            if ridx_d is None or ridx_d.numel() < total.value:
                ridx_d = torch.empty(max(1, int(total.value * 1.1)), dtype=torch.int32,
                                     device=dev)
            capi.check(L.pptk_rnear_fill(ctypes.c_void_p(tree.handle), ctypes.byref(qpt), dmax,
                                         ctypes.c_void_p(off_d.data_ptr()),
                                         ctypes.c_void_p(ridx_d.data_ptr()), capi.MEM_DEVICE,
                                         sh))
            ev[4].record(stream)
            nfb = ctypes.c_int64()
            capi.check(L.pptk_normals(ctypes.c_void_p(tree.handle), k, float("inf"), None, 0, 1,
                                      prec, ctypes.c_void_p(nrm_d.data_ptr()), None, None,
                                      ctypes.byref(nfb), capi.MEM_DEVICE, sh))
            ev[5].record(stream)
        tree.free()
        state["rtotal"] = int(total.value)
        state["nfb"] = int(nfb.value)
        if record is not None:
            stream.synchronize()
            for i, p in enumerate(phases):
                record.setdefault(p, []).append(ev[i].elapsed_time(ev[i + 1]))

    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step(None)
    torch.cuda.synchronize()
    rec = {}
    log("bench: timing %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(rec)
        log("bench: step %d done" % i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    ph = {p: float(np.mean(v)) for p, v in rec.items()}

    ab = algorithmic_bytes(n, k, m, state["rtotal"])
    dom = max(PHASE_KERNEL, key=lambda p: ph[p])
    achieved = ab[dom] / (ph[dom] * 1e-3) / 1e9
    traffic = measured_traffic(PHASE_KERNEL[dom])
    out = {
        "metric": METRIC,
        "value": 2 * n / (ms_step * 1e-3),
        "unit": "(kNN queries + normals)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("lidar scene %d float32 points: build + kNN k=%d self query + r-near "
                         "dmax=%g m on %d query points + normals k=%d r=inf (%s)"
                         % (n, k, args.dmax, m, k, args.precision)),
            "n_points": n, "k": k, "rnear_queries": m, "dmax": args.dmax,
            "leaf_size": capi.Tree.build(pts[:1000]).info.leaf_size if args.leaf == 0
            else args.leaf,
        },
        "phases_ms": ph,
        "knn_qps": n / (ph["knn"] * 1e-3),
        "normals_ps": n / (ph["normals"] * 1e-3),
        "rnear_qps": m / ((ph["rnear_count"] + ph["rnear_fill"]) * 1e-3),
        "build_pts_per_s": n / (ph["build"] * 1e-3),
        "rnear_neighbours": state["rtotal"],
        "normals_fp64_fallback_frac": state["nfb"] / n,
        "roofline": {
            "bound": "hbm",
            "kernel": PHASE_KERNEL[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": ab[dom],
        },
    }
    if not args.no_cpu:
        log("bench: CPU baseline")
        rows = idx_d[: min(n, 1_000_000)].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(pts, qpts, k, dmax, gpu_knn_rows=rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
