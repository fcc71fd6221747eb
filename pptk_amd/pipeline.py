"""One pass of the hot path over a device-resident cloud.

The sequence a caller with points already in HBM runs, through the C ABI,
with every output left in device memory (torch tensors):
  1. build the tree from the device point array
     (kdtree._build, pptk/kdtree/kdtree_wrapper.cpp:281-304);
  2. kNN self query of every point (kdtree._query(tree, None, k),
     kdtree-impl.h:957-976): idx (n, k) int32, counts (n,);
  3. r-near of m query points (kdtree._query(tree, queries, k=None, dmax),
     kdtree-impl.h:1008-1026): CSR offsets (m+1,) int64 + lists int32;
  4. PCA normals of every point on the resident tree
     (estimate_normals, estimate_normals.cpp:64-111): (n, 3).
By default steps 2 and 4 are one call, pptk_knn_normals, which writes both
outputs from one search per point (bit for bit what the two separate calls
write); fused=False makes the two separate calls.

`bench.py` times `LidarStep.run` and the full-size parity test checks its
outputs, so the numbers and the parity claims cover the same calls.
"""
import ctypes

import numpy as np

from . import capi


class LidarStep:
    """Build + kNN + r-near + normals on device arrays.

    pts_d: (n, 3) float32 torch tensor on the GPU; q_d: (m, 3) float32 query
    points on the GPU (or None to skip r-near). Outputs: idx_d, cnt_d, off_d,
    ridx_d (the first `rtotal` entries are valid), nrm_d."""

    PHASES_FUSED = ("build", "knn_normals", "rnear_count", "rnear_fill")
    PHASES_SEPARATE = ("build", "knn", "rnear_count", "rnear_fill", "normals")

    def __init__(self, pts_d, q_d, k=16, dmax=0.1, normals_k=None, leaf=0,
                 precision=capi.PREC_MIXED, stream=None, fused=True):
        import torch
        self.torch = torch
        self.dev = pts_d.device
        self.n = int(pts_d.shape[0])
        self.k = int(k)
        self.nk = int(normals_k if normals_k is not None else k)
        self.dmax = float(np.float32(dmax))  # kdtree._query rounds dmax through float32
        self.leaf = int(leaf)
        self.precision = precision
        self.fused = bool(fused)
        self.PHASES = self.PHASES_FUSED if self.fused else self.PHASES_SEPARATE
        self.pts_d = pts_d
        self.q_d = q_d
        self.m = 0 if q_d is None else int(q_d.shape[0])
        self.stream = stream if stream is not None else torch.cuda.Stream(device=self.dev)
        self.sh = ctypes.c_void_p(self.stream.cuda_stream)
        n, m = self.n, self.m
        self.idx_d = torch.empty((n, self.k), dtype=torch.int32, device=self.dev)
        self.cnt_d = torch.empty(n, dtype=torch.int32, device=self.dev)
        self.nrm_d = torch.empty((n, 3), dtype=torch.float32, device=self.dev)
        self.off_d = torch.empty(m + 1, dtype=torch.int64, device=self.dev)
        self.ridx_d = None
        self.rtotal = 0
        self.num_fp64 = 0
        self.leaf_size = None
        self._qall, _ = capi.make_queries(None, mem=capi.MEM_DEVICE)
        self._qall.kind = capi.Q_ALL
        if m:
            self._qpt, _ = capi.make_queries(dev_ptr=q_d.data_ptr(), m=m, kind=capi.Q_POINTS,
                                             dtype=np.float32)
        torch.cuda.synchronize()

    def run(self, events=None):
        """One pass. With `events` (a list of len(PHASES)+1 timing events)
        each phase is bracketed by events recorded on the library's stream."""
        torch, L, sh = self.torch, capi.lib(), self.sh
        ev = events if events is not None else [None] * (len(self.PHASES) + 1)

        def mark(i):
            if ev[i] is not None:
                ev[i].record(self.stream)

        nfb = ctypes.c_int64(0)
        total = ctypes.c_int64(0)
        idx_p = ctypes.c_void_p(self.idx_d.data_ptr())
        cnt_p = ctypes.c_void_p(self.cnt_d.data_ptr())
        nrm_p = ctypes.c_void_p(self.nrm_d.data_ptr())
        with torch.cuda.stream(self.stream):
            mark(0)
            tree = capi.Tree.build(dev_ptr=self.pts_d.data_ptr(), n=self.n, dim=3,
                                   dtype=np.float32, leaf=self.leaf, stream=sh)
            try:
                h = ctypes.c_void_p(tree.handle)
                mark(1)
                if self.fused:
                    capi.check(L.pptk_knn_normals(h, self.k, float("inf"), self.nk, float("inf"),
                                                  1, self.precision, idx_p, cnt_p, None, nrm_p,
                                                  None, None, ctypes.byref(nfb), capi.MEM_DEVICE,
                                                  sh))
                else:
                    capi.check(L.pptk_knn(h, ctypes.byref(self._qall), self.k, float("inf"),
                                          idx_p, cnt_p, None, capi.MEM_DEVICE, sh))
                mark(2)
                if self.m:
                    capi.check(L.pptk_rnear_count(h, ctypes.byref(self._qpt), self.dmax,
                                                  ctypes.c_void_p(self.off_d.data_ptr()),
                                                  ctypes.byref(total), capi.MEM_DEVICE, sh))
                mark(3)
                if self.m:
                    if self.ridx_d is None or self.ridx_d.numel() < total.value:
                        self.ridx_d = torch.empty(max(1, int(total.value * 1.1)),
                                                  dtype=torch.int32, device=self.dev)
                    capi.check(L.pptk_rnear_fill(h, ctypes.byref(self._qpt), self.dmax,
                                                 ctypes.c_void_p(self.off_d.data_ptr()),
                                                 ctypes.c_void_p(self.ridx_d.data_ptr()),
                                                 capi.MEM_DEVICE, sh))
                mark(4)
                if not self.fused:
                    capi.check(L.pptk_normals(h, self.nk, float("inf"), None, 0, 1,
                                              self.precision, nrm_p, None, None,
                                              ctypes.byref(nfb), capi.MEM_DEVICE, sh))
                    mark(5)
                if self.leaf_size is None:
                    self.leaf_size = tree.info.leaf_size
            finally:
                tree.free()
        self.rtotal = int(total.value)
        self.num_fp64 = int(nfb.value)

    def host_outputs(self):
        """(idx, cnt, off, lists, normals) copied to host numpy arrays."""
        self.stream.synchronize()
        lists = (self.ridx_d[: self.rtotal].cpu().numpy() if self.ridx_d is not None
                 else np.zeros(0, np.int32))
        return (self.idx_d.cpu().numpy(), self.cnt_d.cpu().numpy(), self.off_d.cpu().numpy(),
                lists, self.nrm_d.cpu().numpy())
