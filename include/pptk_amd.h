/*
 * pptk_amd.h -- C ABI of the MI355X (gfx950) k-d tree and normal-estimation
 * library `libpptk_amd.so`.
 *
 * Plain C: no HIP, torch or numpy types cross this boundary. Buffers are raw
 * pointers plus sizes; a `pptk_mem` argument says whether a buffer lives in
 * host memory (the library copies it) or in device memory (used in place).
 * Every entry point returns a `pptk_status`; on failure `pptk_last_error()`
 * returns a message for the calling thread. Status codes map one-to-one onto
 * the Python exceptions raised by pptk's CPython extensions:
 *   PPTK_E_VALUE -> ValueError, PPTK_E_TYPE -> TypeError,
 *   PPTK_E_RUNTIME / PPTK_E_DEVICE -> RuntimeError, PPTK_E_MEMORY ->
 *   MemoryError, PPTK_E_NOTIMPL -> NotImplementedError.
 *
 * Results contract (identical to pptk's CPU k-d tree on the same input):
 *   - squared distance d = ((0 + t0*t0) + t1*t1) + t2*t2 with
 *     t_i = (DistT)p_i - (DistT)q_i, evaluated in DistT without FMA
 *     contraction (pptk/kdtree/src/kdtree-impl.h:331-339). DistT is double for
 *     float64 points and float for float32 and every integer dtype
 *     (pptk/kdtree/src/accumulator.h:12-27).
 *   - a neighbour qualifies iff d < r*r with r = (DistT)dmax
 *     (kdtree-impl.h:365, 881).
 *   - kNN rows hold the k smallest (d, original index) pairs in ascending
 *     order; ties in d are broken by the smaller original point index.
 *   - self queries exclude the query point itself (kdtree-impl.h:931-947).
 */
#ifndef PPTK_AMD_H_
#define PPTK_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PPTK_OK = 0,
  PPTK_E_VALUE = 1,
  PPTK_E_TYPE = 2,
  PPTK_E_RUNTIME = 3,
  PPTK_E_MEMORY = 4,
  PPTK_E_DEVICE = 5,
  PPTK_E_NOTIMPL = 6
} pptk_status;

/* Point / query element types: the ten dtypes pptk's kd-tree dispatches on
 * (pptk/kdtree/kdtree_wrapper.cpp:15-41). */
typedef enum {
  PPTK_F32 = 0,
  PPTK_F64 = 1,
  PPTK_I8 = 2,
  PPTK_I16 = 3,
  PPTK_I32 = 4,
  PPTK_I64 = 5,
  PPTK_U8 = 6,
  PPTK_U16 = 7,
  PPTK_U32 = 8,
  PPTK_U64 = 9
} pptk_dtype;

typedef enum { PPTK_MEM_HOST = 0, PPTK_MEM_DEVICE = 1 } pptk_mem;

/* Opaque device-resident k-d tree. Owns its device buffers. Immutable after
 * build, so concurrent queries on one tree are safe. */
typedef struct pptk_tree pptk_tree;

/* Build parameters. Replaces pointkd::BuildParams (pptk/kdtree/src/kdtree.h:29-42).
 * max_leaf_size: points per leaf bucket; <= 0 picks the tuned default,
 *   otherwise clamped to [1, 256]. The bucket size changes speed only, never
 *   results.
 * num_proc, empty_split: accepted for signature parity and ignored (the GPU
 *   tree uses tight per-node boxes, which trim empty space without
 *   empty-split nodes). */
typedef struct {
  int max_leaf_size;
  int num_proc;
  double empty_split;
} pptk_build_params;

/* Query set for pptk_knn / pptk_rnear_*. Replaces the three query forms of
 * kdtree._query (kdtree_wrapper.cpp:161-248):
 *   PPTK_Q_ALL     : original indices 0 .. num_valid-1 (queries=None,
 *                    kdtree_wrapper.cpp:176-179), self queries;
 *   PPTK_Q_INDICES : m original indices (already range-checked and made
 *                    non-negative by the caller), self queries
 *                    (kdtree_wrapper.cpp:237-239);
 *   PPTK_Q_POINTS  : m x dim row-major points of `dtype`
 *                    (kdtree_wrapper.cpp:186-236). With mem == DEVICE the
 *                    dtype must equal the tree's distance type. */
typedef enum { PPTK_Q_ALL = 0, PPTK_Q_INDICES = 1, PPTK_Q_POINTS = 2 } pptk_query_kind;

typedef struct {
  pptk_query_kind kind;
  int64_t m;
  const int32_t* indices;
  const void* points;
  pptk_dtype dtype;
  pptk_mem mem;
} pptk_queries;

typedef struct {
  int dim;
  pptk_dtype dtype;      /* dtype the tree was built from */
  pptk_dtype dist_dtype; /* PPTK_F32 or PPTK_F64 */
  int64_t num_input;     /* rows passed to build */
  int64_t num_valid;     /* rows without NaN/inf */
  int leaf_size;         /* maximum points per leaf */
  int64_t num_leaves;
  int64_t num_nodes;
  int depth;             /* number of node levels */
  double bbox_lo[4];
  double bbox_hi[4];
  int64_t device_bytes;  /* device memory held by the tree */
} pptk_tree_info;

/* Message of the last failure on this thread ("" if none). */
const char* pptk_last_error(void);

/* Number of visible HIP devices (0 on a machine without a GPU). */
int pptk_device_count(void);

/* Release device memory cached by the library's allocator. */
pptk_status pptk_trim(void);

/* Build a tree. Replaces kdtree._build -> KdTree ctor -> impl::BuildTree
 * (kdtree_wrapper.cpp:281-304, kdtree-impl.h:773-818): NaN/inf rows are
 * excluded (kdtree-impl.h:53-75), the tree keeps its own reordered copy.
 * pts: n x dim row-major, contiguous. dim in {2,3,4}. stream: hipStream_t or
 * NULL for the library stream. Returns after the tree is complete. */
pptk_status pptk_tree_build(const void* pts, int64_t n, int dim, pptk_dtype dtype,
                            const pptk_build_params* params, pptk_mem mem,
                            void* stream, pptk_tree** out);

/* Free a tree (the capsule destructor, kdtree_wrapper.cpp:57-74). NULL is a no-op. */
void pptk_tree_free(pptk_tree* tree);

/* 1 if `p` is a live tree handle produced by this library, else 0. */
int pptk_tree_check(const void* p);

pptk_status pptk_tree_get_info(const pptk_tree* tree, pptk_tree_info* out);

/* Copy the tree's arrays out, mirroring KdTree::points()/indices()/
 * reverse_indices() (kdtree.h:231-234). Any pointer may be NULL.
 *   points  : num_valid x dim in tree order, dist dtype
 *   reverse : num_valid int32, tree position -> original index
 *   forward : num_input int32, original index -> tree position or -1
 *   boxes   : num_nodes x 2 x dim dist dtype, node i = [lo(dim), hi(dim)]
 *   nodes   : num_nodes x 4 int32, node i = [a, b, child, parent]: tree
 *             positions [a, b), first child (children are child, child+1)
 *             or -1 for a leaf, parent or -1 for the root (node 0). */
pptk_status pptk_tree_export(const pptk_tree* tree, void* points, int32_t* reverse,
                             int32_t* forward, void* boxes, int32_t* nodes, pptk_mem mem);

/* Persistent tree format: the whole tree as one flat little-endian byte
 * buffer (header, coordinates, row maps, nodes, boxes, leaf map), restorable
 * without a rebuild. Stands where pptk's unused flat node array does:
 * SmallNode (small_node.h:13-45) from KdTree::SmallNodes()
 * (kdtree-impl.h:605-627, 863-871). serialize: buf NULL returns the size in
 * *out_size; otherwise writes it (capacity >= size). deserialize validates
 * every index on the host before any device upload and fails with
 * PPTK_E_VALUE on a damaged buffer. */
pptk_status pptk_tree_serialize(const pptk_tree* tree, void* buf, int64_t capacity,
                                int64_t* out_size);
pptk_status pptk_tree_deserialize(const void* buf, int64_t size, void* stream,
                                  pptk_tree** out);

/* k nearest neighbours with distance < dmax. Replaces
 * KdTree::KNearestNeighbors / KNearestNeighborsSelf batch calls
 * (kdtree-impl.h:873-976) behind kdtree._query(k>0).
 * dmax: pass the value already rounded through float32 as kdtree._query does
 * (pptk/include/python_util.h:393-400); INFINITY for no bound.
 * out_idx: m x k int32 (row-major), original indices ascending by (d, index),
 * -1 padded. out_cnt: m int32 neighbour counts. out_dist: NULL or m x k of the
 * tree's dist dtype (squared distances, +inf padded). For PPTK_Q_ALL, m is
 * the tree's num_valid. Output buffers live in `out_mem`. */
pptk_status pptk_knn(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                     int32_t* out_idx, int32_t* out_cnt, void* out_dist,
                     pptk_mem out_mem, void* stream);

/* Allocation-free kNN. No counterpart in pptk: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same results and arguments as pptk_knn
 * with queries (PPTK_Q_INDICES / PPTK_Q_POINTS) and outputs in device memory;
 * it calls no hipMalloc/hipFree. Its temporaries live in one 64 MiB
 * device array per GPU, held from the library's load to the process's end:
 * point and index query sets of 4096 or more queries are ordered in it, so
 * they are limited to 2,739,072 queries per call (PPTK_E_VALUE beyond, with
 * the limit in the message); k > 32 searches keep their heaps in it. Self
 * queries of every point (PPTK_Q_ALL) need no scratch and have no limit.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other threads, other streams) race and must be serialised by the
 * caller. Returns after the results are complete on `stream`. */
pptk_status pptk_knn_noalloc(const pptk_tree* tree, const pptk_queries* q, int k, double dmax,
                             int32_t* out_idx, int32_t* out_cnt, void* out_dist, void* stream);

/* r-near neighbours (d < dmax*dmax), two phases. Replaces
 * KdTree::RNearNeighbors / RNearNeighborsSelf (kdtree-impl.h:978-1069)
 * behind kdtree._query(k=None). Phase 1 writes CSR offsets (m+1 int64,
 * offsets[m] = total) and the total into *out_total (host). Phase 2 writes
 * the neighbour lists (original indices) given those offsets. Within a list,
 * neighbours appear in tree order; compare lists as sets. A list is written
 * only inside [offsets[i], offsets[i+1]) clamped to [0, total). With host
 * offsets a negative total is an error; with offsets and out_idx in device
 * memory the kernels read the total themselves (no wait for a copy to the
 * host) and a negative total writes nothing. The fill right after the count
 * of the same queries at the same dmax replays the count's traversal instead
 * of walking the tree again (same output). */
pptk_status pptk_rnear_count(const pptk_tree* tree, const pptk_queries* q, double dmax,
                             int64_t* out_offsets, int64_t* out_total, pptk_mem out_mem,
                             void* stream);
pptk_status pptk_rnear_fill(const pptk_tree* tree, const pptk_queries* q, double dmax,
                            const int64_t* offsets, int32_t* out_idx, pptk_mem out_mem,
                            void* stream);

/* Normal estimation precision modes. */
typedef enum {
  PPTK_PREC_AUTO = 0,  /* FP64 for float64 points, MIXED for float32 points */
  PPTK_PREC_MIXED = 1, /* FP64 covariance centred on the query point; FP32
                          eigensolve; FP64 eigensolve of the same covariance
                          where (l1-l0)/l2 < 1e-3 */
  PPTK_PREC_FP64 = 2   /* FP64 covariance and eigensolve */
} pptk_precision;

/* PCA normals on a built 3-D float tree. Replaces the per-point loop of
 * estimate_normals<T> (pptk/processing/estimate_normals/estimate_normals.cpp:
 * 64-111): neighbourhood = k nearest including the point itself with d < r*r
 * (or, for k == -1, all points with d < r*r), C = X^T X / n of the centred
 * neighbourhood, eigen-decomposition with eigenvalues ascending.
 * rows: NULL -> one row per input point (original order, num_input rows);
 *   else m original indices (the subsample, already range-checked).
 * num_eigen 1: out_evecs m x 3 (eigenvector of the smallest eigenvalue),
 *   out_evals m; num_eigen 3: out_evecs m x 3 x 3 with [i,j,:] the j-th
 *   eigenvector, out_evals m x 3. Output element type = the tree's dtype
 *   (float32 or float64). Eigenvector sign: the component of largest
 *   magnitude is positive. Rows whose point is NaN/inf or has an empty
 *   neighbourhood get NaN vectors/values and size 0.
 * out_evals / out_sizes may be NULL. out_num_fp64: NULL or receives the
 * number of rows that took the FP64 eigensolve. */
pptk_status pptk_normals(const pptk_tree* tree, int k, double r, const int32_t* rows,
                         int64_t m, int num_eigen, pptk_precision precision,
                         void* out_evecs, void* out_evals, int32_t* out_sizes,
                         int64_t* out_num_fp64, pptk_mem mem, void* stream);

/* Self-query kNN and PCA normals of one tree from ONE search per point. No
 * single counterpart in pptk, where a caller that wants both makes two calls,
 * kdtree._query(tree, k=k, dmax=dmax) (kdtree_wrapper.cpp:281-304, the
 * queries=None self query) and estimate_normals(P, nk, r)
 * (estimate_normals.cpp:64-111, which builds and searches a tree of its own):
 * this call writes exactly what
 *   pptk_knn(tree, {PPTK_Q_ALL}, k, dmax, out_idx, out_cnt, out_dist, mem, stream)
 *   pptk_normals(tree, nk, r, NULL, 0, num_eigen, precision, out_evecs,
 *                out_evals, out_sizes, out_num_fp64, mem, stream)
 * write, with the arguments, output shapes and errors of those two calls
 * (k >= 1 and nk >= 1 here). Float32 trees with max(k, nk) <= 32 search
 * once; other trees run the two calls' searches. */
pptk_status pptk_knn_normals(const pptk_tree* tree, int k, double dmax, int nk, double r,
                             int num_eigen, pptk_precision precision, int32_t* out_idx,
                             int32_t* out_cnt, void* out_dist, void* out_evecs,
                             void* out_evals, int32_t* out_sizes, int64_t* out_num_fp64,
                             pptk_mem mem, void* stream);

/* ------------------------------------------------------------------------
 * Batched clouds. No counterpart in pptk, where each cloud is one
 * kdtree._build plus kdtree._query(None, k) plus estimate_normals call
 * (kdtree_wrapper.cpp:281-304, kdtree-impl.h:957-976,
 * estimate_normals.cpp:64-111). A batch plan fixes the cloud sizes and
 * allocates every device buffer once; each run builds one tree per cloud
 * (a forest over the concatenated points), runs the self-query kNN of every
 * point within its cloud and the PCA normals of every point, on one stream,
 * with no host synchronisation inside. With use_graph the whole run is
 * captured once into a hip graph and replayed by later runs with the same
 * buffers. Results per cloud are those of the single-cloud calls on that
 * cloud alone: kNN rows identical, normals from the same neighbourhood.
 * ------------------------------------------------------------------------ */
typedef struct pptk_batch pptk_batch;

typedef struct {
  int k;               /* self-query kNN k, 0..32 (0 = no kNN output) */
  int normals_k;       /* normals neighbourhood size incl. the point, 0..32
                          (0 = no normals); dim must be 3 */
  double dmax;         /* distance bound of both (INFINITY = none) */
  int max_leaf_size;   /* bucket size 1..256, 0 = tuned default */
  int precision;       /* pptk_precision for the normals */
  int use_graph;       /* 1 = capture the run into a hip graph and replay it */
  int query_block;     /* threads per block of the query kernel: 64, 128 or
                          256 (0 = 256) */
  int timing;          /* 1 = run eagerly (no graph) and time the build and
                          query phases with hip events (pptk_batch_info) */
} pptk_batch_params;

typedef struct {
  int64_t num_clouds;
  int64_t num_points;    /* total input rows */
  int64_t num_valid;     /* rows without NaN/inf (after a run) */
  int64_t num_nodes;     /* forest nodes (after a run) */
  int depth;             /* deepest node level + 1 (after a run) */
  int leaf_size;
  int morton_bits;       /* quantisation bits per axis inside a cloud */
  int graph_nodes;       /* kernel launches per run */
  int64_t device_bytes;  /* device memory held by the plan */
  int64_t num_fp64;      /* rows that took the FP64 eigensolve (last run) */
  double build_ms;       /* last timed run: forest build (all kernels before
                            the query kernel) */
  double query_ms;       /* last timed run: fused kNN + normals kernel */
} pptk_batch_info;

/* Plan for clouds of the given sizes (host array of num_clouds counts,
 * total at most 2^30). dim in {2,3,4}; dtype PPTK_F32 or PPTK_F64. */
pptk_status pptk_batch_create(const int64_t* cloud_sizes, int64_t num_clouds, int dim,
                              pptk_dtype dtype, const pptk_batch_params* params,
                              pptk_batch** out);

/* One run. All pointers are device pointers:
 *   pts      : num_points x dim, clouds concatenated in order
 *   out_idx  : num_points x k int32, neighbour indices LOCAL to the cloud,
 *              ascending by (d, index), -1 padded; NULL if k == 0
 *   out_cnt  : NULL or num_points int32 neighbour counts
 *   out_nrm  : num_points x 3, eigenvector of the smallest eigenvalue (dtype
 *              of the points), NaN for invalid rows; NULL if normals_k == 0
 *   out_eval : NULL or num_points smallest eigenvalues
 * Returns after the run completes on `stream` (NULL = library stream). */
pptk_status pptk_batch_run(pptk_batch* b, const void* pts, int32_t* out_idx, int32_t* out_cnt,
                           void* out_nrm, void* out_eval, void* stream);

pptk_status pptk_batch_get_info(const pptk_batch* b, pptk_batch_info* out);

/* Free a plan (and its graph). NULL is a no-op. */
void pptk_batch_free(pptk_batch* b);

/* Building blocks exported for testing and for callers that need them. */

/* Stable ascending sort of (key, value) pairs by the key bits
 * [0, end_bit), in place. Buffers in `mem`. */
pptk_status pptk_sort_pairs_u64(uint64_t* keys, uint32_t* vals, int64_t n, int end_bit,
                                pptk_mem mem);

/* Exclusive prefix sum of n int32 counts into n+1 int64 offsets. */
pptk_status pptk_exclusive_scan_i32(const int32_t* in, int64_t* out, int64_t n,
                                    pptk_mem mem);

#ifdef __cplusplus
}
#endif

#endif /* PPTK_AMD_H_ */
