#pragma once
#include "common.hpp"

#include <mutex>

namespace pk {
// The processing order of the last ordered query set served by a tree, kept
// for the next call on the same query array with the same count (an r-near
// count pass and then its fill pass). Any permutation of the queries gives
// the same results, so a stale entry (same array, new contents) only costs
// locality. An entry is published after the stream that made it has
// synchronised.
struct OrderCache {
  std::mutex mu;
  const void* src = nullptr;
  int64_t m = -1;
  int mode = -1;
  std::shared_ptr<DevBuf> order;
};

// The event lists of the last r-near count on a single-cloud tree (the nodes
// at which each wave of queries took a whole range or scanned a leaf), kept
// for the fill of the same queries at the same radius so that the fill does
// not walk the tree again. The entry keeps a copy of the query data and the
// processing order the count used; the fill replays it only if its own
// queries equal the copy (compared on the device), and walks the tree
// otherwise. An entry is published after the stream that made it has
// synchronised and is never written afterwards; the fill that replays it
// removes it from the cache.
struct RnearEvents {
  const void* src = nullptr;   // the caller's query array
  int64_t m = -1;
  int mode = -1;
  double r = -1.0;             // radius in the tree's distance type
  int pk = 0;                  // queries per count wave
  int64_t nw = 0;              // count waves
  int64_t nchunks = 0;         // chunks of 64 events in the pool
  int64_t chunks_used = 0;     // chunks the count took (> nchunks: some waves not kept)
  std::shared_ptr<DevBuf> order;  // processing order of the count (null: input order)
  DevBuf queries;              // the query data the count read (points or indices)
  size_t qbytes = 0;
  DevBuf node, mask, link, hdr, head;
};
struct RnearCache {
  std::mutex mu;
  std::shared_ptr<RnearEvents> last;
  double chunks_per_wave = 0;  // pool size the last count on this tree needed
};
}  // namespace pk

// Host-side tree handle (opaque `pptk_tree` of the C ABI).
struct pptk_tree {
  static constexpr uint64_t kMagic = 0x70707462676b6474ull;
  uint64_t magic = kMagic;
  int dim = 3;
  pptk_dtype dtype = PPTK_F32;
  pptk_dtype dist = PPTK_F32;
  int64_t n_input = 0;
  int64_t n_valid = 0;
  int leaf = 16;
  int depth = 0;          // number of node levels
  int64_t nleaves = 0;
  int64_t nn = 1;         // number of nodes
  int64_t nroots = 1;     // one root per cloud (forest builds)
  double lo[4] = {0, 0, 0, 0};
  double hi[4] = {0, 0, 0, 0};
  double mlo[4] = {0, 0, 0, 0};     // Morton frame (exact copies of the
  double mscale[4] = {0, 0, 0, 0};  // distance-type values used in the build)
  pk::DevBuf pts;   // PointRec<D,dim>[n_valid]
  pk::DevBuf orig;  // int32[n_valid]
  pk::DevBuf fwd;   // int32[n_input]
  pk::DevBuf box;   // BoxRec<D,dim>[nn]
  pk::DevBuf node;  // NodeRec[nn]
  pk::DevBuf leafof;  // int32[n_valid]: leaf node holding each tree position
  mutable pk::OrderCache order_cache;
  mutable pk::RnearCache rnear_cache;
};

namespace pk {

const pptk_tree* check_tree(const pptk_tree* t);

// Adds a tree made outside pptk_tree_build (a deserialised one) to the set
// of live handles.
void register_tree(pptk_tree* t);

template <typename D, int DIM>
TreeView<D, DIM> tree_view(const pptk_tree* t) {
  TreeView<D, DIM> v;
  v.pts = t->pts.as<const PointRec<D, DIM>>();
  v.orig = t->orig.as<const int32_t>();
  v.fwd = t->fwd.as<const int32_t>();
  v.box = t->box.as<const BoxRec<D, DIM>>();
  v.node = t->node.as<const NodeRec>();
  v.leafof = t->leafof.as<const int32_t>();
  v.n = (int32_t)t->n_valid;
  v.leaf = t->leaf;
  v.nn = (int32_t)t->nn;
  v.n_input = t->n_input;
  v.n_dev = nullptr;
  v.nroots = (int32_t)t->nroots;
  return v;
}

// Calls f(Tag<D>{}, integral_constant<int, DIM>{}) for the tree's distance
// type D and dimension.
template <typename F>
void dispatch_tree(const pptk_tree* t, F&& f) {
  const bool dbl = t->dist == PPTK_F64;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  switch (t->dim) {
    case 2: if (dbl) f(Tag<double>{}, I2{}); else f(Tag<float>{}, I2{}); break;
    case 3: if (dbl) f(Tag<double>{}, I3{}); else f(Tag<float>{}, I3{}); break;
    case 4: if (dbl) f(Tag<double>{}, I4{}); else f(Tag<float>{}, I4{}); break;
    default: fail(PPTK_E_RUNTIME, "invalid k-d tree dim = %d", t->dim);
  }
}

// Tuned default bucket size.
int default_leaf_size(int dim, bool dbl, int64_t n);

}  // namespace pk
