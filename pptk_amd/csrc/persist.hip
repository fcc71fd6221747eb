// Persistent tree format: a device-resident tree as one flat byte buffer, to
// keep trees across process restarts (a server's resident trees) or send
// them over a wire without rebuilding.
//
// pptk keeps the same idea as an unused flat node array: SmallNode
// (pptk/kdtree/src/small_node.h:13-45) filled by KdTree::SmallNodes()
// (pptk/kdtree/src/kdtree-impl.h:605-627, 863-871), which is never written
// out. This format stores this library's tree instead (Morton-cell nodes with
// tight boxes, see build.hip), little-endian, every section 8-byte aligned:
//
//   FileHeader  (magic "PPTKTREE", version, shape, counts, box, Morton frame,
//                section offsets)
//   coords  n_valid x dim   distance type, tree order
//   orig    n_valid         int32, tree position -> original row
//   fwd     n_input         int32, original row -> tree position or -1
//   nodes   nn x 4          int32 [a, b, child, parent]
//   boxes   nn x 2 x dim    distance type [lo(dim), hi(dim)]
//   leafof  n_valid         int32, tree position -> leaf node
//
// Loading validates every index on the host before anything reaches the
// device (node ranges, child pairs, parents, leaf cover, the two row maps,
// depth and leaf count) and every node box (finite, holding its points and
// its children's boxes), so a damaged or hostile buffer is rejected with an
// error instead of steering a kernel out of bounds or pruning wrongly.
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "tree.hpp"

namespace pk {

namespace {

constexpr char kFileMagic[8] = {'P', 'P', 'T', 'K', 'T', 'R', 'E', 'E'};
constexpr uint32_t kFileVersion = 1;

struct FileHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t dim, dtype, dist, leaf, depth, pad;
  int64_t n_input, n_valid, nn, nleaves, nroots;
  double lo[4], hi[4], mlo[4], mscale[4];
  int64_t off_coords, off_orig, off_fwd, off_nodes, off_boxes, off_leafof, total;
};

int64_t align8(int64_t x) { return (x + 7) & ~int64_t(7); }

void layout(FileHeader& h) {
  const int64_t ds = h.dist == PPTK_F64 ? 8 : 4;
  int64_t o = align8(sizeof(FileHeader));
  h.off_coords = o; o = align8(o + h.n_valid * h.dim * ds);
  h.off_orig = o;   o = align8(o + h.n_valid * 4);
  h.off_fwd = o;    o = align8(o + h.n_input * 4);
  h.off_nodes = o;  o = align8(o + h.nn * 16);
  h.off_boxes = o;  o = align8(o + h.nn * 2 * h.dim * ds);
  h.off_leafof = o; o = align8(o + h.n_valid * 4);
  h.total = o;
}

template <typename D>
bool all_finite(const D* p, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!(p[i] - p[i] == D(0))) return false;
  return true;
}

// Host-side validation of a deserialised tree; fails with PPTK_E_VALUE.
void validate(const FileHeader& h, const uint8_t* buf) {
  const int32_t* orig = (const int32_t*)(buf + h.off_orig);
  const int32_t* fwd = (const int32_t*)(buf + h.off_fwd);
  const NodeRec* nodes = (const NodeRec*)(buf + h.off_nodes);
  const int32_t* leafof = (const int32_t*)(buf + h.off_leafof);
  const int64_t nv = h.n_valid, ni = h.n_input, nn = h.nn;
  auto bad = [](const char* what) { fail(PPTK_E_VALUE, "invalid tree buffer: %s", what); };
  std::vector<int8_t> depth((size_t)nn, -1);
  int max_depth = 0;
  for (int64_t i = 0; i < nn; ++i) {
    const NodeRec& r = nodes[i];
    if (r.a < 0 || r.a > r.b || r.b > nv) bad("node range");
    if (i < h.nroots) {
      if (r.parent != -1) bad("root with a parent");
      depth[i] = 0;
    } else {
      if (r.parent < 0 || r.parent >= i) bad("node parent");
      const NodeRec& p = nodes[r.parent];
      if (!(p.child == i || p.child + 1 == i)) bad("parent does not list the node");
      if (depth[r.parent] < 0) bad("unreachable node");
      const int d = depth[r.parent] + 1;
      if (d > 120) bad("tree deeper than 120 levels");
      depth[i] = (int8_t)d;
      max_depth = d > max_depth ? d : max_depth;
    }
    if (r.child != -1) {
      const int64_t c = r.child;
      if (c <= i || (c & 1) == 0 || c + 1 >= nn) bad("child pair");
      const NodeRec& l = nodes[c];
      const NodeRec& rr = nodes[c + 1];
      if (l.parent != i || rr.parent != i) bad("child parents");
      if (l.a != r.a || l.b != rr.a || rr.b != r.b) bad("child ranges");
    }
  }
  if (h.nroots != 1 || nodes[0].a != 0 || nodes[0].b != nv) bad("root range");
  for (int64_t p = 0; p < nv; ++p) {
    const int32_t lf = leafof[p];
    if (lf < 0 || lf >= nn) bad("leafof index");
    const NodeRec& r = nodes[lf];
    if (r.child != -1 || p < r.a || p >= r.b) bad("leafof does not hold the position");
    const int32_t o = orig[p];
    if (o < 0 || o >= ni || fwd[o] != p) bad("row maps disagree");
  }
  int64_t mapped = 0;
  for (int64_t i = 0; i < ni; ++i) {
    const int32_t p = fwd[i];
    if (p == -1) continue;
    if (p < 0 || p >= nv || orig[p] != i) bad("forward map");
    ++mapped;
  }
  if (mapped != nv) bad("forward map count");
  int64_t leaves = 0;
  for (int64_t i = 0; i < nn; ++i) leaves += nodes[i].child == -1;
  if (h.depth != max_depth + 1) bad("depth field");
  if (h.nleaves != leaves) bad("leaf count field");
}

// Node boxes steer pruning, so a box that does not hold its node's points
// would silently drop neighbours. Every non-empty node's box must be finite
// with lo <= hi, a leaf's box must contain its points and an inner node's
// box both child boxes. An empty node must carry the empty box
// (lo = +inf, hi = -inf), which prunes it for every query.
template <typename D>
void validate_boxes(const FileHeader& h, const uint8_t* buf) {
  const D* coords = (const D*)(buf + h.off_coords);
  const D* boxes = (const D*)(buf + h.off_boxes);
  const NodeRec* nodes = (const NodeRec*)(buf + h.off_nodes);
  const int dim = h.dim;
  const D inf = std::numeric_limits<D>::infinity();
  auto bad = [](const char* what) { fail(PPTK_E_VALUE, "invalid tree buffer: %s", what); };
  auto lo = [&](int64_t i, int d) { return boxes[i * 2 * dim + d]; };
  auto hi = [&](int64_t i, int d) { return boxes[i * 2 * dim + dim + d]; };
  for (int64_t i = 0; i < h.nn; ++i) {
    const NodeRec& r = nodes[i];
    if (r.a == r.b) {
      for (int d = 0; d < dim; ++d)
        if (!(lo(i, d) == inf && hi(i, d) == -inf)) bad("box of an empty node");
      continue;
    }
    for (int d = 0; d < dim; ++d) {
      const D l = lo(i, d), u = hi(i, d);
      if (!(l - l == D(0)) || !(u - u == D(0)) || !(l <= u)) bad("node box");
    }
    if (r.child == -1) {
      for (int32_t p = r.a; p < r.b; ++p)
        for (int d = 0; d < dim; ++d) {
          const D c = coords[(int64_t)p * dim + d];
          if (c < lo(i, d) || c > hi(i, d)) bad("leaf box does not contain its points");
        }
    } else {
      for (int64_t c = r.child; c <= (int64_t)r.child + 1; ++c) {
        if (nodes[c].a == nodes[c].b) continue;
        for (int d = 0; d < dim; ++d)
          if (lo(c, d) < lo(i, d) || hi(c, d) > hi(i, d))
            bad("node box does not contain a child box");
      }
    }
  }
}

template <typename D, int DIM>
void upload(pptk_tree* t, const FileHeader& h, const uint8_t* buf, hipStream_t s) {
  const int64_t nv = h.n_valid, nn = h.nn;
  const D* coords = (const D*)(buf + h.off_coords);
  const D* boxes = (const D*)(buf + h.off_boxes);
  std::vector<PointRec<D, DIM>> rec((size_t)std::max<int64_t>(nv, 1));
  const int32_t* orig = (const int32_t*)(buf + h.off_orig);
  for (int64_t p = 0; p < nv; ++p) {
    PointRec<D, DIM> r;
    std::memset(&r, 0, sizeof(r));
    for (int d = 0; d < DIM; ++d) r.c[d] = coords[p * DIM + d];
    r.orig = orig[p];
    rec[(size_t)p] = r;
  }
  std::vector<BoxRec<D, DIM>> box((size_t)nn);
  for (int64_t i = 0; i < nn; ++i) {
    BoxRec<D, DIM> b;
    std::memset(&b, 0, sizeof(b));
    for (int d = 0; d < DIM; ++d) {
      b.lo[d] = boxes[i * 2 * DIM + d];
      b.hi[d] = boxes[i * 2 * DIM + DIM + d];
    }
    box[(size_t)i] = b;
  }
  t->pts.reset((size_t)std::max<int64_t>(nv, 1) * sizeof(PointRec<D, DIM>));
  t->orig.reset((size_t)std::max<int64_t>(nv, 1) * 4);
  t->fwd.reset((size_t)std::max<int64_t>(h.n_input, 1) * 4);
  t->leafof.reset((size_t)std::max<int64_t>(nv, 1) * 4);
  t->node.reset((size_t)nn * sizeof(NodeRec));
  t->box.reset((size_t)nn * sizeof(BoxRec<D, DIM>));
  if (nv > 0) {
    PK_HIP(hipMemcpyAsync(t->pts.get(), rec.data(), (size_t)nv * sizeof(PointRec<D, DIM>),
                          hipMemcpyHostToDevice, s));
    PK_HIP(hipMemcpyAsync(t->orig.get(), orig, (size_t)nv * 4, hipMemcpyHostToDevice, s));
    PK_HIP(hipMemcpyAsync(t->leafof.get(), buf + h.off_leafof, (size_t)nv * 4,
                          hipMemcpyHostToDevice, s));
  }
  if (h.n_input > 0)
    PK_HIP(hipMemcpyAsync(t->fwd.get(), buf + h.off_fwd, (size_t)h.n_input * 4,
                          hipMemcpyHostToDevice, s));
  PK_HIP(hipMemcpyAsync(t->node.get(), buf + h.off_nodes, (size_t)nn * sizeof(NodeRec),
                        hipMemcpyHostToDevice, s));
  PK_HIP(hipMemcpyAsync(t->box.get(), box.data(), (size_t)nn * sizeof(BoxRec<D, DIM>),
                        hipMemcpyHostToDevice, s));
  PK_HIP(hipStreamSynchronize(s));  // the host staging vectors go out of scope
}

}  // namespace

pptk_status api_tree_serialize(const pptk_tree* tree, void* buf, int64_t capacity,
                               int64_t* out_size) {
  return guard([&] {
    check_tree(tree);
    if (!out_size) fail(PPTK_E_VALUE, "null size output");
    if (tree->nroots != 1) fail(PPTK_E_NOTIMPL, "only single-cloud trees are serialised");
    FileHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kFileMagic, 8);
    h.version = kFileVersion;
    h.header_bytes = sizeof(FileHeader);
    h.dim = tree->dim;
    h.dtype = tree->dtype;
    h.dist = tree->dist;
    h.leaf = tree->leaf;
    h.depth = tree->depth;
    h.n_input = tree->n_input;
    h.n_valid = tree->n_valid;
    h.nn = tree->nn;
    h.nleaves = tree->nleaves;
    h.nroots = tree->nroots;
    for (int d = 0; d < 4; ++d) {
      h.lo[d] = tree->lo[d];
      h.hi[d] = tree->hi[d];
      h.mlo[d] = tree->mlo[d];
      h.mscale[d] = tree->mscale[d];
    }
    layout(h);
    *out_size = h.total;
    if (!buf) return;
    if (capacity < h.total) fail(PPTK_E_VALUE, "buffer of %lld bytes, %lld needed",
                                 (long long)capacity, (long long)h.total);
    uint8_t* b = (uint8_t*)buf;
    std::memset(b, 0, (size_t)h.total);
    std::memcpy(b, &h, sizeof(h));
    const pptk_status st = pptk_tree_export(tree, b + h.off_coords, (int32_t*)(b + h.off_orig),
                                            (int32_t*)(b + h.off_fwd), b + h.off_boxes,
                                            (int32_t*)(b + h.off_nodes), PPTK_MEM_HOST);
    if (st != PPTK_OK) fail(st, "%s", pptk_last_error());
    hipStream_t s = lib_stream();
    if (tree->n_valid > 0) {
      PK_HIP(hipMemcpyAsync(b + h.off_leafof, tree->leafof.get(), (size_t)tree->n_valid * 4,
                            hipMemcpyDeviceToHost, s));
      PK_HIP(hipStreamSynchronize(s));
    }
  });
}

pptk_status api_tree_deserialize(const void* buf, int64_t size, void* stream, pptk_tree** out) {
  return guard([&] {
    if (!out) fail(PPTK_E_VALUE, "null output handle");
    *out = nullptr;
    if (!buf || size < (int64_t)sizeof(FileHeader))
      fail(PPTK_E_VALUE, "invalid tree buffer: too short");
    FileHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kFileMagic, 8) != 0) fail(PPTK_E_VALUE, "invalid tree buffer: magic");
    if (h.version != kFileVersion || h.header_bytes != sizeof(FileHeader))
      fail(PPTK_E_VALUE, "unsupported tree buffer version %u", h.version);
    if (h.dim < 2 || h.dim > 4 || !dtype_valid(h.dtype) ||
        h.dist != (h.dtype == PPTK_F64 ? PPTK_F64 : PPTK_F32) || h.leaf < 1 || h.leaf > 256)
      fail(PPTK_E_VALUE, "invalid tree buffer: shape fields");
    if (h.n_input < 0 || h.n_input > (1LL << 30) || h.n_valid < 0 || h.n_valid > h.n_input ||
        h.nn < 1 || h.nn > 2 * std::max<int64_t>(h.n_valid, 1) || h.nroots < 1 || h.nroots > h.nn)
      fail(PPTK_E_VALUE, "invalid tree buffer: counts");
    FileHeader want = h;
    layout(want);
    if (want.off_coords != h.off_coords || want.off_orig != h.off_orig ||
        want.off_fwd != h.off_fwd || want.off_nodes != h.off_nodes ||
        want.off_boxes != h.off_boxes || want.off_leafof != h.off_leafof ||
        want.total != h.total || h.total > size)
      fail(PPTK_E_VALUE, "invalid tree buffer: section layout");
    const uint8_t* b = (const uint8_t*)buf;
    validate(h, b);
    const int64_t nc = h.n_valid * h.dim;
    const bool finite = h.dist == PPTK_F64 ? all_finite((const double*)(b + h.off_coords), nc)
                                           : all_finite((const float*)(b + h.off_coords), nc);
    if (!finite) fail(PPTK_E_VALUE, "invalid tree buffer: coordinates");
    if (h.dist == PPTK_F64) validate_boxes<double>(h, b);
    else validate_boxes<float>(h, b);
    hipStream_t s = pick_stream(stream);
    std::unique_ptr<pptk_tree> t(new pptk_tree());
    t->dim = h.dim;
    t->dtype = (pptk_dtype)h.dtype;
    t->dist = (pptk_dtype)h.dist;
    t->n_input = h.n_input;
    t->n_valid = h.n_valid;
    t->leaf = h.leaf;
    t->depth = h.depth;
    t->nleaves = h.nleaves;
    t->nn = h.nn;
    t->nroots = h.nroots;
    for (int d = 0; d < 4; ++d) {
      t->lo[d] = h.lo[d];
      t->hi[d] = h.hi[d];
      t->mlo[d] = h.mlo[d];
      t->mscale[d] = h.mscale[d];
    }
    dispatch_tree(t.get(), [&](auto dt, auto dc) {
      using D = typename decltype(dt)::type;
      constexpr int DIM = decltype(dc)::value;
      upload<D, DIM>(t.get(), h, b, s);
    });
    register_tree(t.get());
    *out = t.release();
  });
}

}  // namespace pk

extern "C" pptk_status pptk_tree_serialize(const pptk_tree* tree, void* buf, int64_t capacity,
                                           int64_t* out_size) {
  return pk::api_tree_serialize(tree, buf, capacity, out_size);
}

extern "C" pptk_status pptk_tree_deserialize(const void* buf, int64_t size, void* stream,
                                             pptk_tree** out) {
  return pk::api_tree_deserialize(buf, size, stream, out);
}
