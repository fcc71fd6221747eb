// Persistent tree format: layout and host-side validation, plain C++ with no
// device code (the library's loader uses it, and it also builds on its own
// for sanitizer checks of the parser).
//
// pptk keeps the same idea as an unused flat node array: SmallNode
// (pptk/kdtree/src/small_node.h:13-45) filled by KdTree::SmallNodes()
// (pptk/kdtree/src/kdtree-impl.h:605-627, 863-871), which is never written
// out. This format stores this library's tree instead (Morton-cell nodes with
// tight boxes), little-endian, every section 8-byte aligned:
//
//   Header      (magic "PPTKTREE", version, shape, counts, box, Morton frame,
//                section offsets)
//   coords  n_valid x dim   distance type, tree order
//   orig    n_valid         int32, tree position -> original row
//   fwd     n_input         int32, original row -> tree position or -1
//   nodes   nn x 4          int32 [a, b, child, parent]
//   boxes   nn x 2 x dim    distance type [lo(dim), hi(dim)]
//   leafof  n_valid         int32, tree position -> leaf node
//
// check() validates every field and index before anything is used (section
// layout against the buffer size, node ranges, child pairs, parents, leaf
// cover, the two row maps, depth and leaf count, finite coordinates) and
// every node box (finite, holding its points and its children's boxes), so a
// damaged or hostile buffer is rejected instead of steering a kernel out of
// bounds or pruning wrongly.
#pragma once

#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/pptk_amd.h"

namespace pk {
namespace fmt {

// Reports an invalid buffer and does not return. The library raises
// PPTK_E_VALUE with this message; a standalone checker may throw anything.
[[noreturn]] void invalid(const std::string& msg);

constexpr char kMagic[8] = {'P', 'P', 'T', 'K', 'T', 'R', 'E', 'E'};
constexpr uint32_t kVersion = 1;

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int32_t dim, dtype, dist, leaf, depth, pad;
  int64_t n_input, n_valid, nn, nleaves, nroots;
  double lo[4], hi[4], mlo[4], mscale[4];
  int64_t off_coords, off_orig, off_fwd, off_nodes, off_boxes, off_leafof, total;
};

// A node as stored: points [a, b) in tree order, first child of the adjacent
// pair (child, child + 1) or -1, parent or -1.
struct Node {
  int32_t a, b, child, parent;
};

inline int64_t align8(int64_t x) { return (x + 7) & ~int64_t(7); }

// Section offsets and total size for the header's shape and counts.
inline void layout(Header& h) {
  const int64_t ds = h.dist == PPTK_F64 ? 8 : 4;
  int64_t o = align8(sizeof(Header));
  h.off_coords = o; o = align8(o + h.n_valid * h.dim * ds);
  h.off_orig = o;   o = align8(o + h.n_valid * 4);
  h.off_fwd = o;    o = align8(o + h.n_input * 4);
  h.off_nodes = o;  o = align8(o + h.nn * 16);
  h.off_boxes = o;  o = align8(o + h.nn * 2 * h.dim * ds);
  h.off_leafof = o; o = align8(o + h.n_valid * 4);
  h.total = o;
}

inline void bad(const char* what) { invalid(std::string("invalid tree buffer: ") + what); }

template <typename D>
inline bool all_finite(const D* p, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!(p[i] - p[i] == D(0))) return false;
  return true;
}

// A typed view of a section. Every offset is a multiple of 8 and check()
// requires an 8-byte aligned buffer, so every section is aligned for its
// element type.
template <typename T>
inline const T* section(const uint8_t* buf, int64_t off) {
  return reinterpret_cast<const T*>(buf + off);
}

// Node structure, row maps, depth and leaf count.
inline void validate_nodes(const Header& h, const uint8_t* buf) {
  const int32_t* orig = section<int32_t>(buf, h.off_orig);
  const int32_t* fwd = section<int32_t>(buf, h.off_fwd);
  const Node* nodes = section<Node>(buf, h.off_nodes);
  const int32_t* leafof = section<int32_t>(buf, h.off_leafof);
  const int64_t nv = h.n_valid, ni = h.n_input, nn = h.nn;
  std::vector<int8_t> depth((size_t)nn, -1);
  int max_depth = 0;
  for (int64_t i = 0; i < nn; ++i) {
    const Node& r = nodes[i];
    if (r.a < 0 || r.a > r.b || r.b > nv) bad("node range");
    if (i < h.nroots) {
      if (r.parent != -1) bad("root with a parent");
      depth[i] = 0;
    } else {
      if (r.parent < 0 || r.parent >= i) bad("node parent");
      const Node& p = nodes[r.parent];
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
      const Node& l = nodes[c];
      const Node& rr = nodes[c + 1];
      if (l.parent != i || rr.parent != i) bad("child parents");
      if (l.a != r.a || l.b != rr.a || rr.b != r.b) bad("child ranges");
    }
  }
  if (h.nroots != 1 || nodes[0].a != 0 || nodes[0].b != nv) bad("root range");
  for (int64_t p = 0; p < nv; ++p) {
    const int32_t lf = leafof[p];
    if (lf < 0 || lf >= nn) bad("leafof index");
    const Node& r = nodes[lf];
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
inline void validate_boxes(const Header& h, const uint8_t* buf) {
  const D* coords = section<D>(buf, h.off_coords);
  const D* boxes = section<D>(buf, h.off_boxes);
  const Node* nodes = section<Node>(buf, h.off_nodes);
  const int dim = h.dim;
  const D inf = std::numeric_limits<D>::infinity();
  auto lo = [&](int64_t i, int d) { return boxes[i * 2 * dim + d]; };
  auto hi = [&](int64_t i, int d) { return boxes[i * 2 * dim + dim + d]; };
  for (int64_t i = 0; i < h.nn; ++i) {
    const Node& r = nodes[i];
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

// Validates a whole buffer of `size` bytes; returns its header.
inline Header check(const void* buf, int64_t size) {
  if (!buf || size < (int64_t)sizeof(Header)) bad("too short");
  if (reinterpret_cast<uintptr_t>(buf) % 8 != 0) bad("not 8-byte aligned");
  Header h;
  std::memcpy(&h, buf, sizeof(h));
  if (std::memcmp(h.magic, kMagic, 8) != 0) bad("magic");
  if (h.version != kVersion || h.header_bytes != sizeof(Header))
    invalid("unsupported tree buffer version " + std::to_string(h.version));
  const bool dtype_ok = h.dtype >= PPTK_F32 && h.dtype <= PPTK_U64;
  if (h.dim < 2 || h.dim > 4 || !dtype_ok ||
      h.dist != (h.dtype == PPTK_F64 ? PPTK_F64 : PPTK_F32) || h.leaf < 1 || h.leaf > 256)
    bad("shape fields");
  if (h.n_input < 0 || h.n_input > (int64_t(1) << 30) || h.n_valid < 0 ||
      h.n_valid > h.n_input || h.nn < 1 || h.nn > 2 * (h.n_valid > 1 ? h.n_valid : 1) ||
      h.nroots < 1 || h.nroots > h.nn)
    bad("counts");
  Header want = h;
  layout(want);
  if (want.off_coords != h.off_coords || want.off_orig != h.off_orig ||
      want.off_fwd != h.off_fwd || want.off_nodes != h.off_nodes ||
      want.off_boxes != h.off_boxes || want.off_leafof != h.off_leafof ||
      want.total != h.total || h.total > size)
    bad("section layout");
  const uint8_t* b = (const uint8_t*)buf;
  validate_nodes(h, b);
  const int64_t nc = h.n_valid * h.dim;
  const bool finite = h.dist == PPTK_F64 ? all_finite(section<double>(b, h.off_coords), nc)
                                         : all_finite(section<float>(b, h.off_coords), nc);
  if (!finite) bad("coordinates");
  if (h.dist == PPTK_F64) validate_boxes<double>(h, b);
  else validate_boxes<float>(h, b);
  return h;
}

}  // namespace fmt
}  // namespace pk
