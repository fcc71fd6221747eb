// Drives the persisted-tree parser (pptk_amd/csrc/persist_format.hpp, the
// host-side validation behind pptk_tree_deserialize) under AddressSanitizer
// and UndefinedBehaviorSanitizer:
//   * intact buffers of small trees (float and double, 2-4 dimensions, with
//     invalid input rows) pass;
//   * targeted damage to every section (ranges, child pairs, parents, row
//     maps, leaf map, depth and leaf-count fields, boxes, coordinates,
//     counts, offsets, magic, version) is rejected;
//   * every truncation, a misaligned copy and 30,000 seeded random byte and
//     word corruptions either pass or are rejected, with no read outside the
//     buffer and no undefined behaviour (the sanitizers abort otherwise).
// Test infrastructure: built and run by tests/test_sanitizers_cpu.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <vector>

#include "persist_format.hpp"

namespace pk {
namespace fmt {
void invalid(const std::string& msg) { throw std::runtime_error(msg); }
}  // namespace fmt
}  // namespace pk

using pk::fmt::Header;
using pk::fmt::Node;

static int g_failures = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);  \
      std::printf(__VA_ARGS__);                         \
      std::printf("\n");                                \
      if (++g_failures > 20) std::exit(1);              \
    }                                                   \
  } while (0)

// A buffer with its section views (8-byte words keep it aligned).
struct Buf {
  std::vector<uint64_t> words;
  int64_t size = 0;
  uint8_t* data() { return (uint8_t*)words.data(); }
  Header& h() { return *(Header*)words.data(); }
  template <typename T> T* at(int64_t off) { return (T*)(data() + off); }
};

// A valid tree over n points (`bad_rows` of them NaN, left out): points in
// tree order sorted by the first coordinate, ranges halved until <= leaf,
// children appended in pairs from node 1, tight boxes.
template <typename D>
static Buf make_tree(int n, int dim, int leaf, int bad_rows, std::mt19937& rng) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<D> in((size_t)n * dim);
  for (auto& x : in) x = (D)u(rng);
  std::vector<int> valid;
  for (int i = 0; i < n; ++i) {
    if (i % 7 == 3 && bad_rows > 0) {
      in[(size_t)i * dim] = std::numeric_limits<D>::quiet_NaN();
      --bad_rows;
      continue;
    }
    valid.push_back(i);
  }
  std::sort(valid.begin(), valid.end(),
            [&](int a, int b) { return in[(size_t)a * dim] < in[(size_t)b * dim]; });
  const int64_t nv = (int64_t)valid.size();
  std::vector<Node> nodes(1, Node{0, (int32_t)nv, -1, -1});
  std::vector<int> depth(1, 0);
  for (size_t i = 0; i < nodes.size(); ++i) {
    if (nodes[i].b - nodes[i].a <= leaf) continue;
    const int32_t a = nodes[i].a, b = nodes[i].b, m = a + (b - a) / 2;
    nodes[i].child = (int32_t)nodes.size();
    nodes.push_back(Node{a, m, -1, (int32_t)i});
    nodes.push_back(Node{m, b, -1, (int32_t)i});
    depth.push_back(depth[i] + 1);
    depth.push_back(depth[i] + 1);
  }
  Header h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, pk::fmt::kMagic, 8);
  h.version = pk::fmt::kVersion;
  h.header_bytes = sizeof(Header);
  h.dim = dim;
  h.dtype = sizeof(D) == 8 ? PPTK_F64 : PPTK_F32;
  h.dist = h.dtype;
  h.leaf = leaf;
  h.depth = *std::max_element(depth.begin(), depth.end()) + 1;
  h.n_input = n;
  h.n_valid = nv;
  h.nn = (int64_t)nodes.size();
  h.nleaves = std::count_if(nodes.begin(), nodes.end(), [](const Node& r) { return r.child < 0; });
  h.nroots = 1;
  pk::fmt::layout(h);
  Buf buf;
  buf.words.assign((size_t)(h.total + 7) / 8, 0);
  buf.size = h.total;
  std::memcpy(buf.data(), &h, sizeof(h));
  D* coords = buf.at<D>(h.off_coords);
  int32_t* orig = buf.at<int32_t>(h.off_orig);
  int32_t* fwd = buf.at<int32_t>(h.off_fwd);
  Node* nd = buf.at<Node>(h.off_nodes);
  D* boxes = buf.at<D>(h.off_boxes);
  int32_t* leafof = buf.at<int32_t>(h.off_leafof);
  for (int i = 0; i < n; ++i) fwd[i] = -1;
  for (int64_t p = 0; p < nv; ++p) {
    for (int d = 0; d < dim; ++d) coords[p * dim + d] = in[(size_t)valid[p] * dim + d];
    orig[p] = valid[p];
    fwd[valid[p]] = (int32_t)p;
  }
  const D inf = std::numeric_limits<D>::infinity();
  for (int64_t i = (int64_t)nodes.size() - 1; i >= 0; --i) {
    nd[i] = nodes[i];
    for (int d = 0; d < dim; ++d) {
      D lo = inf, hi = -inf;
      for (int32_t p = nodes[i].a; p < nodes[i].b; ++p) {
        lo = std::min(lo, coords[(int64_t)p * dim + d]);
        hi = std::max(hi, coords[(int64_t)p * dim + d]);
      }
      boxes[i * 2 * dim + d] = lo;
      boxes[i * 2 * dim + dim + d] = hi;
    }
    if (nodes[i].child < 0)
      for (int32_t p = nodes[i].a; p < nodes[i].b; ++p) leafof[p] = (int32_t)i;
  }
  return buf;
}

static bool passes(const uint8_t* data, int64_t size) {
  try {
    pk::fmt::check(data, size);
    return true;
  } catch (const std::runtime_error&) {
    return false;
  }
}

template <typename D>
static void damage_cases(Buf base, int dim) {
  const Header h0 = base.h();
  CHECK(passes(base.data(), base.size), "intact buffer rejected");
  auto rejects = [&](const char* what, const std::function<void(Buf&)>& f) {
    Buf b = base;
    f(b);
    CHECK(!passes(b.data(), b.size), "damage not rejected: %s", what);
  };
  const int64_t nn = h0.nn, nv = h0.n_valid;
  const int64_t inner = [&] {
    for (int64_t i = 0; i < nn; ++i)
      if (base.at<Node>(h0.off_nodes)[i].child >= 0) return i;
    return int64_t(-1);
  }();
  const int64_t leaf = [&] {
    for (int64_t i = nn - 1; i >= 0; --i)
      if (base.at<Node>(h0.off_nodes)[i].child < 0 && base.at<Node>(h0.off_nodes)[i].b >
                                                          base.at<Node>(h0.off_nodes)[i].a)
        return i;
    return int64_t(-1);
  }();
  rejects("magic", [](Buf& b) { b.h().magic[0] = 'X'; });
  rejects("version", [](Buf& b) { b.h().version = 2; });
  rejects("dim", [](Buf& b) { b.h().dim = 7; });
  rejects("leaf", [](Buf& b) { b.h().leaf = 0; });
  rejects("n_valid > n_input", [](Buf& b) { b.h().n_valid = b.h().n_input + 1; });
  rejects("huge n_input", [](Buf& b) { b.h().n_input = int64_t(1) << 40; });
  rejects("nn", [](Buf& b) { b.h().nn = b.h().nn + 2; });
  rejects("offset", [](Buf& b) { b.h().off_fwd += 8; });
  rejects("total", [](Buf& b) { b.h().total += 8; });
  rejects("depth field", [](Buf& b) { b.h().depth += 1; });
  rejects("leaf count", [](Buf& b) { b.h().nleaves += 1; });
  rejects("short", [](Buf& b) { b.size -= 8; });
  rejects("orig", [&](Buf& b) { b.at<int32_t>(h0.off_orig)[0] = (int32_t)h0.n_input; });
  if (nv > 0) {
    rejects("fwd", [&](Buf& b) {
      b.at<int32_t>(h0.off_fwd)[b.at<int32_t>(h0.off_orig)[0]] = -1;  // a valid row unmapped
    });
    rejects("leafof", [&](Buf& b) { b.at<int32_t>(h0.off_leafof)[nv - 1] = (int32_t)nn; });
    rejects("coordinate NaN", [&](Buf& b) {
      b.at<D>(h0.off_coords)[0] = std::numeric_limits<D>::quiet_NaN();
    });
  }
  if (inner >= 0) {
    rejects("child pair", [&](Buf& b) { b.at<Node>(h0.off_nodes)[inner].child += 2; });
    rejects("child range", [&](Buf& b) {
      Node* nd = b.at<Node>(h0.off_nodes);
      nd[nd[inner].child].b -= 1;
    });
    rejects("parent", [&](Buf& b) {
      Node* nd = b.at<Node>(h0.off_nodes);
      nd[nd[inner].child].parent = (int32_t)nn;
    });
    rejects("inner box", [&](Buf& b) {
      D* bx = b.at<D>(h0.off_boxes);
      bx[inner * 2 * dim + dim] = bx[inner * 2 * dim];  // hi = lo
    });
  }
  if (leaf >= 0) {
    rejects("leaf box", [&](Buf& b) {
      D* bx = b.at<D>(h0.off_boxes);
      const Node nd = b.at<Node>(h0.off_nodes)[leaf];
      D lo = bx[leaf * 2 * dim];
      D c = b.at<D>(h0.off_coords)[(int64_t)nd.a * dim];
      bx[leaf * 2 * dim] = c > lo ? c : lo + (D)1e-3;  // lo above a point
      if (bx[leaf * 2 * dim] > bx[leaf * 2 * dim + dim]) bx[leaf * 2 * dim + dim] = bx[leaf * 2 * dim];
      // shrink until some point falls outside
      for (int32_t p = nd.a; p < nd.b; ++p) {
        const D x = b.at<D>(h0.off_coords)[(int64_t)p * dim];
        if (x < bx[leaf * 2 * dim]) return;
      }
      bx[leaf * 2 * dim] = std::nextafter(bx[leaf * 2 * dim], (D)10);
      bx[leaf * 2 * dim + dim] = std::max(bx[leaf * 2 * dim + dim], bx[leaf * 2 * dim]);
    });
    rejects("box inf", [&](Buf& b) {
      b.at<D>(h0.off_boxes)[leaf * 2 * dim + dim] = std::numeric_limits<D>::infinity();
    });
  }
  // misaligned copy: rejected before any typed read
  {
    std::vector<uint8_t> raw((size_t)base.size + 1);
    std::memcpy(raw.data() + 1, base.data(), (size_t)base.size);
    CHECK(!passes(raw.data() + 1, base.size), "misaligned buffer accepted");
  }
  // every truncation
  for (int64_t sz = 0; sz < base.size; sz += 4) {
    std::vector<uint64_t> w((size_t)(sz + 7) / 8 + 1, 0);
    std::memcpy(w.data(), base.data(), (size_t)sz);
    CHECK(!passes((const uint8_t*)w.data(), sz), "truncated buffer (%lld bytes) accepted",
          (long long)sz);
  }
}

// Random corruptions: single bytes, and 32-bit words set to small or extreme
// values (indices near the valid range are the likely dangerous ones).
static void fuzz(const Buf& base, std::mt19937& rng, int iters) {
  std::uniform_int_distribution<int64_t> pos(0, base.size - 1);
  std::uniform_int_distribution<int> byte(0, 255);
  const int32_t specials[] = {-2, -1, 0, 1, 2, 3, 7, 8, 63, 64, 255, 256, 1 << 20,
                              INT32_MAX, INT32_MIN};
  int accepted = 0;
  for (int it = 0; it < iters; ++it) {
    Buf b = base;
    const int nflip = 1 + (int)(rng() % 3);
    for (int j = 0; j < nflip; ++j) {
      if (rng() & 1) {
        b.data()[pos(rng)] = (uint8_t)byte(rng);
      } else {
        const int64_t p = pos(rng) & ~int64_t(3);
        if (p + 4 <= b.size) {
          const int32_t v = specials[rng() % (sizeof(specials) / sizeof(specials[0]))];
          std::memcpy(b.data() + p, &v, 4);
        }
      }
    }
    accepted += passes(b.data(), b.size) ? 1 : 0;
  }
  std::printf("fuzz: %d of %d corrupted buffers accepted (padding, in-range values)\n", accepted,
              iters);
}

int main() {
  std::mt19937 rng(11);
  int built = 0;
  for (int dim = 2; dim <= 4; ++dim) {
    for (int n : {1, 2, 9, 60}) {
      for (int leaf : {1, 3, 16}) {
        Buf f = make_tree<float>(n, dim, leaf, n / 5, rng);
        damage_cases<float>(f, dim);
        Buf d = make_tree<double>(n, dim, leaf, n / 5, rng);
        damage_cases<double>(d, dim);
        built += 2;
        if (n == 60 && leaf == 3) {
          fuzz(f, rng, 5000);
          fuzz(d, rng, 5000);
        }
      }
    }
  }
  if (g_failures) {
    std::printf("persist_check: %d failures\n", g_failures);
    return 1;
  }
  std::printf("persist_check: ok (%d trees)\n", built);
  return 0;
}
