// Persistent tree format: a device-resident tree as one flat byte buffer, to
// keep trees across process restarts (a server's resident trees) or send
// them over a wire without rebuilding. The layout and the host-side
// validation of a loaded buffer live in persist_format.hpp (plain C++, also
// built on its own under the sanitizers); this file writes buffers from a
// device tree and uploads validated ones.
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "persist_format.hpp"
#include "tree.hpp"

namespace pk {
namespace fmt {
void invalid(const std::string& msg) { fail(PPTK_E_VALUE, "%s", msg.c_str()); }
}  // namespace fmt

namespace {

using FileHeader = fmt::Header;
static_assert(sizeof(fmt::Node) == sizeof(NodeRec), "stored node = NodeRec");

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
    std::memcpy(h.magic, fmt::kMagic, 8);
    h.version = fmt::kVersion;
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
    fmt::layout(h);
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
    // the parser reads typed sections: a misaligned buffer is copied first
    std::vector<uint64_t> aligned;
    const uint8_t* b = (const uint8_t*)buf;
    if (buf && size > 0 && reinterpret_cast<uintptr_t>(buf) % 8 != 0) {
      aligned.resize((size_t)(size + 7) / 8);
      std::memcpy(aligned.data(), buf, (size_t)size);
      b = (const uint8_t*)aligned.data();
    }
    const FileHeader h = fmt::check(b, size);
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
