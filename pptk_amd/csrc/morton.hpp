// Morton (Z-order) codes shared by the tree build and the query sort.
// Coordinates are quantised on one scale for every axis (cubic cells) to
// 21 bits per axis in 3-D (31 in 2-D, 15 in 4-D) and bit-interleaved.
#pragma once
#include "common.hpp"

namespace pk {

template <int DIM> struct MortonBits { static constexpr int value = DIM == 2 ? 31 : DIM == 3 ? 21 : 15; };

// Bounding box, valid-point count and quantisation scale of a point set.
template <typename D>
struct BuildInfo {
  uint32_t n_valid;
  uint32_t pad;
  D lo[4];
  D hi[4];
  D scale[4];
};

__device__ __forceinline__ uint64_t spread3_21(uint32_t a) {
  uint64_t x = a & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint64_t spread2_31(uint32_t a) {
  uint64_t x = a & 0x7fffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

template <typename D, int DIM>
__device__ __forceinline__ uint64_t morton_code(const D* c, const BuildInfo<D>& in) {
  constexpr int bits = MortonBits<DIM>::value;
  const D maxq = (D)((1u << bits) - 1u);
  // maxq rounds up in float for 31 bits; clamp again as an integer.
  constexpr uint32_t umax = (1u << bits) - 1u;
  uint32_t u[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    D x = (c[d] - in.lo[d]) * in.scale[d];
    x = x > D(0) ? x : D(0);
    x = x < maxq ? x : maxq;
    const uint32_t ux = (uint32_t)x;
    u[d] = ux < umax ? ux : umax;
  }
  if constexpr (DIM == 3) {
    return spread3_21(u[0]) | (spread3_21(u[1]) << 1) | (spread3_21(u[2]) << 2);
  } else if constexpr (DIM == 2) {
    return spread2_31(u[0]) | (spread2_31(u[1]) << 1);
  } else {
    uint64_t code = 0;
    for (int b = 0; b < bits; ++b)
#pragma unroll
      for (int d = 0; d < DIM; ++d) code |= (uint64_t)((u[d] >> b) & 1u) << (b * DIM + d);
    return code;
  }
}

}  // namespace pk
