// Device helpers shared by the BVH builders (kern/tlas.hip, kern/blas.hip):
// Morton bit spreading, boxes and the writer of one 8-wide node in the
// arena format of dlayout.h.
#ifndef BVH_DEV_H
#define BVH_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../dlayout.h"

namespace bvhdev {

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

struct Box {
  float lo[3], hi[3];
};

__device__ __forceinline__ float half_area(const Box &b) {
  const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
  return x * y + y * z + z * x;
}

// One 8-wide node: bounds and words of `cnt` children, empty slots after.
// (All loops fully unrolled so the small arrays stay in registers.)
__device__ inline void write_node(dnode *out, const Box *bx, const uint32_t *word, uint32_t cnt) {
#pragma unroll
  for (uint32_t c = 0; c < 8; c++) {
    const bool e = c >= cnt;
    out->minx[c] = e ? 3.402823466e38f : bx[c].lo[0];
    out->maxx[c] = e ? -3.402823466e38f : bx[c].hi[0];
    out->miny[c] = e ? 3.402823466e38f : bx[c].lo[1];
    out->maxy[c] = e ? -3.402823466e38f : bx[c].hi[1];
    out->minz[c] = e ? 3.402823466e38f : bx[c].lo[2];
    out->maxz[c] = e ? -3.402823466e38f : bx[c].hi[2];
    out->child[c] = e ? 0u : word[c];
  }
#pragma unroll
  for (uint32_t o = 0; o < 8; o++) {
    float dist[8];
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) {
      // near corner along the octant direction (bvh8.c ordering)
      const float x = (o & 1) ? bx[c].lo[0] : bx[c].hi[0];
      const float y = (o & 2) ? bx[c].lo[1] : bx[c].hi[1];
      const float z = (o & 4) ? bx[c].lo[2] : bx[c].hi[2];
      dist[c] = c >= cnt ? 3.402823466e38f
                         : ((o & 1) ? x : -x) + ((o & 2) ? y : -y) + ((o & 4) ? z : -z);
    }
    // rank = position in the stable descending order (0 farthest, 7 nearest)
    uint32_t inv = 0;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++) {
      uint32_t r = 0;
#pragma unroll
      for (uint32_t k = 0; k < 8; k++)
        r += (dist[k] > dist[c] || (dist[k] == dist[c] && k < c)) ? 1u : 0u;
      inv |= r << (3 * c);
    }
    out->inv[o] = inv;
  }
}

}  // namespace bvhdev

#endif
