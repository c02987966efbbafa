// Device TLAS build: the instance-level 8-wide BVH from the instances' world
// AABBs, written in the arena's node format (dlayout.h) so the traversal
// kernel reads it unchanged.
//
// The reference rebuilds its TLAS on the host every frame (rend_prepdynamic,
// rend.c:3073-3090, with the SAH-binned build_bvh8, rend.c:458-709).  Here the
// whole build runs in one 1024-thread workgroup, which is enough for the
// 12-bit instance id limit (4,096 instances) and needs no grid-wide
// synchronisation:
//   1. Morton codes (10 bits per axis) of the instance centroids inside their
//      bounds, made unique by the instance index in the low bits; bitonic sort
//      in LDS.
//   2. Binary radix tree over the sorted keys (Karras 2012, "Maximizing
//      parallelism in the construction of BVHs, octrees and k-d trees").
//   3. Internal node bounds level by level, deepest level first, kept in LDS
//      (they replace the keys, which the tree no longer needs; the whole
//      build lives in 152 KB of LDS).
//   4. Collapse to 8-wide nodes breadth-first: starting from a binary node's
//      two children, the child with the largest surface area is replaced by
//      its two children until there are eight or only leaves remain (the same
//      rule as the reference's horizontal phase, rend.c:491-573).  Child
//      bounds, child words (byte offset of a child node, or NODE_LEAF |
//      instance id) and the per-octant traversal ranks follow the host repack
//      (bvh8.c), empty slots hold the reference's empty box and a zero word.
// The tree differs from the SAH tree, so hits are the same (up to which of
// two instances at exactly equal distance is reported) but visit counts are
// not.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../dlayout.h"
#include "bvh_dev.h"

namespace {

using bvhdev::Box;
using bvhdev::half_area;
using bvhdev::spread10;
using bvhdev::write_node;

constexpr uint32_t kMax = 4096;       // instances (12-bit instance ids)
constexpr uint32_t kThreads = 1024;
constexpr uint16_t kLeaf = 0x8000u;   // child code: leaf (sorted position) flag

struct BuildLds {
  union {
    uint64_t key[kMax];               // sorted Morton | instance keys (steps 1-2)
    float nb[kMax][6];                // internal node bounds (steps 3-4)
  };
  uint16_t inst[kMax];                // instance at each sorted position
  uint16_t left[kMax], right[kMax];   // binary children (kLeaf | position)
  uint16_t parent[kMax];              // parent of an internal node
  uint8_t depth[kMax];                // internal node depth
  uint16_t q[2][kMax];                // breadth-first queues of binary nodes
  float red[6][kThreads / 64];
  uint32_t cnt, maxd, err;
};

// Longest common prefix of sorted keys i and j, -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t *key, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll(key[i] ^ key[j]);
}

__device__ __forceinline__ Box leaf_box(const float4 *ab, const uint16_t *inst, uint32_t pos) {
  const uint32_t k = inst[pos];
  const float4 a = ab[2 * k], b = ab[2 * k + 1];
  return {{a.x, a.y, a.z}, {b.x, b.y, b.z}};
}

__device__ __forceinline__ Box node_box(const float (*nb)[6], uint32_t i) {
  return {{nb[i][0], nb[i][1], nb[i][2]}, {nb[i][3], nb[i][4], nb[i][5]}};
}

__device__ __forceinline__ Box child_box(const float4 *ab, const BuildLds &S, uint16_t c) {
  return (c & kLeaf) ? leaf_box(ab, S.inst, c & ~kLeaf) : node_box(S.nb, c);
}

__device__ __forceinline__ float child_area(const BuildLds &S, uint16_t c) {
  return (c & kLeaf) ? -1.0f : half_area(node_box(S.nb, c));
}

__device__ __forceinline__ float block_min(BuildLds &S, int slot, float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63u) == 0) S.red[slot][threadIdx.x >> 6] = v;
  __syncthreads();
  float r = S.red[slot][0];
  for (uint32_t w = 1; w < kThreads / 64; w++) r = fminf(r, S.red[slot][w]);
  return r;
}

__global__ __launch_bounds__(1024) void k_tlas_build(const float4 *__restrict__ ab, uint32_t n,
                                                     dnode *__restrict__ out,
                                                     uint32_t *__restrict__ info) {
  __shared__ BuildLds S;
  const uint32_t tid = threadIdx.x;
  if (n < 2) {
    // root with zero or one instance leaf (build_bvh8 also makes a root node)
    if (tid == 0) {
      Box b[8] = {};
      uint32_t w[8] = {DNODE_LEAF, 0, 0, 0, 0, 0, 0, 0};
      if (n) {
        const float4 a = ab[0], c = ab[1];
        b[0] = {{a.x, a.y, a.z}, {c.x, c.y, c.z}};
      }
      write_node(out, b, w, n);
      info[0] = 1;
      info[1] = 0;
    }
    return;
  }
  // ---- 1. centroid bounds and keys
  float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  for (uint32_t i = tid; i < n; i += kThreads) {
    const float4 a = ab[2 * i], b = ab[2 * i + 1];
    const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    for (int k = 0; k < 3; k++) {
      lo[k] = fminf(lo[k], c[k]);
      hi[k] = fmaxf(hi[k], c[k]);
    }
  }
  float cmin[3], cmax[3];
  for (int k = 0; k < 3; k++) {
    cmin[k] = block_min(S, k, lo[k]);
    cmax[k] = -block_min(S, 3 + k, -hi[k]);
  }
  uint32_t npow = 2;
  while (npow < n) npow <<= 1;
  for (uint32_t i = tid; i < npow; i += kThreads) {
    uint64_t key = ~0ull;
    if (i < n) {
      const float4 a = ab[2 * i], b = ab[2 * i + 1];
      const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
      uint32_t q[3];
      for (int k = 0; k < 3; k++) {
        const float ext = cmax[k] - cmin[k];
        float f = ext > 0.0f ? (c[k] - cmin[k]) / ext * 1023.0f : 0.0f;
        f = fminf(fmaxf(f, 0.0f), 1023.0f);  // also maps NaN to 0
        q[k] = (uint32_t)f;
      }
      const uint32_t m = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
      key = ((uint64_t)m << 32) | i;
    }
    S.key[i] = key;
  }
  if (tid == 0) {
    S.cnt = 0;
    S.maxd = 0;
    S.err = 0;
  }
  __syncthreads();
  // ---- bitonic sort, ascending
  for (uint32_t k = 2; k <= npow; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = tid; t < npow / 2; t += kThreads) {
        const uint32_t i = 2 * j * (t / j) + (t % j);  // lower element of the pair
        const uint32_t l = i + j;
        const bool up = (i & k) == 0;
        const uint64_t a = S.key[i], b = S.key[l];
        if ((a > b) == up) {
          S.key[i] = b;
          S.key[l] = a;
        }
      }
      __syncthreads();
    }
  }
  // ---- 2. radix tree: internal nodes 0..n-2, root 0
  const int ni = (int)n;
  for (uint32_t t = tid; t + 1 < n; t += kThreads) {
    const int i = (int)t;
    const int d = delta(S.key, ni, i, i + 1) - delta(S.key, ni, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(S.key, ni, i, i - d);
    int lmax = 2;
    while (delta(S.key, ni, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int s = lmax >> 1; s > 0; s >>= 1)
      if (delta(S.key, ni, i, i + (l + s) * d) > dmin) l += s;
    const int j = i + l * d;
    const int dn = delta(S.key, ni, i, j);
    int s = 0;
    for (int w = (l + 1) >> 1;; w = (w + 1) >> 1) {
      if (delta(S.key, ni, i, i + (s + w) * d) > dn) s += w;
      if (w <= 1) break;
    }
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int first = i < j ? i : j, last = i < j ? j : i;
    const uint16_t lc = first == g ? (uint16_t)(kLeaf | g) : (uint16_t)g;
    const uint16_t rc = last == g + 1 ? (uint16_t)(kLeaf | (g + 1)) : (uint16_t)(g + 1);
    S.left[i] = lc;
    S.right[i] = rc;
    if (!(lc & kLeaf)) S.parent[lc] = (uint16_t)i;
    if (!(rc & kLeaf)) S.parent[rc] = (uint16_t)i;
  }
  for (uint32_t p = tid; p < n; p += kThreads) S.inst[p] = (uint16_t)S.key[p];
  if (tid == 0) S.parent[0] = 0xffffu;
  __syncthreads();
  // ---- 3. depths, then bounds from the deepest level up (the keys are dead
  // from here on: S.nb overwrites them)
  for (uint32_t i = tid; i + 1 < n; i += kThreads) {
    uint32_t d = 0, p = i;
    while (p != 0 && p < kMax && d < kMax) {
      p = S.parent[p];
      d++;
    }
    if (d > 255 || p != 0) {
      atomicOr(&S.err, 1u);
      d = 255;
    }
    S.depth[i] = (uint8_t)d;
    atomicMax(&S.maxd, d);
  }
  __syncthreads();
  if (S.err) {
    if (tid == 0) {
      info[0] = 0;
      info[1] = 1;
    }
    return;
  }
  for (int lv = (int)S.maxd; lv >= 0; lv--) {
    for (uint32_t i = tid; i + 1 < n; i += kThreads) {
      if (S.depth[i] != (uint32_t)lv) continue;
      const Box a = child_box(ab, S, S.left[i]);
      const Box b = child_box(ab, S, S.right[i]);
      for (int k = 0; k < 3; k++) {
        S.nb[i][k] = fminf(a.lo[k], b.lo[k]);
        S.nb[i][3 + k] = fmaxf(a.hi[k], b.hi[k]);
      }
    }
    __syncthreads();
  }
  // ---- 4. collapse to 8-wide nodes, breadth first
  uint32_t qn = 1, base = 0, cur = 0;
  if (tid == 0) S.q[0][0] = 0;
  __syncthreads();
  for (uint32_t level = 0; qn > 0 && level < kMax; level++) {
    for (uint32_t e = tid; e < qn; e += kThreads) {
      // child codes and areas (leaves -1); indexed only by unrolled
      // constants so they stay in registers
      uint16_t ch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      float area[8] = {-1.0f, -1.0f, -1.0f, -1.0f, -1.0f, -1.0f, -1.0f, -1.0f};
      const uint16_t b = S.q[cur][e];
      ch[0] = S.left[b];
      ch[1] = S.right[b];
      area[0] = child_area(S, ch[0]);
      area[1] = child_area(S, ch[1]);
      uint32_t cnt = 2;
#pragma unroll
      for (uint32_t it = 0; it < 6; it++) {  // at most 6 splits from 2 to 8
        int best = -1;
        float ba = -1.0f;
#pragma unroll
        for (uint32_t k = 0; k < 8; k++)
          if (k < cnt && !(ch[k] & kLeaf) && area[k] > ba) {
            ba = area[k];
            best = (int)k;
          }
        if (best < 0) break;
        uint16_t x = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; k++)
          if ((int)k == best) x = ch[k];
        const uint16_t l = S.left[x], r = S.right[x];
        const float al = child_area(S, l), ar = child_area(S, r);
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) {
          if ((int)k == best) { ch[k] = l; area[k] = al; }
          if (k == cnt) { ch[k] = r; area[k] = ar; }
        }
        cnt++;
      }
      Box bx[8] = {};
      uint32_t word[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (uint32_t k = 0; k < 8; k++) {
        if (k >= cnt) continue;
        bx[k] = child_box(ab, S, ch[k]);
        if (ch[k] & kLeaf) {
          word[k] = DNODE_LEAF | (uint32_t)S.inst[ch[k] & ~kLeaf];
        } else {
          const uint32_t slot = atomicAdd(&S.cnt, 1u);
          S.q[cur ^ 1][slot] = ch[k];
          word[k] = (base + qn + slot) * (uint32_t)sizeof(dnode);
        }
      }
      write_node(out + base + e, bx, word, cnt);
    }
    __syncthreads();
    base += qn;
    qn = S.cnt;
    cur ^= 1;
    __syncthreads();
    if (tid == 0) S.cnt = 0;
    __syncthreads();
  }
  if (tid == 0) {
    info[0] = base;
    info[1] = qn ? 2u : 0u;
  }
}

}  // namespace

extern "C" int rk_tlas_build(const void *aabbs, unsigned n, void *nodes, unsigned *info,
                             void *stream) {
  if (n > kMax) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tlas_build, dim3(1), dim3(kThreads), 0, (hipStream_t)stream,
                     (const float4 *)aabbs, n, (dnode *)nodes, info);
  return (int)hipGetLastError();
}
