// Device BLAS build (opt-in, rend_hip_set_blas_build(REND_BLAS_GPU)): the
// triangle-level 8-wide BVH of one mesh built on the GPU and written in the
// arena format of dlayout.h, so the traversal kernel reads it unchanged.
//
// The reference builds every BLAS on the host with its SAH-binned build_bvh8
// (rend.c:458-709, called from rend_prepstatic, rend.c:3033-3071); that takes
// about half a second for a million triangles.  Here:
//   1. centroid bounds (block reductions, ordered-integer atomics);
//   2. 30-bit Morton codes of the centroids, made unique by the triangle
//      index in the low 32 bits, sorted with a device radix sort;
//   3. the binary radix tree over the sorted keys (Karras 2012, "Maximizing
//      parallelism in the construction of BVHs, octrees and k-d trees");
//   4. node bounds bottom-up: each leaf walks towards the root and the second
//      of two siblings to arrive computes the parent (device-scope fences
//      around a per-node counter);
//   5. collapse to 8-wide nodes breadth-first, one launch per level: a node
//      starts from a binary node's two children and replaces the child of
//      largest surface area that holds more than 4 triangles by its two
//      children until it has eight; children of at most 4 triangles become
//      leaves (the reference's leaf size, rend.c:521-538);
//   6. relocation into the arena: nodes first (256-byte aligned), then the
//      leaves, child words rewritten as byte offsets from the root.
// Leaves hold v0 and the edges v1 - v0, v2 - v0 in FP32 and the local
// triangle indices, unused lanes repeating the last triangle, as the
// reference's embed_leaf4 (rend.c:416-452) computes them, so every
// triangle test gives the reference's bits.  The tree differs from the SAH
// tree: hits are the same up to which of two triangles at exactly equal
// distance is reported (and rays that graze box boundaries of different
// extents), visit counts are not.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "../dlayout.h"
#include "bvh_dev.h"

namespace {

using bvhdev::Box;
using bvhdev::half_area;
using bvhdev::spread10;
using bvhdev::write_node;

constexpr uint32_t kLeafMax = 4;  // triangles per leaf (rend.c:521-538)

// 48-byte triangle record (rtri: v0, v1, v2 with a pad word each).
struct Tri {
  float4 a, b, c;
};

// Order-preserving float <-> uint32 (for atomic min/max of bounds).
__device__ __forceinline__ uint32_t ord_enc(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ Box tri_box(const Tri &t) {
  Box b;
  b.lo[0] = fminf(fminf(t.a.x, t.b.x), t.c.x);
  b.lo[1] = fminf(fminf(t.a.y, t.b.y), t.c.y);
  b.lo[2] = fminf(fminf(t.a.z, t.b.z), t.c.z);
  b.hi[0] = fmaxf(fmaxf(t.a.x, t.b.x), t.c.x);
  b.hi[1] = fmaxf(fmaxf(t.a.y, t.b.y), t.c.y);
  b.hi[2] = fmaxf(fmaxf(t.a.z, t.b.z), t.c.z);
  return b;
}

// ---- 1. centroid bounds ------------------------------------------------------

__global__ __launch_bounds__(256) void k_blas_cbounds(const Tri *__restrict__ tris, uint32_t n,
                                                      uint32_t *__restrict__ cb) {
  float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const Box b = tri_box(tris[i]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float c = 0.5f * (b.lo[k] + b.hi[k]);
      if (c == c) {
        lo[k] = fminf(lo[k], c);
        hi[k] = fmaxf(hi[k], c);
      }
    }
  }
  __shared__ float red[6][4];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
    }
  }
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      red[k][w] = lo[k];
      red[3 + k][w] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = (int)threadIdx.x;
    float l = red[k][0], h = red[3 + k][0];
    for (uint32_t j = 1; j < blockDim.x / 64u; j++) {
      l = fminf(l, red[k][j]);
      h = fmaxf(h, red[3 + k][j]);
    }
    atomicMin(&cb[k], ord_enc(l));
    atomicMax(&cb[3 + k], ord_enc(h));
  }
}

// ---- 2. keys -------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_blas_keys(const Tri *__restrict__ tris, uint32_t n,
                                                   const uint32_t *__restrict__ cb,
                                                   uint64_t *__restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Box b = tri_box(tris[i]);
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float lo = ord_dec(cb[k]), hi = ord_dec(cb[3 + k]);
    const float ext = hi - lo;
    const float c = 0.5f * (b.lo[k] + b.hi[k]);
    float f = ext > 0.0f ? (c - lo) / ext * 1023.0f : 0.0f;
    f = fminf(fmaxf(f, 0.0f), 1023.0f);  // also maps NaN to 0
    q[k] = (uint32_t)f;
  }
  const uint32_t m = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  keys[i] = ((uint64_t)m << 32) | i;
}

// ---- 3. binary radix tree --------------------------------------------------------
//
// Codes: 0 .. n-2 internal nodes (0 is the root), n-1 .. 2n-2 leaves (sorted
// position p has code n - 1 + p).

__device__ __forceinline__ int lcp(const uint64_t *__restrict__ key, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll(key[i] ^ key[j]);
}

__global__ __launch_bounds__(256) void k_blas_radix(const uint64_t *__restrict__ key, uint32_t n,
                                                    uint2 *__restrict__ child,
                                                    uint2 *__restrict__ range,
                                                    uint32_t *__restrict__ parent) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int ni = (int)n;
  if (i >= ni - 1) return;
  const int d = lcp(key, ni, i, i + 1) - lcp(key, ni, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = lcp(key, ni, i, i - d);
  int lmax = 2;
  while (lcp(key, ni, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (lcp(key, ni, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dn = lcp(key, ni, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    const int t = (l + div - 1) / div;
    if (lcp(key, ni, i, i + (s + t) * d) > dn) s += t;
    if (t <= 1) break;
  }
  const int g = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t left = lo == g ? (uint32_t)(ni - 1 + g) : (uint32_t)g;
  const uint32_t right = hi == g + 1 ? (uint32_t)(ni - 1 + g + 1) : (uint32_t)(g + 1);
  child[i] = make_uint2(left, right);
  range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
  parent[left] = (uint32_t)i;
  parent[right] = (uint32_t)i;
}

// ---- 4. bounds, bottom-up ----------------------------------------------------------

__global__ __launch_bounds__(256) void k_blas_boxes(const Tri *__restrict__ tris,
                                                    const uint64_t *__restrict__ key, uint32_t n,
                                                    const uint2 *__restrict__ child,
                                                    const uint32_t *__restrict__ parent,
                                                    float *box, uint32_t *visit) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t code = n - 1u + p;
  const Box b = tri_box(tris[(uint32_t)key[p]]);
  float *o = &box[(size_t)code * 6u];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o[k] = b.lo[k];
    o[3 + k] = b.hi[k];
  }
  while (code != 0u) {
    const uint32_t par = parent[code];
    // release this subtree's box, then count the arrival at the parent: the
    // second child to arrive sees the first one's box
    __threadfence();
    if (atomicAdd(&visit[par], 1u) == 0u) return;
    __threadfence();
    const uint2 c = child[par];
    const float *l = &box[(size_t)c.x * 6u], *r = &box[(size_t)c.y * 6u];
    float *q = &box[(size_t)par * 6u];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      q[k] = fminf(l[k], r[k]);
      q[3 + k] = fmaxf(l[3 + k], r[3 + k]);
    }
    code = par;
  }
}

// ---- 5. collapse to 8-wide nodes, one level per launch ---------------------------

struct Collapse {
  const Tri *tris;
  const uint64_t *key;
  uint32_t n;
  const uint2 *child;
  const uint2 *range;
  const float *box;
  dnode *nodes;        // temporary node array (root at index 0)
  dleaf *leaves;       // temporary leaf array
  const uint2 *qin;    // (binary code, node index) of this level
  uint2 *qout;         // next level
  uint32_t *ctr;       // [0] nodes, [1] leaves, [2] next-level entries
};

__device__ __forceinline__ uint32_t code_count(const Collapse &C, uint32_t code) {
  if (code >= C.n - 1u) return 1u;
  const uint2 r = C.range[code];
  return r.y - r.x + 1u;
}

__device__ __forceinline__ Box code_box(const Collapse &C, uint32_t code) {
  const float *b = &C.box[(size_t)code * 6u];
  return {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
}

// A leaf of the sorted triangles [first, first + cnt) (cnt 1..4).
__device__ void write_leaf(const Collapse &C, dleaf *lf, uint32_t first, uint32_t cnt) {
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    const uint32_t p = first + (k < cnt ? k : cnt - 1u);
    const uint32_t id = (uint32_t)C.key[p];
    const Tri t = C.tris[id];
    lf->v0x[k] = t.a.x;
    lf->v0y[k] = t.a.y;
    lf->v0z[k] = t.a.z;
    lf->e0x[k] = t.b.x - t.a.x;
    lf->e0y[k] = t.b.y - t.a.y;
    lf->e0z[k] = t.b.z - t.a.z;
    lf->e1x[k] = t.c.x - t.a.x;
    lf->e1y[k] = t.c.y - t.a.y;
    lf->e1z[k] = t.c.z - t.a.z;
    lf->id[k] = id;
  }
  lf->tricnt = cnt;
#pragma unroll
  for (int k = 0; k < 7; k++) lf->pad[k] = 0u;
}

__global__ __launch_bounds__(64) void k_blas_collapse(Collapse C, uint32_t qn) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= qn) return;
  const uint2 ent = C.qin[e];
  uint32_t code[8];
  uint32_t cnt = 0;
  if (ent.x < C.n - 1u) {
    const uint2 c = C.child[ent.x];
    code[0] = c.x;
    code[1] = c.y;
    cnt = 2;
  } else {
    code[0] = ent.x;  // a one-triangle tree
    cnt = 1;
  }
  // expand the largest child of more than kLeafMax triangles until 8
  while (cnt < 8u) {
    int best = -1;
    float ba = -1.0f;
    for (uint32_t k = 0; k < cnt; k++) {
      if (code_count(C, code[k]) <= kLeafMax) continue;
      const float a = half_area(code_box(C, code[k]));
      if (best < 0 || a > ba) {
        best = (int)k;
        ba = a;
      }
    }
    if (best < 0) break;
    const uint2 c = C.child[code[best]];
    code[best] = c.x;
    code[cnt++] = c.y;
  }
  Box bx[8];
  uint32_t word[8];
  for (uint32_t k = 0; k < 8; k++) {
    if (k >= cnt) {
      word[k] = 0u;
      continue;
    }
    bx[k] = code_box(C, code[k]);
    const uint32_t m = code_count(C, code[k]);
    if (m <= kLeafMax) {
      const uint32_t first = code[k] >= C.n - 1u ? code[k] - (C.n - 1u) : C.range[code[k]].x;
      const uint32_t li = atomicAdd(&C.ctr[1], 1u);
      write_leaf(C, &C.leaves[li], first, m);
      word[k] = DNODE_LEAF | li;
    } else {
      const uint32_t ni = atomicAdd(&C.ctr[0], 1u);
      C.qout[atomicAdd(&C.ctr[2], 1u)] = make_uint2(code[k], ni);
      word[k] = ni;
    }
  }
  write_node(&C.nodes[ent.y], bx, word, cnt);
}

// ---- 6. relocation into the arena -----------------------------------------------

__global__ __launch_bounds__(256) void k_blas_reloc(const dnode *__restrict__ nodes, uint32_t nn,
                                                    const dleaf *__restrict__ leaves, uint32_t nl,
                                                    uint8_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) {
    dnode d = nodes[i];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t w = d.child[c];
      // empty slots keep 0; the root is never a child, so 0 is free
      d.child[c] = (w & DNODE_LEAF) ? (DNODE_LEAF | (nn * 256u + (w & ~DNODE_LEAF) * 192u))
                                    : w * 256u;
    }
    ((dnode *)out)[i] = d;
  } else if (i - nn < nl) {
    const uint32_t k = i - nn;
    *(dleaf *)(out + (size_t)nn * 256u + (size_t)k * 192u) = leaves[k];
  }
}

// Centroid bounds to their identities, the counters (the root node is
// allocated) and the first level's queue: the root.
__global__ void k_blas_init(uint32_t *cb, uint32_t *ctr, uint2 *q, uint32_t root) {
  if (threadIdx.x < 3) {
    cb[threadIdx.x] = 0xffffffffu;
    cb[3 + threadIdx.x] = 0u;
    ctr[threadIdx.x] = threadIdx.x == 0 ? 1u : 0u;
  }
  if (threadIdx.x == 0) q[0] = make_uint2(root, 0u);
}

__global__ void k_blas_next_level(uint32_t *ctr) {
  if (threadIdx.x == 0) ctr[2] = 0u;
}

// A root without children (a mesh without triangles).
__global__ void k_blas_empty(dnode *root) {
  if (threadIdx.x == 0) {
    Box bx[8];
    uint32_t word[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    write_node(root, bx, word, 0);
  }
}

__global__ void k_blas_fill(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

// ---- C-ABI ------------------------------------------------------------------------

namespace {
// Workspace layout for n triangles (bytes, each part 256-byte aligned).
struct Ws {
  size_t keys, keys2, child, range, parent, box, visit, q0, q1, nodes, leaves, ctl, sort, total;
};

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

Ws ws_layout(uint32_t n, size_t sort_bytes) {
  const size_t m = n ? n : 1;
  const size_t maxnodes = m / 2 + 2;
  Ws w;
  size_t o = 0;
  w.keys = o;   o += al(m * 8);
  w.keys2 = o;  o += al(m * 8);
  w.child = o;  o += al(m * 8);
  w.range = o;  o += al(m * 8);
  w.parent = o; o += al(2 * m * 4);
  w.box = o;    o += al(2 * m * 24);
  w.visit = o;  o += al(m * 4);
  w.q0 = o;     o += al(maxnodes * 8);
  w.q1 = o;     o += al(maxnodes * 8);
  w.nodes = o;  o += al(maxnodes * sizeof(dnode));
  w.leaves = o; o += al(m * sizeof(dleaf));
  w.ctl = o;    o += al(16 * 4);
  w.sort = o;   o += al(sort_bytes);
  w.total = o;
  return w;
}

size_t sort_temp_bytes(uint32_t n) {
  size_t b = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, b, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                    (int)(n ? n : 1), 0, 62);
  return b;
}
}  // namespace

// Device workspace bytes for a BLAS of n triangles.
extern "C" unsigned long long rk_blas_ws_bytes(unsigned n) {
  return ws_layout(n, sort_temp_bytes(n)).total;
}

// Builds the BLAS of n triangles (rtri records at tris) into the workspace;
// on return (stream synchronised) counts[0] = nodes, counts[1] = leaves.
// Returns a hipError_t value, or -1 when the node array would overflow.
extern "C" int rk_blas_build(const void *tris, unsigned n, void *ws, unsigned counts[2],
                             void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t sb = sort_temp_bytes(n);
  const Ws w = ws_layout(n, sb);
  uint8_t *b = (uint8_t *)ws;
  const Tri *t = (const Tri *)tris;
  uint64_t *keys = (uint64_t *)(b + w.keys), *keys2 = (uint64_t *)(b + w.keys2);
  uint2 *child = (uint2 *)(b + w.child), *range = (uint2 *)(b + w.range);
  uint32_t *parent = (uint32_t *)(b + w.parent), *visit = (uint32_t *)(b + w.visit);
  float *box = (float *)(b + w.box);
  uint2 *q[2] = {(uint2 *)(b + w.q0), (uint2 *)(b + w.q1)};
  dnode *nodes = (dnode *)(b + w.nodes);
  dleaf *leaves = (dleaf *)(b + w.leaves);
  uint32_t *ctl = (uint32_t *)(b + w.ctl);  // [0..5] centroid bounds, [8..10] counters
  const uint32_t maxnodes = (n ? n : 1) / 2 + 2;
  if (n == 0) {
    hipLaunchKernelGGL(k_blas_empty, dim3(1), dim3(64), 0, s, nodes);
    counts[0] = 1;
    counts[1] = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return (int)e;
  }
  const uint32_t gb = (n + 255u) / 256u;
  // root code: internal node 0, or the one leaf of a one-triangle mesh
  hipLaunchKernelGGL(k_blas_init, dim3(1), dim3(64), 0, s, ctl, ctl + 8, q[0],
                     n > 1 ? 0u : n - 1u);
  hipLaunchKernelGGL(k_blas_cbounds, dim3(gb < 1024u ? gb : 1024u), dim3(256), 0, s, t, n, ctl);
  hipLaunchKernelGGL(k_blas_keys, dim3(gb), dim3(256), 0, s, t, n, ctl, keys);
  size_t tb = sb;
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(b + w.sort, tb, keys, keys2, (int)n, 0, 62, s);
  if (e != hipSuccess) return (int)e;
  if (n > 1) {
    hipLaunchKernelGGL(k_blas_radix, dim3((n - 1u + 255u) / 256u), dim3(256), 0, s, keys2, n,
                       child, range, parent);
    hipLaunchKernelGGL(k_blas_fill, dim3(256), dim3(256), 0, s, visit, (size_t)n, 0u);
  }
  hipLaunchKernelGGL(k_blas_boxes, dim3(gb), dim3(256), 0, s, t, keys2, n, child, parent, box,
                     visit);
  // level-by-level collapse; the host reads each level's size
  uint32_t h[3] = {1u, 0u, 0u};
  e = hipGetLastError();
  uint32_t qn = 1, cur = 0;
  Collapse C;
  C.tris = t;
  C.key = keys2;
  C.n = n;
  C.child = child;
  C.range = range;
  C.box = box;
  C.nodes = nodes;
  C.leaves = leaves;
  C.ctr = ctl + 8;
  while (e == hipSuccess && qn) {
    C.qin = q[cur];
    C.qout = q[cur ^ 1u];
    hipLaunchKernelGGL(k_blas_collapse, dim3((qn + 63u) / 64u), dim3(64), 0, s, C, qn);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, ctl + 8, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) break;
    if (h[0] > maxnodes || h[1] > n) return -1;
    qn = h[2];
    cur ^= 1u;
    if (qn) {
      hipLaunchKernelGGL(k_blas_next_level, dim3(1), dim3(64), 0, s, ctl + 8);
      e = hipGetLastError();
    }
  }
  if (e != hipSuccess) return (int)e;
  counts[0] = h[0];
  counts[1] = h[1];
  return 0;
}

// Copies a built BLAS (rk_blas_build's workspace and counts) to out:
// counts[0] nodes of 256 bytes, then counts[1] leaves of 192 bytes.
extern "C" int rk_blas_reloc(const void *ws, unsigned n, const unsigned counts[2], void *out,
                             void *stream) {
  const Ws w = ws_layout(n, sort_temp_bytes(n));
  const uint8_t *b = (const uint8_t *)ws;
  const uint32_t tot = counts[0] + counts[1];
  hipLaunchKernelGGL(k_blas_reloc, dim3((tot + 255u) / 256u), dim3(256), 0, (hipStream_t)stream,
                     (const dnode *)(b + w.nodes), counts[0], (const dleaf *)(b + w.leaves),
                     counts[1], (uint8_t *)out);
  return (int)hipGetLastError();
}
