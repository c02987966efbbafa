// Closest-hit and any-hit traversal of the two-level BVH8 on CDNA4 (gfx950).
//
// Semantics follow the reference's single-ray traversal operation by
// operation: intersect_tlas (rend.c:1976-2155), the instance transform
// mulpos/muldir (rend.c:171-187), intersect_blas (rend.c:714-968) and the
// any-hit pair intersect_any_tlas/blas (rend.c:2883-2998, 1797-1973):
//   * slab test t = fma(plane, 1/d, -o/d) with x86 MAXPS/MINPS rules (the
//     second operand wins on NaN) and an ordered <= test;
//   * children pushed in the node's per-octant rank order, nearest visited
//     first; stack entries whose entry distance exceeds the current t are
//     skipped when popped (the reference compacts them eagerly, which removes
//     the same entries);
//   * Moeller-Trumbore with the reference's FMA shapes and the host CPU's
//     RCPSS reciprocal reproduced from a 2048-entry table;
//   * among hits in a leaf, the smallest t and, on ties, the highest lane.
// The file is compiled with -ffp-contract=off; fused operations are explicit.
//
// Execution model: one ray per lane, wave64, persistent workgroups.  The rays
// are split into RK_NSEG segments, one per XCD; a wave claims RK_CHUNK rays
// at a time from its own XCD's segment with one atomic (and steals from the
// other segments once its own is drained), then hands claimed rays to idle
// lanes whenever at least `refill` lanes are idle (ballot + mbcnt prefix gives
// each idle lane its ray).  The stack's first DEPTH entries live in LDS,
// deeper entries in a per-thread global spill area; DEPTH and the waves per
// SIMD the register budget must allow are compile-time variants.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../dlayout.h"
#include "../rcp_table.h"
#include "rend_rays.h"

#define FLT_MAXF 3.402823466e+38f

// RCPSS mantissa tables (12 leading input bits): the one captured on the
// build host (an Intel Xeon, for results reproduced from fixtures made
// there) and this host's, captured by the library at set-up
// (rk_trace_set_host_rcp); RK_TRACE_XEON selects the former.
__constant__ uint32_t c_rcp_xeon[4096] = RCP_TABLE_INIT;
__constant__ uint16_t c_rcp_host[4096];

namespace {

struct RayS {
  float ox, oy, oz, dx, dy, dz;
  float idx, idy, idz, rx, ry, rz;
  uint32_t oct;
  // 1/d and o/d all finite: slab distances are then never NaN, so IEEE
  // max/min give the same comparisons as the x86 operand rules
  bool fin;
};

typedef __attribute__((address_space(1))) uint32_t gu32;  // global memory

__device__ __forceinline__ float xmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float xmin(float a, float b) { return a < b ? a : b; }

__device__ __forceinline__ void ray_setup(RayS &r, float ox, float oy, float oz,
                                          float dx, float dy, float dz) {
  r.ox = ox; r.oy = oy; r.oz = oz;
  r.dx = dx; r.dy = dy; r.dz = dz;
  r.idx = 1.0f / dx;
  r.idy = 1.0f / dy;
  r.idz = 1.0f / dz;
  r.rx = ox * r.idx;
  r.ry = oy * r.idy;
  r.rz = oz * r.idz;
  r.oct = (dz >= 0.0f ? 4u : 0u) | (dy >= 0.0f ? 2u : 0u) | (dx >= 0.0f ? 1u : 0u);
  r.fin = __builtin_isfinite(r.idx) & __builtin_isfinite(r.idy) & __builtin_isfinite(r.idz) &
          __builtin_isfinite(r.rx) & __builtin_isfinite(r.ry) & __builtin_isfinite(r.rz);
}

// An instance-space direction component can cancel to exactly zero (about
// one in 1e8 products of a random rotation); 1/0 = inf then meets o*inf and
// the slab distances go NaN, which the x86 rules count as a hit on every box,
// so the ray walks its whole BLAS.  Path-tracing callers ask for such a
// component to be replaced by a signed 1e-20 (geometrically the same
// axis-parallel ray); the reference-exact entry points leave it alone.
__device__ __forceinline__ float nudge0(float x) {
  return x == 0.0f ? __builtin_copysignf(1e-20f, x) : x;
}

// x86 RCPSS from a captured mantissa table indexed by the 12 leading
// mantissa bits (rule proven exhaustively by gen_rcp_table.c on Intel and
// AMD cores).  Entries keep 12 significant bits (the low 11 mantissa bits of
// every result are zero), so LDS holds them as 16-bit values.
__device__ __forceinline__ float rcp_table(float x, const uint16_t *tab) {
  uint32_t u = __float_as_uint(x);
  uint32_t s = u & 0x80000000u;
  uint32_t e = (u >> 23) & 0xffu;
  uint32_t m = u & 0x7fffffu;
  uint32_t r = s | ((253u - e) << 23) | ((uint32_t)tab[m >> 11] << 11);
  r = (e >= 253u) ? s : r;
  r = (e == 255u && m != 0u) ? (u | 0x00400000u) : r;
  r = (e == 0u) ? (s | 0x7f800000u) : r;
  return __uint_as_float(r);
}

// Sticky error bits for the host, and the first failing ray's record.
__device__ __noinline__ void report_error(uint32_t *diag, uint32_t code, const rray *rays,
                                          uint32_t rid, uint32_t iter) {
  atomicOr(&diag[0], code);
  if (atomicCAS(&diag[1], 0u, 1u) == 0u) {
    const uint4 *r = (const uint4 *)&rays[rid];
    uint4 *d = (uint4 *)&diag[4];
    d[0] = r[0];
    d[1] = r[1];
    d[2] = make_uint4(rid, iter, code, 0u);
  }
}

// One step of a quad reduction to the nearest hit: exchange with the lane
// given by the DPP quad permutation CTRL; hi names the lane holding the
// higher triangle index of the pair, which wins equal distances.  No value
// is NaN (a miss carries FLT_MAX), so the comparisons are plain.
template <int CTRL>
__device__ __forceinline__ void quad_nearest(uint32_t hi, float &tb, float &u, float &v,
                                             uint32_t &id) {
  const float pt = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(tb), CTRL, 0xf, 0xf, false));
  const float pu = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(u), CTRL, 0xf, 0xf, false));
  const float pv = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
  const uint32_t pid = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)id, CTRL, 0xf, 0xf, false);
  const bool take = hi ? (pt < tb) : (pt <= tb);
  tb = take ? pt : tb;
  u = take ? pu : u;
  v = take ? pv : v;
  id = take ? pid : id;
}

// Loads a node and runs the slab test of its 8 children: entry distances
// into tmn, child words into ch, the octant's rank word into invw; returns
// the hit mask indexed by rank.  FAST: every slab distance is known to be
// finite (no NaN), so IEEE max/min (v_max3/v_min3) make the same decisions
// as the reference's x86 MAXPS/MINPS operand rules.
template <bool FAST, bool HOIST>
__device__ __forceinline__ uint32_t slab8(const uint8_t *arena, uint32_t nofs, const RayS &cr,
                                          float t, float tmn[8], uint32_t ch[8],
                                          uint32_t &invw) {
  // The node sits at the 32-bit arena offset nofs (256-byte aligned).  Every
  // address is the arena base plus a 32-bit offset, which the loads take as
  // a scalar base and a 32-bit vector offset (no 64-bit address arithmetic
  // per lane).  Near/far bound rows of each axis: 32-byte rows (min, max).
  const uint32_t ax = (cr.oct & 1u) ? 0u : 32u, ay = (cr.oct & 2u) ? 64u : 96u,
                 az = (cr.oct & 4u) ? 128u : 160u;
  auto row = [&](uint32_t o) { return (const float4 *)(arena + (uint32_t)(nofs + o)); };
  float nxv[8], fxv[8], nyv[8], fyv[8], nzv[8], fzv[8];
  const float4 *nx = row(ax), *fx = row(ax ^ 32u), *ny = row(ay), *fy = row(ay ^ 32u),
               *nz = row(az), *fz = row(az ^ 32u);
  const uint4 *cw = (const uint4 *)(arena + (uint32_t)(nofs + (uint32_t)offsetof(dnode, child)));
#pragma unroll
  for (int k = 0; k < 2; k++) {
    *(float4 *)&nxv[4 * k] = nx[k];
    *(float4 *)&fxv[4 * k] = fx[k];
    *(float4 *)&nyv[4 * k] = ny[k];
    *(float4 *)&fyv[4 * k] = fy[k];
    *(float4 *)&nzv[4 * k] = nz[k];
    *(float4 *)&fzv[4 * k] = fz[k];
    *(uint4 *)&ch[4 * k] = cw[k];
  }
  invw = *(const uint32_t *)(arena + (uint32_t)(nofs + 4u * cr.oct +
                                                (uint32_t)offsetof(dnode, inv)));
  uint32_t hmr = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const float t0x = __builtin_fmaf(nxv[c], cr.idx, -cr.rx);
    const float t0y = __builtin_fmaf(nyv[c], cr.idy, -cr.ry);
    const float t0z = __builtin_fmaf(nzv[c], cr.idz, -cr.rz);
    const float t1x = __builtin_fmaf(fxv[c], cr.idx, -cr.rx);
    const float t1y = __builtin_fmaf(fyv[c], cr.idy, -cr.ry);
    const float t1z = __builtin_fmaf(fzv[c], cr.idz, -cr.rz);
    float lo, hi;
    if (FAST) {
      lo = __builtin_fmaxf(__builtin_fmaxf(__builtin_fmaxf(t0x, t0y), t0z), 0.0f);
      hi = __builtin_fminf(__builtin_fminf(__builtin_fminf(t1x, t1y), t1z), t);
    } else {
      lo = xmax(xmax(xmax(t0x, t0y), t0z), 0.0f);
      hi = xmin(xmin(xmin(t1x, t1y), t1z), t);
    }
    tmn[c] = lo;
    const uint32_t rank = (invw >> (3 * c)) & 7u;
    hmr |= (lo <= hi ? 1u : 0u) << rank;
  }
  // The child words are only read when a child is hit; left alone, the
  // compiler sinks their load into that branch, after the slab test, and a
  // visit waits for a second memory round trip.  HOIST keeps the load with
  // the bounds (-2 to -5 % per launch) where the register budget allows it:
  // with the host-table reciprocal the kernel then spills to scratch.
  if (HOIST)
    asm volatile("" ::"v"(ch[0]), "v"(ch[1]), "v"(ch[2]), "v"(ch[3]), "v"(ch[4]), "v"(ch[5]),
                 "v"(ch[6]), "v"(ch[7]));
  return hmr;
}

// DEPTH: LDS stack entries per lane; MINW: minimum waves per SIMD the
// register allocation must allow (occupancy variants; LDS per block is
// 8 KB table + (DEPTH + 1) * 2 KB).
template <bool CLOSEST, bool COUNT, bool RCP_IEEE, int DEPTH, int MINW>
__global__ __launch_bounds__(REND_TRACE_BLOCK, MINW)
void k_trace(const uint8_t *__restrict__ arena, uint64_t tlas,
             const dinst *__restrict__ insts, const rray *__restrict__ rays,
             rhit *__restrict__ hits, uint8_t *__restrict__ occ, uint32_t n,
             uint32_t *__restrict__ ctl, uint2 *__restrict__ spill,
             unsigned long long *__restrict__ stats, uint32_t refill,
             uint32_t max_iter, const uint32_t *__restrict__ n_dev,
             uint32_t *__restrict__ diag, uint32_t flags) {
  __shared__ uint16_t s_tab[RCP_IEEE ? 2 : 4096];
  if (n_dev) n = *n_dev;  // queue length produced by an earlier kernel
  // stack rows [0, DEPTH) per lane, plus one junk row that branch-free
  // pushes write for children they do not keep.  Entry words (child offset)
  // and entry distances are separate arrays: a push then stores two
  // independent registers, where a packed pair would first be copied into
  // adjacent registers.  The any-hit pass keeps no distances.
  // One array, so a distance sits at a constant offset from its entry
  // word and both stores share one address.
  __shared__ uint32_t s_stk[CLOSEST ? 2 : 1][DEPTH + 1][REND_TRACE_BLOCK];
  if (!RCP_IEEE) {
    const bool xeon = (flags & RK_TRACE_XEON) != 0u;
    for (uint32_t i = threadIdx.x; i < 4096; i += REND_TRACE_BLOCK)
      s_tab[i] = xeon ? (uint16_t)(c_rcp_xeon[i] >> 11) : c_rcp_host[i];
    __syncthreads();
  }

  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tid = threadIdx.x;
  // Stack entries [0, DEPTH) live in LDS, deeper ones in the per-thread
  // spill area (interleaved across threads so equal depths coalesce).
  uint2 *const myspill = spill + blockIdx.x * REND_TRACE_BLOCK + threadIdx.x;
  const size_t sstride = (size_t)gridDim.x * REND_TRACE_BLOCK;
  // LDS is always addressed through a clamped row (deep entries touch the
  // junk row) so LDS and spill accesses never share a generic pointer
  auto put = [&](uint32_t i, uint32_t o, float d) {
    const uint32_t row = i < (uint32_t)DEPTH ? i : (uint32_t)DEPTH;
    s_stk[0][row][tid] = o;
    if (CLOSEST) s_stk[CLOSEST ? 1 : 0][row][tid] = __float_as_uint(d);
    if (i >= (uint32_t)DEPTH) myspill[(i - DEPTH) * sstride] = make_uint2(o, __float_as_uint(d));
  };

  // Lane states.  A lane waits in its state until the wave runs that
  // state's step; every ray still makes exactly the reference's sequence of
  // visits, only the interleaving across lanes changes:
  //   NODE        ofs is an interior node of the current tree
  //   LEAF        ofs is a BLAS leaf
  //   SETUP_INST  ofs is a TLAS leaf: transform the world ray, enter the BLAS
  //   SETUP_WORLD the world ray's slab constants are needed (a TLAS node
  //               popped after a BLAS)
  //   SETUP_NEW   a new ray at the TLAS root: the refill loaded its record
  //               into the world-ray registers and its tmax into a register
  //               of its own (both unread by the node step, which so never
  //               waits for that load); the next setup step sets t and the
  //               slab constants
  enum : uint32_t { ST_IDLE = 0, ST_NODE = 1, ST_LEAF = 2, ST_SETUP_INST = 3,
                    ST_SETUP_WORLD = 4, ST_SETUP_NEW = 5 };
  // Leaf steps spread the waiting lanes' triangles over the wave (below) in
  // the host-table kernels.  The IEEE-reciprocal kernels (the path
  // tracer's) load each node's child words with its bounds; with the
  // exchange's registers on top of those the closest-hit one spills, and
  // without the early child load path-traced frames were no faster, so they
  // test each leaf in its own lane.
  constexpr bool SPREAD_LEAF = !RCP_IEEE;
  // leaf and setup steps run for at least this many waiting lanes, or once
  // fewer lanes than this are left at interior nodes
  const uint32_t kStepBatch = (flags >> 8) & 0xffu ? (flags >> 8) & 0xffu : RK_STEP_BATCH;
  uint32_t state = ST_IDLE;
  bool exhausted = false;  // wave-uniform
  // wave-uniform claim state: rays [lnext, lend) are claimed but unassigned
  uint32_t lnext = 0, lend = 0, tried = 0;
  const uint32_t myseg = blockIdx.x % RK_NSEG;
  uint32_t rid = 0;
  RayS cr;                    // ray of the current tree (world or instance)
  float wo[3] = {0, 0, 0}, wd[3] = {0, 0, 0};  // world ray
  float t = 0.0f, tu = 0.0f, tv = 0.0f, tnew = 0.0f;
  uint32_t hid = 0xffffffffu;
  uint32_t ofs = 0, sp = 0, blas_sp = 0, instid = 0, iter = 0;
  uint32_t base = (uint32_t)tlas;  // arena offset of the current tree (< 4 GiB)
  bool in_blas = false;
  bool occluded = false;
  uint64_t c_nodes = 0, c_leaves = 0, c_insts = 0, c_rays = 0;
  uint64_t c_witer = 0, c_lanes = 0, c_leafx = 0, c_setupx = 0;

  // next state for a lane at entry ofs of the current tree
  auto classify = [&](uint32_t o) -> uint32_t {
    return !(o & DNODE_LEAF) ? ST_NODE : (in_blas ? ST_LEAF : ST_SETUP_INST);
  };

  for (;;) {
    const uint64_t m_live = __ballot(state != ST_IDLE);
    if (COUNT && m_live) {
      c_witer++;
      c_lanes += (uint64_t)__popcll(m_live);
    }

    bool need_pop = false;
    bool done = false;
    uint32_t err = 0;  // error code of a lane that ends on an error

    // ---- node step -------------------------------------------------------
    const uint64_t m_node = __ballot(state == ST_NODE);
    if (m_node && state == ST_NODE) {
      if (++iter > max_iter) {
        err = RK_ERR_ITER;
        done = true;
      } else {
        float tmn[8];
        uint32_t ch[8], invw;
        constexpr bool HOIST_CH = RCP_IEEE;  // slab8 already consumes the child words
        // the path is chosen before the node is read, so each is one block
        const uint32_t hmr = __ballot(!(cr.fin && t == t) && state == ST_NODE) == 0
                                 ? slab8<true, RCP_IEEE>(arena, base + ofs, cr, t, tmn, ch, invw)
                                 : slab8<false, RCP_IEEE>(arena, base + ofs, cr, t, tmn, ch, invw);
        if (COUNT) c_nodes++;

        const uint32_t hitcnt = __builtin_popcount(hmr);
        if (hitcnt == 0) {
          need_pop = true;
        } else if (sp + hitcnt > REND_STACK_MAX) {
          // The reference asserts here (rend.c:834); report instead of
          // writing past the stack.
          err = RK_ERR_STACK;
          done = true;
        } else {
          // Highest rank among hits (nearest) is the next node, the others
          // are pushed farthest first.  The any-hit pass uses the same order:
          // its t_far never shrinks, so the set of leaves it can reach, and
          // with it the occlusion result, does not depend on the visiting
          // order (the reference pushes unordered, rend.c:1884-1886).
          uint32_t next = 0;
          const uint32_t keym = hmr;
          const uint32_t top = 31u - __builtin_clz(keym);
          if (__ballot(sp + 7u > (uint32_t)DEPTH && state == ST_NODE) == 0) {
            // no lane can reach the spill area: branch-free pushes,
            // children that are not kept go to the junk row
#pragma unroll
            for (int c = 0; c < 8; c++) {
              const uint32_t key = (invw >> (3 * c)) & 7u;
              const bool hit = (keym >> key) & 1u;
              const uint32_t pos = __builtin_popcount(keym & ((1u << key) - 1u));
              const uint32_t slot = (hit && key != top) ? sp + pos : (uint32_t)DEPTH;
              s_stk[0][slot][tid] = ch[c];
              if (CLOSEST) s_stk[CLOSEST ? 1 : 0][slot][tid] = __float_as_uint(tmn[c]);
              next = key == top ? ch[c] : next;
            }
          } else {
#pragma unroll
            for (int c = 0; c < 8; c++) {
              const uint32_t key = (invw >> (3 * c)) & 7u;
              if ((keym >> key) & 1u) {
                if (key == top) {
                  next = ch[c];
                } else {
                  const uint32_t pos = __builtin_popcount(keym & ((1u << key) - 1u));
                  put(sp + pos, ch[c], CLOSEST ? tmn[c] : 0.0f);
                }
              }
            }
          }
          sp += hitcnt - 1;
          ofs = next;
          state = classify(next);
        }
        // Waves whose lanes all missed never read the child words; left in
        // flight, their loads would make the next node step, which may reuse
        // those registers, first wait for every outstanding memory access
        // (this iteration's refill loads included).  Consumed here, they
        // arrive with the bounds and cost no wait of their own.
        if (!HOIST_CH)
          asm volatile("" ::"v"(ch[0]), "v"(ch[1]), "v"(ch[2]), "v"(ch[3]), "v"(ch[4]),
                       "v"(ch[5]), "v"(ch[6]), "v"(ch[7]));
      }
    }

    // ---- leaf step: run when enough lanes wait at leaves, or when few
    // lanes are left at interior nodes --------------------------------------
    const uint64_t m_leaf = __ballot(state == ST_LEAF && !need_pop && !done);
    // lanes that stay in the node phase: at an interior node, or about to
    // pop (most pops lead to another interior node)
    const uint64_t m_node2 = __ballot((state == ST_NODE || need_pop) && !done);
    const uint32_t n_node = (uint32_t)__popcll(m_node2);
    const bool leaf_step =
        m_leaf && ((uint32_t)__popcll(m_leaf) >= kStepBatch || n_node < kStepBatch);
    if (SPREAD_LEAF && leaf_step) {
      // The waiting lanes' leaves are spread over the whole wave, one
      // triangle per lane: the q-th waiting lane's four triangles go to lanes
      // 4q .. 4q+3 (16 leaves per pass), each lane runs the reference's
      // intersection arithmetic for its triangle with the owner's ray, and
      // a quad reduction picks the nearest hit with the reference's tie rule
      // (the highest-indexed of the equal minima).  Every lane of the wave
      // takes part in the exchanges, so this block is wave-uniform.
      if (COUNT) c_leafx++;
      bool part = state == ST_LEAF && !need_pop && !done;
      if (part && ++iter > max_iter) {
        err = RK_ERR_ITER;
        done = true;
        part = false;
      }
      if (COUNT && part) c_leaves++;
      // lane-derived values are recomputed here from an opaque copy of the
      // lane index: hoisted out of the loop they would hold registers
      // through the node step
      uint32_t lid = lane;
      asm volatile("" : "+v"(lid));
      const uint64_t mp = __ballot(part);
      const uint32_t nleaf = (uint32_t)__popcll(mp);
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(mp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mp, 0u));
      const uint32_t other = lid - rank;  // non-waiting lanes below this one
      // lane q of tab holds the lane id of the q-th waiting lane: waiting
      // lanes send their id to lane rank, the others to nleaf + other, so
      // every lane receives exactly one value
      const uint32_t tab = (uint32_t)__builtin_amdgcn_ds_permute(
          (int)((part ? rank : nleaf + other) << 2), (int)lid);
      const uint32_t lofs = base + (ofs & ~DNODE_LEAF);
      const uint32_t k = lid & 3u;
      for (uint32_t q0 = 0; q0 < nleaf; q0 += 16u) {
        const uint32_t q = q0 + (lid >> 2);
        const bool act = q < nleaf;
        const int src = (int)__shfl((int)tab, act ? (int)q : 0);
        const float ox = __shfl(cr.ox, src), oy = __shfl(cr.oy, src), oz = __shfl(cr.oz, src);
        const float dx = __shfl(cr.dx, src), dy = __shfl(cr.dy, src), dz = __shfl(cr.dz, src);
        const float tcur = __shfl(t, src);
        const uint32_t lo = (uint32_t)__shfl((int)lofs, src);
        float tb = FLT_MAXF, u = 0.0f, v = 0.0f;
        uint32_t id = 0;
        bool ok = false;
        if (act) {
          // this lane's column of the SoA leaf: the arena base plus a 32-bit
          // offset (scalar base, vector offset, constant field offsets)
          const float *lf = (const float *)(arena + (uint32_t)(lo + 4u * k));
          const float V0x = lf[0], V0y = lf[4], V0z = lf[8];
          const float E0x = lf[12], E0y = lf[16], E0z = lf[20];
          const float E1x = lf[24], E1y = lf[28], E1z = lf[32];
          if (CLOSEST) id = ((const uint32_t *)lf)[36];
          const float pvx = __builtin_fmaf(dy, E1z, -(dz * E1y));
          const float pvy = __builtin_fmaf(dz, E1x, -(dx * E1z));
          const float pvz = __builtin_fmaf(dx, E1y, -(dy * E1x));
          const float tvx = ox - V0x;
          const float tvy = oy - V0y;
          const float tvz = oz - V0z;
          const float det = __builtin_fmaf(E0x, pvx, __builtin_fmaf(E0y, pvy, E0z * pvz));
          const float qvx = __builtin_fmaf(tvy, E0z, -(tvz * E0y));
          const float qvy = __builtin_fmaf(tvz, E0x, -(tvx * E0z));
          const float qvz = __builtin_fmaf(tvx, E0y, -(tvy * E0x));
          const float idet = RCP_IEEE ? 1.0f / det : rcp_table(det, s_tab);
          u = idet * __builtin_fmaf(tvx, pvx, __builtin_fmaf(tvy, pvy, tvz * pvz));
          v = idet * __builtin_fmaf(dx, qvx, __builtin_fmaf(dy, qvy, dz * qvz));
          const float tk = idet * __builtin_fmaf(E1x, qvx, __builtin_fmaf(E1y, qvy, E1z * qvz));
          ok = (u >= 0.0f) & (v >= 0.0f) & ((u + v) <= 1.0f) & (tk > 0.0f) & (tk < tcur);
          tb = ok ? tk : FLT_MAXF;
        }
        const uint64_t okm = __ballot(ok);
        if (CLOSEST) {
          // nearest of the quad; on equal distances the higher triangle
          // index wins (the reference keeps the last of the equal minima)
          quad_nearest<0xB1>(k & 1u, tb, u, v, id);
          quad_nearest<0x4E>(k & 2u, tb, u, v, id);
        }
        // owners of this pass read their quad's result from its first lane
        const bool mine = part && rank >= q0 && rank < q0 + 16u;
        const uint32_t lead = mine ? (rank - q0) << 2 : 0u;
        const bool qhit = ((okm >> lead) & 0xfull) != 0ull;
        // (a lane owns a leaf in one pass only, so its t can change here)
        if (CLOSEST) {
          const float qt = __shfl(tb, (int)lead), qu = __shfl(u, (int)lead),
                      qv = __shfl(v, (int)lead);
          const uint32_t qid = (uint32_t)__shfl((int)id, (int)lead);
          if (mine) {
            if (qhit) {
              t = qt;
              tu = qu;
              tv = qv;
              hid = (qid << 12) | instid;
            }
            need_pop = true;
          }
        } else if (mine) {
          if (qhit) { done = true; occluded = true; }
          else need_pop = true;
        }
      }
    }

    if (!SPREAD_LEAF && leaf_step) {
      // each waiting lane tests its own leaf's four triangles
      if (COUNT) c_leafx++;
      if (state == ST_LEAF && !need_pop && !done) {
        if (++iter > max_iter) {
          err = RK_ERR_ITER;
          done = true;
        } else {
          const dleaf *lf = (const dleaf *)(arena + (base + (ofs & ~DNODE_LEAF)));
          float4 v0x = *(const float4 *)lf->v0x, v0y = *(const float4 *)lf->v0y,
                 v0z = *(const float4 *)lf->v0z;
          float4 e0x = *(const float4 *)lf->e0x, e0y = *(const float4 *)lf->e0y,
                 e0z = *(const float4 *)lf->e0z;
          float4 e1x = *(const float4 *)lf->e1x, e1y = *(const float4 *)lf->e1y,
                 e1z = *(const float4 *)lf->e1z;
          const float *V0x = (const float *)&v0x, *V0y = (const float *)&v0y,
                      *V0z = (const float *)&v0z;
          const float *E0x = (const float *)&e0x, *E0y = (const float *)&e0y,
                      *E0z = (const float *)&e0z;
          const float *E1x = (const float *)&e1x, *E1y = (const float *)&e1y,
                      *E1z = (const float *)&e1z;
          if (COUNT) c_leaves++;

          float tt[4], uu[4], vv[4];
          uint32_t mask = 0;
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const float pvx = __builtin_fmaf(cr.dy, E1z[k], -(cr.dz * E1y[k]));
            const float pvy = __builtin_fmaf(cr.dz, E1x[k], -(cr.dx * E1z[k]));
            const float pvz = __builtin_fmaf(cr.dx, E1y[k], -(cr.dy * E1x[k]));
            const float tvx = cr.ox - V0x[k];
            const float tvy = cr.oy - V0y[k];
            const float tvz = cr.oz - V0z[k];
            const float det = __builtin_fmaf(E0x[k], pvx,
                              __builtin_fmaf(E0y[k], pvy, E0z[k] * pvz));
            const float qvx = __builtin_fmaf(tvy, E0z[k], -(tvz * E0y[k]));
            const float qvy = __builtin_fmaf(tvz, E0x[k], -(tvx * E0z[k]));
            const float qvz = __builtin_fmaf(tvx, E0y[k], -(tvy * E0x[k]));
            const float idet = RCP_IEEE ? 1.0f / det : rcp_table(det, s_tab);
            const float u = idet * __builtin_fmaf(tvx, pvx, __builtin_fmaf(tvy, pvy, tvz * pvz));
            const float v = idet * __builtin_fmaf(cr.dx, qvx, __builtin_fmaf(cr.dy, qvy, cr.dz * qvz));
            const float tk = idet * __builtin_fmaf(E1x[k], qvx, __builtin_fmaf(E1y[k], qvy, E1z[k] * qvz));
            const bool ok = (u >= 0.0f) & (v >= 0.0f) & ((u + v) <= 1.0f) &
                            (tk > 0.0f) & (tk < t);
            tt[k] = tk; uu[k] = u; vv[k] = v;
            mask |= (ok ? 1u : 0u) << k;
          }
          if (!CLOSEST) {
            if (mask) { done = true; occluded = true; }
            else need_pop = true;
          } else {
            if (mask) {
              float tb[4];
#pragma unroll
              for (int k = 0; k < 4; k++) tb[k] = ((mask >> k) & 1u) ? tt[k] : FLT_MAXF;
              const float m = xmin(xmin(tb[0], tb[1]), xmin(tb[2], tb[3]));
              int sel = 0;
#pragma unroll
              for (int k = 0; k < 4; k++) if (tb[k] == m) sel = k;
              float su = uu[0], sv = vv[0], stt = tb[0];
#pragma unroll
              for (int k = 1; k < 4; k++)
                if (sel == k) { su = uu[k]; sv = vv[k]; stt = tb[k]; }
              t = stt;
              tu = su;
              tv = sv;
              hid = (lf->id[sel] << 12) | instid;
            }
            need_pop = true;
          }
        }
      }
    }

    // ---- setup step: instance entry (transform, then the BLAS ray's slab
    // constants) and world-ray constants (new rays, TLAS re-entry) ----------
    const bool in_setup = state == ST_SETUP_INST || state == ST_SETUP_WORLD ||
                          state == ST_SETUP_NEW;
    const uint64_t m_setup = __ballot(in_setup && !need_pop && !done);
    if (m_setup && ((uint32_t)__popcll(m_setup) >= kStepBatch || n_node < kStepBatch)) {
      if (COUNT) c_setupx++;
      if (in_setup && !need_pop && !done) {
        // One slab set-up (three IEEE divisions) serves both kinds: the
        // world ray as it is, or the world ray in the instance's space.
        float ox = wo[0], oy = wo[1], oz = wo[2], dx = wd[0], dy = wd[1], dz = wd[2];
        if (state == ST_SETUP_NEW) {
          t = RCP_IEEE ? tu : tnew;
          tu = 0.0f;
        }
        if (state == ST_SETUP_INST) {
          if (++iter > max_iter) {
            err = RK_ERR_ITER;
            done = true;
          } else {
            // TLAS leaf: enter the instance's BLAS with the transformed ray.
            instid = ofs & ~DNODE_LEAF;
            const dinst *in = insts + instid;
            const float4 m0 = *(const float4 *)&in->inv[0];
            const float4 m1 = *(const float4 *)&in->inv[4];
            const float4 m2 = *(const float4 *)&in->inv[8];
            const uint64_t blas = in->blas;
            // wait for the BLAS offset here, beside the matrix rows: left
            // outstanding, the next node step (which needs it for its
            // address) would have to drain the whole vector-memory counter,
            // hit stores and refill loads included
            asm volatile("" ::"v"(blas));
            if (COUNT) c_insts++;
            ox = m0.x * wo[0] + m0.y * wo[1] + m0.z * wo[2] + m0.w;
            oy = m1.x * wo[0] + m1.y * wo[1] + m1.z * wo[2] + m1.w;
            oz = m2.x * wo[0] + m2.y * wo[1] + m2.z * wo[2] + m2.w;
            dx = m0.x * wd[0] + m0.y * wd[1] + m0.z * wd[2];
            dy = m1.x * wd[0] + m1.y * wd[1] + m1.z * wd[2];
            dz = m2.x * wd[0] + m2.y * wd[1] + m2.z * wd[2];
            if (flags & RK_TRACE_NUDGE) {
              dx = nudge0(dx);
              dy = nudge0(dy);
              dz = nudge0(dz);
            }
            in_blas = true;
            blas_sp = sp;
            base = (uint32_t)blas;
            ofs = 0;
          }
        }
        if (!done) {
          ray_setup(cr, ox, oy, oz, dx, dy, dz);
          state = ST_NODE;
        }
      }
    }

    // ---- pops: after a node without hits or a leaf -------------------------
    if (need_pop && !done) {
      bool world = false;
      for (;;) {
        if (in_blas && sp == blas_sp) {
          // back in the TLAS: the world ray's constants are needed again
          // before the next TLAS node (not for an instance entry or the end)
          in_blas = false;
          base = (uint32_t)tlas;
          world = true;
        }
        if (sp == 0) { done = true; break; }
        --sp;
        const uint32_t row = sp < (uint32_t)DEPTH ? sp : (uint32_t)DEPTH;
        uint2 e = make_uint2(s_stk[0][row][tid], CLOSEST ? s_stk[CLOSEST ? 1 : 0][row][tid] : 0u);
        if (sp >= (uint32_t)DEPTH) {
          // volatile keeps the compiler from merging this load with the LDS
          // one; the global address space keeps it a global (not flat)
          // load: a flat load counts against both the vector-memory and
          // the LDS counters, and while one may be outstanding every later
          // wait in the loop has to drain both counters completely
          const volatile gu32 *q = (const volatile gu32 *)&myspill[(sp - DEPTH) * sstride];
          e.x = q[0];
          e.y = q[1];
        }
        if (!CLOSEST || __uint_as_float(e.y) <= t) { ofs = e.x; break; }
      }
      if (!done) {
        const uint32_t nst = classify(ofs);
        state = (world && nst == ST_NODE) ? ST_SETUP_WORLD : nst;
      }
    }

    if (done) {
      if (err) report_error(diag, err, rays, rid, iter);
      if (CLOSEST) {
        rhit h; h.t = t; h.u = tu; h.v = tv; h.id = hid;
        *(float4 *)&hits[rid] = *(const float4 *)&h;
      } else {
        occ[rid] = occluded ? 1 : 0;
      }
      state = ST_IDLE;
    }

    // ---- refill idle lanes from the wave's claimed chunk ------------------
    // At the end of the loop body: a new ray's record is loaded here and its
    // slab constants computed by the next iteration's SETUP_NEW step, after
    // the node step, so the load's latency overlaps the other lanes' work.
    const uint64_t idle = __ballot(state == ST_IDLE);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle >= refill) {
      // claim a new chunk when the current one is used up (one atomic per
      // RK_CHUNK rays), first from this XCD's segment, then from the others
      while (lnext == lend && !exhausted) {
        const uint32_t seg = (myseg + tried) % RK_NSEG;
        const uint32_t lo = (uint32_t)(((uint64_t)n * seg) / RK_NSEG);
        const uint32_t len = (uint32_t)(((uint64_t)n * (seg + 1)) / RK_NSEG) - lo;
        uint32_t got = 0;
        if (lane == 0) got = atomicAdd(&ctl[seg * RK_SEG_STRIDE], RK_CHUNK);
        got = __shfl(got, 0);
        if (got < len) {
          lnext = lo + got;
          lend = lo + min(got + RK_CHUNK, len);
        } else if (++tried == RK_NSEG) {
          exhausted = true;
        }
      }
      const uint32_t take = min(nidle, lend - lnext);
      if (state == ST_IDLE) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        const uint32_t r = lnext + below;
        if (below < take) {
          rid = r;
          // Origin, direction and tmax should land in the registers the
          // loop carries them in.  Read as one 16-byte load with the origin,
          // tmax is copied out of it, and that copy waits for the record
          // right here; a load of its own (an opaque offset keeps the loads
          // apart) avoids the wait.  The IEEE-reciprocal kernels, at their
          // register limit, would spill for it and keep the single load.
          const float *rr = (const float *)&rays[r];
          uint32_t o3 = 3u;
          if (!RCP_IEEE) asm volatile("" : "+v"(o3));
          wo[0] = rr[0]; wo[1] = rr[1]; wo[2] = rr[2];
          wd[0] = rr[4]; wd[1] = rr[5]; wd[2] = rr[6];
          // tmax, moved to t by the SETUP_NEW step; its own register keeps
          // the hit store (t, u, v, id) from waiting for this load
          if (RCP_IEEE) tu = rr[o3];
          else tnew = rr[o3];
          tv = 0.0f;
          hid = 0xffffffffu;
          ofs = 0; sp = 0; blas_sp = 0; iter = 0;
          base = (uint32_t)tlas;
          in_blas = false;
          occluded = false;
          state = ST_SETUP_NEW;
          if (COUNT) c_rays++;
        }
      }
      lnext += take;
    }
    if (__ballot(state != ST_IDLE) == 0 && exhausted) break;
  }

  if (COUNT) {
    for (int o = 32; o > 0; o >>= 1) {
      c_nodes += __shfl_down(c_nodes, o);
      c_leaves += __shfl_down(c_leaves, o);
      c_insts += __shfl_down(c_insts, o);
      c_rays += __shfl_down(c_rays, o);
    }
    if (lane == 0) {
      atomicAdd(&stats[0], (unsigned long long)c_nodes);
      atomicAdd(&stats[1], (unsigned long long)c_leaves);
      atomicAdd(&stats[2], (unsigned long long)c_insts);
      atomicAdd(&stats[3], (unsigned long long)c_rays);
      // wave-level figures are the same in every lane: lane 0's alone
      atomicAdd(&stats[4], (unsigned long long)c_witer);
      atomicAdd(&stats[5], (unsigned long long)c_lanes);
      atomicAdd(&stats[6], (unsigned long long)c_leafx);
      atomicAdd(&stats[7], (unsigned long long)c_setupx);
    }
  }
}

}  // namespace

namespace {
__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

struct LaunchArgs {
  const uint8_t *arena;
  uint64_t tlas;
  const dinst *insts;
  const rray *rays;
  rhit *hits;
  uint8_t *occ;
  uint32_t n;
  uint32_t *ctl;
  uint2 *spill;
  unsigned long long *stats;
  uint32_t refill, max_iter;
  const uint32_t *n_dev;
  uint32_t *diag;
  uint32_t flags;
};

template <bool C, bool N, bool I, int D, int W>
hipError_t launch_one(uint32_t grid, hipStream_t s, const LaunchArgs &a) {
  hipLaunchKernelGGL((k_trace<C, N, I, D, W>), dim3(grid), dim3(REND_TRACE_BLOCK), 0, s,
                     a.arena, a.tlas, a.insts, a.rays, a.hits, a.occ, a.n, a.ctl,
                     a.spill, a.stats, a.refill, a.max_iter, a.n_dev, a.diag, a.flags);
  return hipGetLastError();
}

template <bool C, bool N, bool I, int D, int W>
int occ_one() {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (k_trace<C, N, I, D, W>),
                                                    REND_TRACE_BLOCK, 0) != hipSuccess)
    return -1;
  return b;
}

// Occupancy variants (LDS stack depth, waves per SIMD); LDS per block is
// 8 KB + (depth + 1) * 2 KB, so 160 KB of LDS holds 4, 5, 6 and 8 blocks.
// The counting kernel exists for variant 0 only.
#define RK_VARIANTS(X) X(0, 15, 4) X(1, 11, 5) X(2, 8, 6) X(3, 5, 8)

template <bool C, bool I>
hipError_t launch_v(int v, uint32_t grid, hipStream_t s, const LaunchArgs &a) {
#define RK_L(idx, d, w) case idx: return launch_one<C, false, I, d, w>(grid, s, a);
  switch (v) { RK_VARIANTS(RK_L) default: break; }
#undef RK_L
  return hipErrorInvalidValue;
}

template <bool C, bool I>
int occ_v(int v) {
#define RK_O(idx, d, w) case idx: return occ_one<C, false, I, d, w>();
  switch (v) { RK_VARIANTS(RK_O) default: break; }
#undef RK_O
  return -1;
}
}  // namespace

// C-ABI between the host library and the traversal kernels.
extern "C" int rk_trace_variants(void) { return RK_TRACE_VARIANTS; }

// This host's RCPSS table (12-bit entries, index = 12 leading mantissa bits).
extern "C" int rk_trace_set_host_rcp(const uint16_t *tab) {
  return (int)hipMemcpyToSymbol(HIP_SYMBOL(c_rcp_host), tab, 4096 * sizeof(uint16_t), 0,
                                hipMemcpyHostToDevice);
}

extern "C" int rk_trace_occupancy(int closest, int count, int rcp_ieee, int variant) {
  if (count) {
    const int k = (closest ? 2 : 0) | (rcp_ieee ? 1 : 0);
    switch (k) {
      case 0: return occ_one<false, true, false, RK_DEPTH0, 4>();
      case 1: return occ_one<false, true, true, RK_DEPTH0, 4>();
      case 2: return occ_one<true, true, false, RK_DEPTH0, 4>();
      default: return occ_one<true, true, true, RK_DEPTH0, 4>();
    }
  }
  const int k = (closest ? 2 : 0) | (rcp_ieee ? 1 : 0);
  switch (k) {
    case 0: return occ_v<false, false>(variant);
    case 1: return occ_v<false, true>(variant);
    case 2: return occ_v<true, false>(variant);
    default: return occ_v<true, true>(variant);
  }
}

// LDS stack depth of a variant (the spill area holds the remaining entries).
extern "C" int rk_trace_lds_depth(int variant) {
#define RK_D(idx, d, w) case idx: return d;
  switch (variant) { RK_VARIANTS(RK_D) default: break; }
#undef RK_D
  return -1;
}

// Zeroes the claim counters with a kernel (safe inside captured graphs), then
// launches the persistent traversal grid.  Returns a hipError_t value.
extern "C" int rk_trace_launch(const void *arena, unsigned long long tlas,
                               const void *insts, const void *rays, void *hits,
                               void *occ, unsigned int n, unsigned int *ctl,
                               void *spill, unsigned long long *stats,
                               int closest, int count, int rcp_ieee, int variant,
                               unsigned int grid, unsigned int refill,
                               unsigned int max_iter, const unsigned int *n_dev,
                               unsigned int *diag, unsigned int flags,
                               void *stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u32, dim3(4), dim3(64), 0, s, ctl, (uint32_t)RK_CTL_ERR);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return (int)e;
  if (n == 0 && !n_dev) return 0;
  LaunchArgs a;
  a.arena = (const uint8_t *)arena;
  a.tlas = tlas;
  a.insts = (const dinst *)insts;
  a.rays = (const rray *)rays;
  a.hits = (rhit *)hits;
  a.occ = (uint8_t *)occ;
  a.n = n;
  a.ctl = ctl;
  a.spill = (uint2 *)spill;
  a.stats = stats;
  a.refill = refill;
  a.max_iter = max_iter;
  a.n_dev = n_dev;
  a.diag = diag;
  a.flags = flags;
  const int k = (closest ? 2 : 0) | (rcp_ieee ? 1 : 0);
  if (count) {
    switch (k) {
      case 0: e = launch_one<false, true, false, RK_DEPTH0, 4>(grid, s, a); break;
      case 1: e = launch_one<false, true, true, RK_DEPTH0, 4>(grid, s, a); break;
      case 2: e = launch_one<true, true, false, RK_DEPTH0, 4>(grid, s, a); break;
      default: e = launch_one<true, true, true, RK_DEPTH0, 4>(grid, s, a); break;
    }
    return (int)e;
  }
  switch (k) {
    case 0: e = launch_v<false, false>(variant, grid, s, a); break;
    case 1: e = launch_v<false, true>(variant, grid, s, a); break;
    case 2: e = launch_v<true, false>(variant, grid, s, a); break;
    default: e = launch_v<true, true>(variant, grid, s, a); break;
  }
  return (int)e;
}

// ---- allocation-free launches ---------------------------------------------
//
// The spill area of rend_hip_trace_noalloc is this file-scope array, not a
// library allocation.  Host code never passes its name to a kernel: the
// device address comes from hipGetSymbolAddress, taken once per device at
// library set-up (rk_scratch_init) and kept per device.  Every call on a
// device shares its array; calls that overlap on one device race on it.

constexpr size_t kScratchEntries =
    (size_t)RK_NOALLOC_BLOCKS * REND_TRACE_BLOCK * (REND_STACK_MAX - RK_NOALLOC_DEPTH);
static __device__ __align__(256) uint2 scratch[kScratchEntries];

constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices];

// Called at library set-up with the device current.  Returns a hipError_t.
extern "C" int rk_scratch_init(void) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  if (g_scratch[dev]) return 0;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return (int)e;
  g_scratch[dev] = p;
  return 0;
}

extern "C" unsigned long long rk_scratch_bytes(void) { return sizeof(scratch); }

// Closest-hit launch of the 16-entry LDS stack variant with the static spill
// area; grid = min(occupancy x CUs, RK_NOALLOC_BLOCKS).  Allocates nothing.
extern "C" int rk_trace_launch_noalloc(const void *arena, unsigned long long tlas,
                                       const void *insts, const void *rays, void *hits,
                                       unsigned int n, unsigned int *ctl, int rcp_ieee,
                                       unsigned int cus, unsigned int refill,
                                       unsigned int max_iter, unsigned int *diag,
                                       unsigned int flags, void *stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev]) return (int)hipErrorNotInitialized;
  const int occ = rcp_ieee ? occ_one<true, false, true, RK_NOALLOC_DEPTH, 4>()
                           : occ_one<true, false, false, RK_NOALLOC_DEPTH, 4>();
  if (occ < 1) return (int)hipErrorInvalidConfiguration;
  uint32_t grid = (uint32_t)occ * cus;
  if (grid > RK_NOALLOC_BLOCKS) grid = RK_NOALLOC_BLOCKS;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u32, dim3(4), dim3(64), 0, s, ctl, (uint32_t)RK_CTL_ERR);
  e = hipGetLastError();
  if (e != hipSuccess || n == 0) return (int)e;
  LaunchArgs a;
  a.arena = (const uint8_t *)arena;
  a.tlas = tlas;
  a.insts = (const dinst *)insts;
  a.rays = (const rray *)rays;
  a.hits = (rhit *)hits;
  a.occ = nullptr;
  a.n = n;
  a.ctl = ctl;
  a.spill = (uint2 *)g_scratch[dev];
  a.stats = nullptr;
  a.refill = refill;
  a.max_iter = max_iter;
  a.n_dev = nullptr;
  a.diag = diag;
  a.flags = flags & ~RK_TRACE_NUDGE;
  e = rcp_ieee ? launch_one<true, false, true, RK_NOALLOC_DEPTH, 4>(grid, s, a)
               : launch_one<true, false, false, RK_NOALLOC_DEPTH, 4>(grid, s, a);
  return (int)e;
}
