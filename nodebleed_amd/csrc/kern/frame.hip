// Frame kernels behind rend_render: camera ray generation, the reference's
// primary-ray shading and progressive accumulation.
//
//   k_camrays   make_camray_pckt (rend.c:3293-3402): jittered pinhole or
//               thin-lens rays; directions normalised exactly (the reference
//               uses the approximate RSQRTPS) and jitter from a counter-based
//               hash instead of xorshift128+ packets, so images agree with the
//               reference statistically, not bit for bit.
//   k_camrays_ref  the reference's own primary rays, bit for bit
//               (xorshift128+ packets, host RSQRTPS table, for a thin lens
//               the host libmvec's cos/sin of the lens angle from a table):
//               what rend_render traces.
//   k_shade_normals, k_shade_normals_ref  trace_rays/calc_nrm
//               (rend.c:3404-3440, 3109-3116): colour (n * 0.5 + 0.5) *
//               albedo, background on a miss, for pixel-major rays or for
//               rays in the reference's block/packet order.
//   accumulate  accum_pckt (rend.c:3565-3580): acc += colour, ARGB8 from
//               acc / (spp + samples) with the reference's 8-bit wrap.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../dlayout.h"
#include "rend_rays.h"

#include "rng.h"

struct RNrm {  // rend.h:73-80
  float n0[3];
  uint32_t mtlid;
  float n1[3];
  uint32_t pad0;
  float n2[3];
  uint32_t pad1;
};

struct RMtl {  // rend.h:54-62
  float col[3];
  float metallic, roughness, ior;
  uint32_t flags, pad0;
};

namespace {

// A direction component of exactly +-0 makes 1/d and o/d infinite; under the
// reference's slab rules (x86 MAXPS/MINPS, NaN operand wins) the NaN slab
// distances let every box test pass and the ray visits the whole tree.
// Generated directions are nudged to +-1e-20 instead (measure zero).
__device__ __forceinline__ float nonzero_dir(float x) {
  return x == 0.0f ? __builtin_copysignf(1e-20f, x) : x;
}


__device__ __forceinline__ void camray(const CamArgs &c, float px, float py,
                                       float u0, float u1, float u2, float u3,
                                       rray &r) {
  float sx = px + (u0 - 0.5f);
  float sy = py + (u1 - 0.5f);
  sx = __builtin_fmaf(sx, c.rw, -0.5f);
  sy = __builtin_fmaf(sy, c.rh, -0.5f);
  sx = sx * c.fovfdist;
  sy = sy * c.fovfdist;
  sx = sx * c.aspect;
  float ox = c.eye[0], oy = c.eye[1], oz = c.eye[2];
  float tx, ty, tz;
  const float rx0 = __builtin_fmaf(c.ri[0], sx, -(c.up[0] * sy));
  const float ry0 = __builtin_fmaf(c.ri[1], sx, -(c.up[1] * sy));
  const float rz0 = __builtin_fmaf(c.ri[2], sx, -(c.up[2] * sy));
  if (c.dof) {
    const float rad = sqrtf(u2);
    const float th = 6.28318531f * u3;
    float s, co;
    sincosf(th, &s, &co);
    const float lx = co * rad, ly = s * rad;
    ox = __builtin_fmaf(__builtin_fmaf(c.ri[0], lx, c.up[0] * ly), c.focrad, c.eye[0]);
    oy = __builtin_fmaf(__builtin_fmaf(c.ri[1], lx, c.up[1] * ly), c.focrad, c.eye[1]);
    oz = __builtin_fmaf(__builtin_fmaf(c.ri[2], lx, c.up[2] * ly), c.focrad, c.eye[2]);
    tx = (__builtin_fmaf(c.fwd[0], c.fdist, c.eye[0]) + rx0) - ox;
    ty = (__builtin_fmaf(c.fwd[1], c.fdist, c.eye[1]) + ry0) - oy;
    tz = (__builtin_fmaf(c.fwd[2], c.fdist, c.eye[2]) + rz0) - oz;
  } else {
    tx = __builtin_fmaf(c.fwd[0], c.fdist, rx0);
    ty = __builtin_fmaf(c.fwd[1], c.fdist, ry0);
    tz = __builtin_fmaf(c.fwd[2], c.fdist, rz0);
  }
  const float il = 1.0f / sqrtf(__builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz)));
  r.ox = ox; r.oy = oy; r.oz = oz; r.tmax = 3.402823466e+38f;
  r.dx = tx * il; r.dy = ty * il; r.dz = tz * il; r.pad = 0.0f;
}

// Rays for pixels of a rw x rh region, spp consecutive rays per pixel.
__global__ void k_camrays(CamArgs c, uint32_t rw, uint32_t rh, uint32_t spp,
                          uint32_t seed, uint32_t frame, rray *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = rw * rh * spp;
  if (i >= n) return;
  const uint32_t p = i / spp;
  const uint32_t x = p % rw, y = p / rw;
  const uint32_t key = rng_key(seed, frame, i);
  rray r;
  camray(c, (float)x, (float)y, rng_u01(key, 0), rng_u01(key, 1),
         rng_u01(key, 2), rng_u01(key, 3), r);
  float4 *o = (float4 *)&out[i];
  o[0] = make_float4(r.ox, r.oy, r.oz, r.tmax);
  o[1] = make_float4(nonzero_dir(r.dx), nonzero_dir(r.dy), nonzero_dir(r.dz), r.pad);
}

__device__ __forceinline__ uint32_t pack8(float a) {
  return (uint32_t)(long long)(255.0f * a) & 0xffu;
}

// The reference's normal tint of one hit (trace_rays/calc_nrm,
// rend.c:3419-3447, 3109-3116): (unit(M^-T n) + 1) * 0.5 * albedo, the
// background on a miss; same operations, same order, no fusion.
__device__ __forceinline__ void shade_hit(const rhit &h, const dinst *__restrict__ insts,
                                          const RNrm *__restrict__ nrms,
                                          const RMtl *__restrict__ mtls, float bgr,
                                          float bgg, float bgb, float &r, float &g,
                                          float &b) {
  r = bgr; g = bgg; b = bgb;
  if (h.t < 3.402823466e+38f) {
    const uint32_t inst = h.id & 0xfffu, tri = h.id >> 12;
    const dinst &in = insts[inst];
    const RNrm &rn = nrms[in.triofs + tri];
    const float u = h.u, v = h.v, w = 1.0f - u - v;
    float nx = rn.n1[0] * u + (rn.n2[0] * v + rn.n0[0] * w);
    float ny = rn.n1[1] * u + (rn.n2[1] * v + rn.n0[1] * w);
    float nz = rn.n1[2] * u + (rn.n2[2] * v + rn.n0[2] * w);
    // inverse transpose of the 3x3 part of globinv (rend.c:3431-3436)
    const float *m = in.inv;
    const float tx = m[0] * nx + m[4] * ny + m[8] * nz;
    const float ty = m[1] * nx + m[5] * ny + m[9] * nz;
    const float tz = m[2] * nx + m[6] * ny + m[10] * nz;
    const float il = 1.0f / sqrtf(tx * tx + ty * ty + tz * tz);
    nx = tx * il; ny = ty * il; nz = tz * il;
    const RMtl &mt = mtls[rn.mtlid];
    r = (nx + 1.0f) * 0.5f * mt.col[0];
    g = (ny + 1.0f) * 0.5f * mt.col[1];
    b = (nz + 1.0f) * 0.5f * mt.col[2];
  }
}

// accum_pckt (rend.c:3565-3580) for one pixel.
__device__ __forceinline__ void accum_px(float *__restrict__ acc, uint32_t *__restrict__ buf,
                                         size_t o, float cr, float cg, float cb,
                                         float invsamples) {
  const float ar = acc[3 * o + 0] + cr;
  const float ag = acc[3 * o + 1] + cg;
  const float ab = acc[3 * o + 2] + cb;
  acc[3 * o + 0] = ar;
  acc[3 * o + 1] = ag;
  acc[3 * o + 2] = ab;
  buf[o] = 0xff000000u | pack8(ar * invsamples) << 16 |
           pack8(ag * invsamples) << 8 | pack8(ab * invsamples);
}

// Shade spp hits per pixel with the reference's normal tint and accumulate.
__global__ void k_shade_normals(const rhit *__restrict__ hits,
                                const dinst *__restrict__ insts,
                                const RNrm *__restrict__ nrms,
                                const RMtl *__restrict__ mtls, float bgr,
                                float bgg, float bgb, uint32_t rw, uint32_t rh,
                                uint32_t width, uint32_t spp, float invsamples,
                                float *__restrict__ acc, uint32_t *__restrict__ buf) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= rw * rh) return;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f;
  for (uint32_t s = 0; s < spp; s++) {
    float r, g, b;
    shade_hit(hits[(size_t)p * spp + s], insts, nrms, mtls, bgr, bgg, bgb, r, g, b);
    cr += r; cg += g; cb += b;
  }
  const uint32_t x = p % rw, y = p / rw;
  accum_px(acc, buf, (size_t)y * width + x, cr, cg, cb, invsamples);
}

struct RTri {  // rend.h:64-71
  float v0[3], p0, v1[3], p1, v2[3], p2;
};

// One cosine-weighted diffuse ray per primary hit, about the geometric
// normal turned towards the incoming ray, origin p + 1e-3 * dir (EPS2,
// util.h:8; rend.c:3281).  Basis: branchless ONB of Duff et al. 2017 as in
// create_onb (rend.c:3126-3133); direction: rand_hemicos (rend.c:3135-3140).
// A missed primary yields a ray with tmax = -1, which misses at the root.
__global__ void k_diffuse_rays(const rray *__restrict__ prim,
                               const rhit *__restrict__ hits,
                               const dinst *__restrict__ insts,
                               const RTri *__restrict__ tris, uint32_t n,
                               uint32_t seed, rray *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rray r = prim[i];
  const rhit h = hits[i];
  rray o;
  o.pad = 0.0f;
  if (h.id == 0xffffffffu) {
    o.ox = r.ox; o.oy = r.oy; o.oz = r.oz; o.tmax = -1.0f;
    o.dx = r.dx; o.dy = r.dy; o.dz = r.dz;
  } else {
    const dinst &in = insts[h.id & 0xfffu];
    const RTri &t = tris[in.triofs + (h.id >> 12)];
    const float e0x = t.v1[0] - t.v0[0], e0y = t.v1[1] - t.v0[1], e0z = t.v1[2] - t.v0[2];
    const float e1x = t.v2[0] - t.v0[0], e1y = t.v2[1] - t.v0[1], e1z = t.v2[2] - t.v0[2];
    float nx = e0y * e1z - e0z * e1y, ny = e0z * e1x - e0x * e1z, nz = e0x * e1y - e0y * e1x;
    const float *m = in.inv;  // normal to world: transpose of the inverse
    float wx = m[0] * nx + m[4] * ny + m[8] * nz;
    float wy = m[1] * nx + m[5] * ny + m[9] * nz;
    float wz = m[2] * nx + m[6] * ny + m[10] * nz;
    const float il = 1.0f / sqrtf(wx * wx + wy * wy + wz * wz);
    wx *= il; wy *= il; wz *= il;
    if (wx * r.dx + wy * r.dy + wz * r.dz > 0.0f) { wx = -wx; wy = -wy; wz = -wz; }
    const uint32_t key = rng_key(seed, 0x5bd1e995u, i);
    const float u0 = rng_u01(key, 0), u1 = rng_u01(key, 1);
    const float sgn = copysignf(1.0f, wz);
    const float a = -1.0f / (sgn + wz);
    const float b = wx * wy * a;
    const float b1x = 1.0f + sgn * wx * wx * a, b1y = sgn * b, b1z = -sgn * wx;
    const float b2x = b, b2y = sgn + wy * wy * a, b2z = -wy;
    float s, c;
    sincosf(6.28318531f * u0, &s, &c);
    const float su = sqrtf(u1), lx = c * su, ly = s * su, lz = sqrtf(1.0f - u1);
    const float dx = b1x * lx + b2x * ly + wx * lz;
    const float dy = b1y * lx + b2y * ly + wy * lz;
    const float dz = b1z * lx + b2z * ly + wz * lz;
    const float px = r.ox + r.dx * h.t, py = r.oy + r.dy * h.t, pz = r.oz + r.dz * h.t;
    o.ox = px + dx * 1e-3f; o.oy = py + dy * 1e-3f; o.oz = pz + dz * 1e-3f;
    o.tmax = 3.402823466e+38f;
    o.dx = dx; o.dy = dy; o.dz = dz;
  }
  float4 *po = (float4 *)&out[i];
  po[0] = make_float4(o.ox, o.oy, o.oz, o.tmax);
  po[1] = make_float4(nonzero_dir(o.dx), nonzero_dir(o.dy), nonzero_dir(o.dz), o.pad);
}

// Plain streaming copy (16 bytes per lane per step): the reference point
// for converting hardware fetch/write counters into HBM bytes.
__global__ void k_copy_f4(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// ---- reference-exact primary rays -----------------------------------------
//
// rend_render (rend.c:3603-3708) walks blocks of bszx x bszy pixels; block
// blk seeds a 4-lane xorshift128+ generator with rand8_init((blk + 13) *
// 1571, samples * 23) (util.c:63-75, lanes 1..3 by the generator's jump),
// and its packets of 4 x 2 pixels, in Morton order (rend.c:155-162), each
// draw two 8-float vectors (randf8, util.c:99-107) for the pixel jitter of
// make_camray_pckt (rend.c:3293-3402).  The pinhole branch is restated with
// the same fused and separate roundings and the host RSQRTPS table, so every
// ray is bit-identical.  One thread per packet replays its block's
// generator 2 k steps ahead.

// RSQRTPS results for the 12 leading mantissa bits at biased exponents 128
// (index 0..4095) and 127 (4096..8191): this host's or the build host's
// capture, as the library uploads them.
__constant__ uint32_t c_rsqrt_tab[8192];

struct Xs128 {
  uint64_t s0, s1;
};

// xorshift128+ with shifts 23, 18, 5 (util.c:29-41): returns s0 + s1 and
// advances the state.
__device__ __forceinline__ uint64_t xs128_next(Xs128 &x) {
  uint64_t a = x.s0;
  const uint64_t b = x.s1;
  const uint64_t res = a + b;
  x.s0 = b;
  a ^= a << 23;
  x.s1 = a ^ b ^ (a >> 18) ^ (b >> 5);
  return res;
}

// The generator's jump polynomial (2^64 steps ahead), as util.c:43-61.
__device__ __forceinline__ Xs128 xs128_jump(Xs128 in) {
  const uint64_t J[2] = {0x8a5cd789635d2dffull, 0x121fd2155c472f96ull};
  Xs128 out = {0, 0};
  for (int i = 0; i < 2; i++)
    for (int b = 0; b < 64; b++) {
      if (J[i] & (1ull << b)) {
        out.s0 ^= in.s0;
        out.s1 ^= in.s1;
      }
      xs128_next(in);
    }
  return out;
}

// x86 RSQRTPS from the captured tables (rule proven over all inputs by
// gen_rsqrt_table.c on Intel and AMD cores).
__device__ __forceinline__ float rsqrt_host(float x) {
  const uint32_t u = __float_as_uint(x);
  const uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
  uint32_t r;
  if (e == 255u) r = m ? (u | 0x00400000u) : (s ? 0xffc00000u : 0u);
  else if (e == 0u) r = s | 0x7f800000u;
  else if (s) r = 0xffc00000u;
  else {
    const uint32_t p = e & 1u;
    const int sh = ((int)e - (p ? 127 : 128)) / 2;
    r = (uint32_t)((int32_t)c_rsqrt_tab[(p << 12) | (m >> 11)] - sh * (1 << 23));
  }
  return __uint_as_float(r);
}

__device__ __forceinline__ uint32_t mort_dec(uint32_t v) {
  v &= 0x55555555u;
  v = (v ^ (v >> 1)) & 0x33333333u;
  v = (v ^ (v >> 2)) & 0x0f0f0f0fu;
  v = (v ^ (v >> 4)) & 0x00ff00ffu;
  return (v ^ (v >> 8)) & 0x0000ffffu;
}

__global__ void k_camrays_ref(CamArgs c, uint32_t bszx, uint32_t bszy, uint32_t blkcntx,
                              uint32_t nblk, uint32_t npk, uint32_t samples,
                              const float2 *__restrict__ lens, rray *__restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * npk) return;
  const uint32_t blk = t / npk, k = t % npk;
  // draws per packet: u0, u1 (pixel jitter), and u2, u3 (lens) for a thin lens
  const int nd = c.dof ? 4 : 2;
  Xs128 g[4];
  g[0].s0 = (uint64_t)(int64_t)(((int)blk + 13) * 1571);
  g[0].s1 = (uint64_t)(samples * 23u);
  for (int l = 1; l < 4; l++) g[l] = xs128_jump(g[l - 1]);
  for (uint32_t i = 0; i < (uint32_t)nd * k; i++)
    for (int l = 0; l < 4; l++) xs128_next(g[l]);
  float u[4][8];
  // mantissa bits of u3, the index of the lens angle 2 pi u3 in the table
  uint32_t m3[8];
  for (int d = 0; d < nd; d++)
    for (int l = 0; l < 4; l++) {
      const uint64_t v = xs128_next(g[l]);
      const uint32_t lo = (uint32_t)v >> 9, hi = (uint32_t)(v >> 32) >> 9;
      u[d][2 * l] = __uint_as_float(lo | 0x3f800000u) - 1.0f;
      u[d][2 * l + 1] = __uint_as_float(hi | 0x3f800000u) - 1.0f;
      if (d == 3) { m3[2 * l] = lo; m3[2 * l + 1] = hi; }
    }
  const uint32_t x = (blk % blkcntx) * bszx + mort_dec(k) * 4u;
  const uint32_t y = (blk / blkcntx) * bszy + mort_dec(k >> 1) * 2u;
  for (int l = 0; l < 8; l++) {
    float sx = (float)x + (float)(l & 3);
    float sy = (float)y + (float)(l >> 2);
    sx = sx + (u[0][l] - 0.5f);
    sy = sy + (u[1][l] - 0.5f);
    sx = __builtin_fmaf(sx, c.rw, -0.5f);
    sy = __builtin_fmaf(sy, c.rh, -0.5f);
    sx = sx * c.fovfdist;
    sy = sy * c.fovfdist;
    sx = sx * c.aspect;
    const float ux = c.up[0] * sy, uy = c.up[1] * sy, uz = c.up[2] * sy;
    const float rx = __builtin_fmaf(c.ri[0], sx, -ux);
    const float ry = __builtin_fmaf(c.ri[1], sx, -uy);
    const float rz = __builtin_fmaf(c.ri[2], sx, -uz);
    float ox = c.eye[0], oy = c.eye[1], oz = c.eye[2], tx, ty, tz;
    if (c.dof) {
      // lens sample (rend.c:3320-3337): radius sqrt(u2), angle 2 pi u3 with
      // cos and sin from the host's libmvec, tabulated over all 2^23 u3
      const float rad = sqrtf(u[2][l]);
      const float2 cs = lens[m3[l]];
      const float lx = cs.x * rad, ly = cs.y * rad;
      ox = __builtin_fmaf(__builtin_fmaf(c.ri[0], lx, c.up[0] * ly), c.focrad, c.eye[0]);
      oy = __builtin_fmaf(__builtin_fmaf(c.ri[1], lx, c.up[1] * ly), c.focrad, c.eye[1]);
      oz = __builtin_fmaf(__builtin_fmaf(c.ri[2], lx, c.up[2] * ly), c.focrad, c.eye[2]);
      // towards the pixel's point on the focal plane (rend.c:3339-3366)
      tx = (__builtin_fmaf(c.fwd[0], c.fdist, c.eye[0]) + rx) - ox;
      ty = (__builtin_fmaf(c.fwd[1], c.fdist, c.eye[1]) + ry) - oy;
      tz = (__builtin_fmaf(c.fwd[2], c.fdist, c.eye[2]) + rz) - oz;
    } else {
      tx = __builtin_fmaf(c.fwd[0], c.fdist, rx);
      ty = __builtin_fmaf(c.fwd[1], c.fdist, ry);
      tz = __builtin_fmaf(c.fwd[2], c.fdist, rz);
    }
    const float rl = rsqrt_host(__builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz)));
    float4 *o = (float4 *)&out[(size_t)t * 8u + (uint32_t)l];
    o[0] = make_float4(ox, oy, oz, 3.402823466e38f);
    o[1] = make_float4(tx * rl, ty * rl, tz * rl, 0.0f);
  }
}

// Normal tint of the hits of the reference's primary rays (k_camrays_ref
// order: ray = (block * packets + packet) * 8 + lane, packets in Morton
// order of 4 x 2 pixels), accumulated as accum_pckt does.  Each pixel of the
// block grid has exactly one ray when the packet grid of a block is 2^a x 2^a
// or 2^(a+1) x 2^a (checked by the host); a pixel outside width x height is
// skipped.
__global__ void k_shade_normals_ref(const rhit *__restrict__ hits,
                                    const dinst *__restrict__ insts,
                                    const RNrm *__restrict__ nrms,
                                    const RMtl *__restrict__ mtls, float bgr, float bgg,
                                    float bgb, uint32_t bszx, uint32_t bszy, uint32_t blkcntx,
                                    uint32_t npk, uint32_t n, uint32_t width, uint32_t height,
                                    float invsamples, float *__restrict__ acc,
                                    uint32_t *__restrict__ buf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pk = i >> 3, lane = i & 7u;
  const uint32_t blk = pk / npk, k = pk % npk;
  const uint32_t x = (blk % blkcntx) * bszx + mort_dec(k) * 4u + (lane & 3u);
  const uint32_t y = (blk / blkcntx) * bszy + mort_dec(k >> 1) * 2u + (lane >> 2);
  if (x >= width || y >= height) return;
  float r, g, b;
  shade_hit(hits[i], insts, nrms, mtls, bgr, bgg, bgb, r, g, b);
  accum_px(acc, buf, (size_t)y * width + x, r, g, b, invsamples);
}

extern "C" int rk_camrays_ref_setup(const unsigned *tab, void *stream) {
  return (int)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rsqrt_tab), tab, 8192 * sizeof(uint32_t), 0,
                                     hipMemcpyHostToDevice, (hipStream_t)stream);
}

// lens: 2^23 (cos, sin) pairs of the thin-lens angle, needed when c->dof.
extern "C" int rk_camrays_ref(const CamArgs *c, unsigned bszx, unsigned bszy, unsigned blkcntx,
                              unsigned nblk, unsigned samples, const void *lens, void *out,
                              void *stream) {
  const unsigned npk = bszx * bszy / 8u, n = nblk * npk;
  if (n == 0) return 0;
  if (c->dof && !lens) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_camrays_ref, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream,
                     *c, bszx, bszy, blkcntx, nblk, npk, samples, (const float2 *)lens,
                     (rray *)out);
  return (int)hipGetLastError();
}

extern "C" int rk_camrays(const CamArgs *c, unsigned rw, unsigned rh,
                          unsigned spp, unsigned seed, unsigned frame,
                          void *out, void *stream) {
  const uint32_t n = rw * rh * spp;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_camrays, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, *c, rw, rh, spp, seed, frame, (rray *)out);
  return (int)hipGetLastError();
}

extern "C" int rk_shade_normals(const void *hits, const void *insts,
                                const void *nrms, const void *mtls,
                                const float bg[3], unsigned rw, unsigned rh,
                                unsigned width, unsigned spp, float invsamples,
                                float *acc, unsigned *buf, void *stream) {
  const uint32_t n = rw * rh;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_shade_normals, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (const rhit *)hits, (const dinst *)insts,
                     (const RNrm *)nrms, (const RMtl *)mtls, bg[0], bg[1], bg[2],
                     rw, rh, width, spp, invsamples, acc, buf);
  return (int)hipGetLastError();
}

extern "C" int rk_diffuse_rays(const void *prim, const void *hits,
                               const void *insts, const void *tris, unsigned n,
                               unsigned seed, void *out, void *stream) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_diffuse_rays, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (const rray *)prim, (const rhit *)hits,
                     (const dinst *)insts, (const RTri *)tris, n, seed, (rray *)out);
  return (int)hipGetLastError();
}

extern "C" int rk_copy(const void *src, void *dst, unsigned long long bytes, void *stream) {
  const size_t n = bytes / 16;
  if (!n) return 0;
  hipLaunchKernelGGL(k_copy_f4, dim3(4096), dim3(256), 0, (hipStream_t)stream,
                     (const float4 *)src, (float4 *)dst, n);
  return (int)hipGetLastError();
}

extern "C" int rk_fill_f32(float *p, unsigned long long n, float v, void *stream) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, p, (size_t)n, v);
  return (int)hipGetLastError();
}

extern "C" int rk_shade_normals_ref(const void *hits, const void *insts, const void *nrms,
                                    const void *mtls, const float bg[3], unsigned bszx,
                                    unsigned bszy, unsigned blkcntx, unsigned nblk,
                                    unsigned width, unsigned height, float invsamples,
                                    float *acc, unsigned *buf, void *stream) {
  const uint32_t npk = bszx * bszy / 8u, n = nblk * npk * 8u;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_shade_normals_ref, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (const rhit *)hits, (const dinst *)insts,
                     (const RNrm *)nrms, (const RMtl *)mtls, bg[0], bg[1], bg[2], bszx, bszy,
                     blkcntx, npk, n, width, height, invsamples, acc, buf);
  return (int)hipGetLastError();
}
