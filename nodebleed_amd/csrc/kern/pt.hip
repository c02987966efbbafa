// Wavefront path tracing kernels (REND_MODE_PATHTRACE and batched views).
//
// Estimator: the reference's trace3 (rend.c:3205-3291) generalised --
// Lambertian BRDF albedo/pi, next-event estimation towards emissive
// triangles with an any-hit shadow ray from p + 1e-3 * l (EPS2), a
// cosine-weighted indirect ray (create_onb + rand_hemicos,
// rend.c:3126-3140), background radiance on a miss (rend.c:3220), emission
// counted where a camera ray sees an emitter, Russian roulette after a
// configurable depth and a maximum path length.  Optionally the indirect
// direction comes from a one-sample MIS mixture of the BSDF and a learned
// vMF-mixture proposal (guide.hip), keeping the estimator unbiased.
//
// One pass = every active path advances by one segment:
//   k_pt_gen      camera rays for a batch of paths (views x pixels x spp)
//   (trace)       closest hit of the ray queue   (trace.hip, count on device)
//   k_pt_shade    emission, NEE shadow ray, next ray; compacts the next-ray,
//                 shadow and training-record queues with wave ballots, mbcnt
//                 and one atomic per queue per block
//   (occluded)    any hit of the shadow queue
//   k_pt_shadow   adds unoccluded NEE contributions to their paths
//   k_pt_resolve  adds path radiance to the view accumulators, ARGB8 output
// Counts live in device memory, so a whole frame is one launch sequence that
// can be captured in a HIP graph.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../dlayout.h"
#include "../ptlayout.h"
#include "rend_rays.h"
#include "rng.h"
#include "guide_dev.h"

namespace {

constexpr float kPi = 3.14159265f;
constexpr float kInvPi = 0.318309886f;
constexpr float kEps = 1e-3f;   // EPS2, util.h:8
constexpr float kBig = 3.402823466e+38f;

struct RNrm {
  float n0[3];
  uint32_t mtlid;
  float n1[3];
  uint32_t pad0;
  float n2[3];
  uint32_t pad1;
};

struct RMtl {
  float col[3];
  float metallic, roughness, ior;
  uint32_t flags, pad0;
};

struct RTri {
  float v0[3], p0, v1[3], p1, v2[3], p2;
};

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// A direction component of exactly +-0 gives infinite 1/d and o/d; under the
// reference's slab rules (x86 MAXPS/MINPS, NaN operand wins) the resulting
// NaN distances make every box test pass, and such a ray visits the whole
// tree.  Sampled directions hit exact zeros with probability ~2^-24 per
// draw (more often next to axis-aligned normals); nudging the component to
// +-1e-20 is a measure-zero change of the sampled distribution.
__device__ __forceinline__ float nonzero_dir(float x) {
  return x == 0.0f ? __builtin_copysignf(1e-20f, x) : x;
}

__device__ __forceinline__ void store_ray(rray *out, uint32_t i, float ox, float oy,
                                          float oz, float tmax, float dx, float dy,
                                          float dz) {
  float4 *o = (float4 *)&out[i];
  o[0] = make_float4(ox, oy, oz, tmax);
  o[1] = make_float4(nonzero_dir(dx), nonzero_dir(dy), nonzero_dir(dz), 0.0f);
}

// Batch layout: path index -> (view, pixel, sample) by the view table's
// prefix sums of w*h*spp.
__device__ __forceinline__ uint32_t find_view(const PtView *views, uint32_t nviews,
                                              uint32_t g) {
  uint32_t lo = 0, hi = nviews;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].first <= g) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void k_pt_gen(PtArgs a, const PtView *__restrict__ views,
                         PtPath *__restrict__ paths, rray *__restrict__ rays,
                         uint32_t *__restrict__ qpath, uint32_t *__restrict__ counts,
                         uint32_t first, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    counts[PT_CNT_CUR] = n;
    counts[PT_CNT_NEXT] = 0;
    counts[PT_CNT_SHADOW] = 0;
    counts[PT_CNT_TRAIN] = 0;
  }
  if (i >= n) return;
  const uint32_t g = first + i;
  const uint32_t vi = find_view(views, a.nviews, g);
  const PtView &v = views[vi];
  const uint32_t local = g - v.first;
  const uint32_t spp = v.spp;
  const uint32_t px = local / spp, s = local % spp;
  const uint32_t x = px % v.w, y = px / v.w;
  const uint32_t key = rng_key(a.seed ^ v.stream, v.sample_base + s, px);
  const CamArgs &c = v.cam;
  const float u0 = rng_u01(key, 0), u1 = rng_u01(key, 1);
  float sx = (float)x + (u0 - 0.5f);
  float sy = (float)y + (u1 - 0.5f);
  sx = __builtin_fmaf(sx, c.rw, -0.5f) * c.fovfdist * c.aspect;
  sy = __builtin_fmaf(sy, c.rh, -0.5f) * c.fovfdist;
  float ox = c.eye[0], oy = c.eye[1], oz = c.eye[2], tx, ty, tz;
  const float rx0 = __builtin_fmaf(c.ri[0], sx, -(c.up[0] * sy));
  const float ry0 = __builtin_fmaf(c.ri[1], sx, -(c.up[1] * sy));
  const float rz0 = __builtin_fmaf(c.ri[2], sx, -(c.up[2] * sy));
  if (c.dof) {
    const float rad = sqrtf(rng_u01(key, 2));
    float sn, cs;
    sincosf(6.28318531f * rng_u01(key, 3), &sn, &cs);
    const float lx = cs * rad, ly = sn * rad;
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
  const float il = 1.0f / sqrtf(tx * tx + ty * ty + tz * tz);
  PtPath p;
  p.thr[0] = p.thr[1] = p.thr[2] = 1.0f;
  p.L[0] = p.L[1] = p.L[2] = 0.0f;
  p.pix = v.acc_ofs + y * v.pitch + x;
  p.key = key;
  p.depth = 0;
  p.pdf_prev = 0.0f;
  p.rec = 0xffffffffu;
  p.rec_prev = 0xffffffffu;
  paths[i] = p;
  store_ray(rays, i, ox, oy, oz, kBig, tx * il, ty * il, tz * il);
  qpath[i] = i;
}

constexpr uint32_t kLightScan = 64;  // up to this many emitters: per-point choice
constexpr float kLightFloor = 0.1f;  // share of the area*luminance choice kept

// Importance of light L for a receiver at p with normal n: power times the
// two cosines at the triangle's centroid over the squared distance.
__device__ __forceinline__ float light_importance(const PtLight &L, const float p[3],
                                                  const float n[3]) {
  float d[3];
  for (int j = 0; j < 3; j++) d[j] = L.v0[j] + (L.e0[j] + L.e1[j]) * (1.0f / 3.0f) - p[j];
  const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float cr = fmaxf(n[0] * d[0] + n[1] * d[1] + n[2] * d[2], 0.0f);
  const float cl = fmaxf(-(L.n[0] * d[0] + L.n[1] * d[1] + L.n[2] * d[2]), 0.0f);
  const float lum = 0.2126f * L.E[0] + 0.7152f * L.E[1] + 0.0722f * L.E[2];
  // cr cl / d2 is cos_r cos_l d2 / d2^2; area keeps small close lights finite
  const float dd = fmaxf(d2, L.area);
  return lum * L.area * cr * cl / (dd * dd);
}

// Uniform point on an emissive triangle.  With few emitters the triangle is
// chosen per receiver: probability (1 - kLightFloor) * importance / total +
// kLightFloor * (area*luminance share), so every emitter keeps a nonzero
// probability and the estimator stays unbiased; with many emitters, by the
// area*luminance CDF alone.
__device__ __forceinline__ void sample_light(const PtLight *__restrict__ lights,
                                             const float *__restrict__ cdf,
                                             uint32_t nl, const float p[3], const float n[3],
                                             float u0, float u1, float u2,
                                             float q[3], float nl3[3], float E[3],
                                             float &pdf_area) {
  uint32_t k = 0;
  float pk = 0.0f;
  if (nl <= kLightScan) {
    float tot = 0.0f;
    for (uint32_t i = 0; i < nl; i++) tot += light_importance(lights[i], p, n);
    const float wi = tot > 0.0f ? (1.0f - kLightFloor) / tot : 0.0f;
    const float wb = tot > 0.0f ? kLightFloor : 1.0f;
    float acc = 0.0f;
    k = nl - 1;
    for (uint32_t i = 0; i < nl; i++) {
      const float pi = wi * light_importance(lights[i], p, n) + wb * lights[i].pick;
      acc += pi;
      if (u0 < acc) { k = i; pk = pi; break; }
      pk = pi;
    }
  } else {
    uint32_t lo = 0, hi = nl;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cdf[mid] <= u0) lo = mid + 1; else hi = mid;
    }
    k = lo < nl ? lo : nl - 1;
    pk = lights[k].pick;
  }
  const PtLight &L = lights[k];
  const float su = sqrtf(u1);
  const float b0 = su * (1.0f - u2), b1 = su * u2;  // uniform on the triangle
  for (int j = 0; j < 3; j++) {
    q[j] = L.v0[j] + L.e0[j] * b0 + L.e1[j] * b1;
    nl3[j] = L.n[j];
    E[j] = L.E[j];
  }
  pdf_area = pk / L.area;
}

// Slots in the training-record, next-ray and shadow queues for one trip of
// a 256-thread block: one atomic per queue per block (every wave of the
// launch appends to the same three counters, so per-wave atomics queue up
// on three addresses), slots in wave then lane order.
struct BlockAppend {
  uint32_t cnt[4][3];
  uint32_t base[3];
};

__device__ __forceinline__ void block_append3(uint32_t *counts, bool w0, bool w1, bool w2,
                                              uint32_t &s0, uint32_t &s1, uint32_t &s2,
                                              BlockAppend &S) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint64_t m0 = __ballot(w0), m1 = __ballot(w1), m2 = __ballot(w2);
  if (lane == 0) {
    S.cnt[wv][0] = (uint32_t)__popcll(m0);
    S.cnt[wv][1] = (uint32_t)__popcll(m1);
    S.cnt[wv][2] = (uint32_t)__popcll(m2);
  }
  __syncthreads();
  if (threadIdx.x < 3u) {
    const uint32_t q = threadIdx.x;
    const uint32_t tot = S.cnt[0][q] + S.cnt[1][q] + S.cnt[2][q] + S.cnt[3][q];
    const uint32_t ctr = q == 0 ? PT_CNT_TRAIN : (q == 1 ? PT_CNT_NEXT : PT_CNT_SHADOW);
    S.base[q] = tot ? atomicAdd(&counts[ctr], tot) : 0u;
  }
  __syncthreads();
  uint32_t b0 = S.base[0], b1 = S.base[1], b2 = S.base[2];
  for (uint32_t k = 0; k < wv; k++) {
    b0 += S.cnt[k][0];
    b1 += S.cnt[k][1];
    b2 += S.cnt[k][2];
  }
  __syncthreads();  // S is reused by the next trip
  const uint32_t l0 = (uint32_t)m0, h0 = (uint32_t)(m0 >> 32);
  const uint32_t l1 = (uint32_t)m1, h1 = (uint32_t)(m1 >> 32);
  const uint32_t l2 = (uint32_t)m2, h2 = (uint32_t)(m2 >> 32);
  s0 = w0 ? b0 + __builtin_amdgcn_mbcnt_hi(h0, __builtin_amdgcn_mbcnt_lo(l0, 0u)) : 0xffffffffu;
  s1 = w1 ? b1 + __builtin_amdgcn_mbcnt_hi(h1, __builtin_amdgcn_mbcnt_lo(l1, 0u)) : 0xffffffffu;
  s2 = w2 ? b2 + __builtin_amdgcn_mbcnt_hi(h2, __builtin_amdgcn_mbcnt_lo(l2, 0u)) : 0xffffffffu;
}

__global__ __launch_bounds__(256) void k_pt_shade(PtArgs a, PtPath *__restrict__ paths,
                           const rray *__restrict__ rays, const rhit *__restrict__ hits,
                           const uint32_t *__restrict__ qpath, rray *__restrict__ nrays,
                           uint32_t *__restrict__ nqpath, rray *__restrict__ srays,
                           uint32_t *__restrict__ spath, float *__restrict__ scontrib,
                           uint32_t *__restrict__ counts, const dinst *__restrict__ insts,
                           const RTri *__restrict__ tris, const RNrm *__restrict__ nrms,
                           const RMtl *__restrict__ mtls, const PtLight *__restrict__ lights,
                           const float *__restrict__ lcdf, const float *__restrict__ guide,
                           GdRecord *__restrict__ train, uint32_t depth) {
  __shared__ BlockAppend s_app;
  const uint32_t n = counts[PT_CNT_CUR];
  const uint32_t stride = gridDim.x * blockDim.x;
  // uniform trip count so every thread of the block takes part in the appends
  const uint32_t trips = (n + stride - 1) / stride;
  for (uint32_t t = 0; t < trips; t++) {
    const uint32_t i = t * stride + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    bool want_next = false, want_shadow = false, want_rec = false;
    GdRecord rec;
    float nro[3], nrd[3], sro[3], srd[3], stmax = 0.0f, sc[3] = {0, 0, 0};
    uint32_t pi = 0;
    PtPath p;
    if (valid) {
      pi = qpath[i];
      p = paths[pi];
      const rray r = rays[i];
      const rhit h = hits[i];
      const uint32_t key = p.key;
      const uint32_t dbase = 8u + depth * 8u;
      if (h.id == 0xffffffffu) {
        // miss: background radiance (rend.c:3220-3221)
        for (int j = 0; j < 3; j++) p.L[j] += p.thr[j] * a.bg[j];
      } else {
        const uint32_t inst = h.id & 0xfffu, tri = h.id >> 12;
        const dinst &in = insts[inst];
        const RTri &T = tris[in.triofs + tri];
        const RNrm &rn = nrms[in.triofs + tri];
        const RMtl &mt = mtls[rn.mtlid];
        const float *m = in.inv;
        // geometric normal to world (inverse transpose)
        const float e0x = T.v1[0] - T.v0[0], e0y = T.v1[1] - T.v0[1], e0z = T.v1[2] - T.v0[2];
        const float e1x = T.v2[0] - T.v0[0], e1y = T.v2[1] - T.v0[1], e1z = T.v2[2] - T.v0[2];
        const float gx0 = e0y * e1z - e0z * e1y, gy0 = e0z * e1x - e0x * e1z,
                    gz0 = e0x * e1y - e0y * e1x;
        float gx = m[0] * gx0 + m[4] * gy0 + m[8] * gz0;
        float gy = m[1] * gx0 + m[5] * gy0 + m[9] * gz0;
        float gz = m[2] * gx0 + m[6] * gy0 + m[10] * gz0;
        float il = 1.0f / sqrtf(gx * gx + gy * gy + gz * gz);
        gx *= il; gy *= il; gz *= il;
        const bool front = gx * r.dx + gy * r.dy + gz * r.dz < 0.0f;
        // emission seen by a camera ray (NEE covers later vertices)
        if ((mt.flags & 0x1000u) && depth == 0 && front)
          for (int j = 0; j < 3; j++) p.L[j] += p.thr[j] * mt.col[j];
        if (!(mt.flags & 0x1000u)) {
          // shading normal (calc_nrm, rend.c:3109-3116), both facing the ray
          const float u = h.u, v = h.v, w = 1.0f - u - v;
          const float sx0 = rn.n1[0] * u + (rn.n2[0] * v + rn.n0[0] * w);
          const float sy0 = rn.n1[1] * u + (rn.n2[1] * v + rn.n0[1] * w);
          const float sz0 = rn.n1[2] * u + (rn.n2[2] * v + rn.n0[2] * w);
          float nx = m[0] * sx0 + m[4] * sy0 + m[8] * sz0;
          float ny = m[1] * sx0 + m[5] * sy0 + m[9] * sz0;
          float nz = m[2] * sx0 + m[6] * sy0 + m[10] * sz0;
          il = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
          nx *= il; ny *= il; nz *= il;
          if (!front) { gx = -gx; gy = -gy; gz = -gz; }
          if (nx * gx + ny * gy + nz * gz < 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
          const float px = r.ox + r.dx * h.t, py = r.oy + r.dy * h.t, pz = r.oz + r.dz * h.t;
          const float alb[3] = {mt.col[0], mt.col[1], mt.col[2]};
          // next-event estimation
          if (a.nlights) {
            float q[3], ln[3], E[3], pdfa;
            const float sp[3] = {px, py, pz}, sn[3] = {nx, ny, nz};
            sample_light(lights, lcdf, a.nlights, sp, sn, rng_u01(key, dbase + 0),
                         rng_u01(key, dbase + 1), rng_u01(key, dbase + 2), q, ln, E, pdfa);
            float lx = q[0] - px, ly = q[1] - py, lz = q[2] - pz;
            const float d2 = lx * lx + ly * ly + lz * lz;
            const float dist = sqrtf(d2);
            lx /= dist; ly /= dist; lz /= dist;
            const float cs = nx * lx + ny * ly + nz * lz;
            const float cl = -(ln[0] * lx + ln[1] * ly + ln[2] * lz);
            if (cs > 0.0f && cl > 0.0f && gx * lx + gy * ly + gz * lz > 0.0f && dist > 4.0f * kEps) {
              const float f = kInvPi * cs * cl / (d2 * pdfa);
              for (int j = 0; j < 3; j++) sc[j] = p.thr[j] * alb[j] * E[j] * f;
              sro[0] = px + lx * kEps; sro[1] = py + ly * kEps; sro[2] = pz + lz * kEps;
              srd[0] = lx; srd[1] = ly; srd[2] = lz;
              stmax = dist * (1.0f - 1e-4f) - 2.0f * kEps;
              want_shadow = (sc[0] + sc[1] + sc[2]) > 0.0f;
            }
          }
          // indirect direction
          if (depth + 1 < a.maxdepth) {
            const float sgn = copysignf(1.0f, nz);
            const float aa = -1.0f / (sgn + nz);
            const float bb = nx * ny * aa;
            const float b1x = 1.0f + sgn * nx * nx * aa, b1y = sgn * bb, b1z = -sgn * nx;
            const float b2x = bb, b2y = sgn + ny * ny * aa, b2z = -ny;
            float dx, dy, dz, wgt;
            bool ok = true;
            if (guide && a.guide_mix > 0.0f && depth < a.guide_depth) {
              // at depth 0 consecutive paths of a pixel share one proposal
              const size_t gi = depth == 0 ? i / a.guide_share : i;
              ok = guide_sample(guide + gi * GD_MIX, a.guide_mix, nx, ny, nz, b1x, b1y,
                                b1z, b2x, b2y, b2z, rng_u01(key, dbase + 3),
                                rng_u01(key, dbase + 4), rng_u01(key, dbase + 5), dx, dy, dz,
                                wgt, p.pdf_prev);
            } else {
              float sn, csn;
              sincosf(6.28318531f * rng_u01(key, dbase + 3), &sn, &csn);
              const float u1 = rng_u01(key, dbase + 4);
              const float su = sqrtf(u1), lx = csn * su, ly = sn * su, lz = sqrtf(1.0f - u1);
              dx = b1x * lx + b2x * ly + nx * lz;
              dy = b1y * lx + b2y * ly + ny * lz;
              dz = b1z * lx + b2z * ly + nz * lz;
              wgt = 1.0f;  // cos/pdf * 1/pi = 1
              p.pdf_prev = lz * kInvPi;
            }
            if (ok && gx * dx + gy * dy + gz * dz > 0.0f && wgt > 0.0f) {
              for (int j = 0; j < 3; j++) p.thr[j] *= alb[j] * wgt;
              bool live = true;
              if (depth + 1 >= a.rr_depth) {
                const float mx = fmaxf(p.thr[0], fmaxf(p.thr[1], p.thr[2]));
                const float qs = fminf(mx, 0.95f);
                if (rng_u01(key, dbase + 6) >= qs) live = false;
                else for (int j = 0; j < 3; j++) p.thr[j] /= qs;
              }
              if (live) {
                want_next = true;
                nro[0] = px + dx * kEps; nro[1] = py + dy * kEps; nro[2] = pz + dz * kEps;
                nrd[0] = dx; nrd[1] = dy; nrd[2] = dz;
                if (train && a.train_stride && (pi % a.train_stride) == 0 &&
                    depth < a.guide_depth) {
                  want_rec = true;
                  rec.pos[0] = px; rec.pos[1] = py; rec.pos[2] = pz;
                  rec.nrm[0] = gx; rec.nrm[1] = gy; rec.nrm[2] = gz;
                  rec.dir[0] = dx; rec.dir[1] = dy; rec.dir[2] = dz;
                  rec.pdf = p.pdf_prev;
                  rec.lbefore = 0.2126f * p.L[0] + 0.7152f * p.L[1] + 0.0722f * p.L[2];
                  rec.thr = 0.2126f * p.thr[0] + 0.7152f * p.thr[1] + 0.0722f * p.thr[2];
                  rec.path = pi;
                  rec.target = 0.0f;
                  rec.l1 = 0.0f;
                }
              }
            }
          }
        }
      }
      p.depth = depth + 1;
    }
    uint32_t rs, ns, ss;
    block_append3(counts, want_rec, want_next, want_shadow, rs, ns, ss, s_app);
    if (valid) {
      p.rec_prev = p.rec;  // the previous vertex's record (its target gathers this vertex's NEE)
      p.rec = 0xffffffffu;
      if (want_rec && rs < a.train_cap) {
        // four 16-byte stores (a 64-byte record per lane)
        float4 *dst = (float4 *)&train[rs];
        dst[0] = make_float4(rec.pos[0], rec.pos[1], rec.pos[2], rec.nrm[0]);
        dst[1] = make_float4(rec.nrm[1], rec.nrm[2], rec.dir[0], rec.dir[1]);
        dst[2] = make_float4(rec.dir[2], rec.pdf, rec.lbefore, rec.thr);
        dst[3] = make_float4(__uint_as_float(rec.path), rec.target, rec.l1, 0.0f);
        p.rec = rs;
      }
      paths[pi] = p;
    }
    if (want_next) {
      store_ray(nrays, ns, nro[0], nro[1], nro[2], kBig, nrd[0], nrd[1], nrd[2]);
      nqpath[ns] = pi;
    }
    if (want_shadow) {
      store_ray(srays, ss, sro[0], sro[1], sro[2], stmax, srd[0], srd[1], srd[2]);
      spath[ss] = pi;
      scontrib[3 * ss + 0] = sc[0];
      scontrib[3 * ss + 1] = sc[1];
      scontrib[3 * ss + 2] = sc[2];
    }
  }
}

__global__ void k_pt_shadow(PtPath *__restrict__ paths, const uint32_t *__restrict__ spath,
                            const float *__restrict__ scontrib,
                            const uint8_t *__restrict__ occ,
                            const uint32_t *__restrict__ counts,
                            GdRecord *__restrict__ rec) {
  const uint32_t n = counts[PT_CNT_SHADOW];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    if (occ[i]) continue;
    PtPath &p = paths[spath[i]];  // one shadow ray per path per pass
    const float c0 = scontrib[3 * i + 0], c1 = scontrib[3 * i + 1], c2 = scontrib[3 * i + 2];
    p.L[0] += c0;
    p.L[1] += c1;
    p.L[2] += c2;
    // direct light of this vertex is not incident radiance along the
    // continuation: keep it out of the training target
    const float lum = 0.2126f * c0 + 0.7152f * c1 + 0.0722f * c2;
    if (rec && p.rec != 0xffffffffu) rec[p.rec].lbefore += lum;
    // ... and is the previous vertex's next-vertex direct light
    if (rec && p.rec_prev != 0xffffffffu) rec[p.rec_prev].l1 += lum;
  }
}

// Moves the next queue to the current one (counts only; buffers alternate
// on the host side) and clears the shadow queue.
__global__ void k_pt_advance(uint32_t *counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    // rays traced this segment: extension queue + shadow queue
    *(unsigned long long *)&counts[4] += (unsigned long long)counts[PT_CNT_CUR] +
                                         counts[PT_CNT_SHADOW];
    counts[PT_CNT_CUR] = counts[PT_CNT_NEXT];
    counts[PT_CNT_NEXT] = 0;
    counts[PT_CNT_SHADOW] = 0;
  }
}

// Path radiance into the accumulators.  The samples of one pixel are
// consecutive paths of a batch, so each wave first sums every run of equal
// pixels (segmented reduction over the lanes: after the step with offset o a
// lane holds the sum of its run's next 2o lanes) and only the first lane of
// a run adds to the accumulator.  Runs can continue in the next wave or
// batch, so the additions stay atomic.
__global__ void k_pt_accum(const PtPath *__restrict__ paths, uint32_t n,
                           float *__restrict__ acc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = lane_id();
  const bool valid = i < n;
  uint32_t pix = 0xffffffffu;
  float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
  if (valid) {
    const PtPath &p = paths[i];
    pix = p.pix;
    l0 = p.L[0];
    l1 = p.L[1];
    l2 = p.L[2];
  }
  const uint32_t prev = __shfl_up(pix, 1);
  const bool head = valid && (lane == 0 || prev != pix);
  // run index of each lane: heads at or below it
  const uint64_t hm = __ballot(head);
  const uint32_t run = __builtin_amdgcn_mbcnt_hi(
      (uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u)) + (head ? 1u : 0u);
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t r2 = __shfl_down(run, o);
    const float a0 = __shfl_down(l0, o), a1 = __shfl_down(l1, o), a2 = __shfl_down(l2, o);
    if (lane + o < 64 && r2 == run) {
      l0 += a0;
      l1 += a1;
      l2 += a2;
    }
  }
  if (head) {
    atomicAdd(&acc[3 * (size_t)pix + 0], l0);
    atomicAdd(&acc[3 * (size_t)pix + 1], l1);
    atomicAdd(&acc[3 * (size_t)pix + 2], l2);
  }
}

__device__ __forceinline__ uint32_t pack8(float a) {
  return (uint32_t)(long long)(255.0f * a) & 0xffu;
}

// ARGB8 of acc * invsamples for one view (accum_pckt, rend.c:3565-3580).
__global__ void k_pt_resolve(const float *__restrict__ acc, uint32_t *__restrict__ buf,
                             uint32_t npx, float invsamples) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  buf[i] = 0xff000000u | pack8(acc[3 * i] * invsamples) << 16 |
           pack8(acc[3 * i + 1] * invsamples) << 8 | pack8(acc[3 * i + 2] * invsamples);
}

}  // namespace

// ---- C-ABI ------------------------------------------------------------------

extern "C" int rk_pt_gen(const PtArgs *a, const void *views, void *paths, void *rays,
                         void *qpath, unsigned *counts, unsigned first, unsigned n,
                         void *stream) {
  hipLaunchKernelGGL(k_pt_gen, dim3((n + 255) / 256 ? (n + 255) / 256 : 1), dim3(256), 0,
                     (hipStream_t)stream, *a, (const PtView *)views, (PtPath *)paths,
                     (rray *)rays, (uint32_t *)qpath, counts, first, n);
  return (int)hipGetLastError();
}

extern "C" int rk_pt_shade(const PtArgs *a, void *paths, const void *rays, const void *hits,
                           const void *qpath, void *nrays, void *nqpath, void *srays,
                           void *spath, float *scontrib, unsigned *counts, const void *insts,
                           const void *tris, const void *nrms, const void *mtls,
                           const void *lights, const float *lcdf, const float *guide,
                           void *train, unsigned depth, unsigned grid, void *stream) {
  hipLaunchKernelGGL(k_pt_shade, dim3(grid), dim3(256), 0, (hipStream_t)stream, *a,
                     (PtPath *)paths, (const rray *)rays, (const rhit *)hits,
                     (const uint32_t *)qpath, (rray *)nrays, (uint32_t *)nqpath,
                     (rray *)srays, (uint32_t *)spath, scontrib, counts,
                     (const dinst *)insts, (const RTri *)tris, (const RNrm *)nrms,
                     (const RMtl *)mtls, (const PtLight *)lights, lcdf, guide,
                     (GdRecord *)train, depth);
  return (int)hipGetLastError();
}

extern "C" int rk_pt_shadow(void *paths, const void *spath, const float *scontrib,
                            const void *occ, const unsigned *counts, void *rec,
                            unsigned grid, void *stream) {
  hipLaunchKernelGGL(k_pt_shadow, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                     (PtPath *)paths, (const uint32_t *)spath, scontrib,
                     (const uint8_t *)occ, counts, (GdRecord *)rec);
  return (int)hipGetLastError();
}

extern "C" int rk_pt_advance(unsigned *counts, void *stream) {
  hipLaunchKernelGGL(k_pt_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, counts);
  return (int)hipGetLastError();
}

extern "C" int rk_pt_accum(const void *paths, unsigned n, float *acc, void *stream) {
  if (!n) return 0;
  hipLaunchKernelGGL(k_pt_accum, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     (const PtPath *)paths, n, acc);
  return (int)hipGetLastError();
}

extern "C" int rk_pt_resolve(const float *acc, unsigned *buf, unsigned npx, float invs,
                             void *stream) {
  if (!npx) return 0;
  hipLaunchKernelGGL(k_pt_resolve, dim3((npx + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, acc, buf, npx, invs);
  return (int)hipGetLastError();
}
