// Learned path-guiding proposal: inference, training and optimiser kernels.
//
// Model: multiresolution hash encoding of the shading position (Mueller,
// Evans, Schied, Keller 2022, "Instant Neural Graphics Primitives": GD_LEVELS
// levels with geometric resolutions GD_NMIN..GD_NMAX, dense indexing while a
// level fits the table, spatial hash x ^ 2654435761 y ^ 805459861 z
// otherwise, trilinear interpolation), concatenated with the shading normal
// and a constant 1, through a 3 x 64 ReLU MLP whose outputs parameterise a
// mixture of GD_LOBES von Mises-Fisher lobes over the sphere.
//
// MLP layers run on MFMA in BF16 with FP32 accumulation
// (v_mfma_f32_16x16x32_bf16): one wave owns 64 samples, activations stay in
// LDS between layers, weights are read as MFMA B fragments from a BF16
// mirror of the FP32 master parameters.  Training minimises the Monte Carlo
// estimate of the KL divergence between the normalised product
// L_i(w) max(0, n.w) and the mixture, -(L cos / q) log p(w), with gradients
// through the mixture, the MLP (MFMA) and the hash grid (atomic scatter),
// and Adam on the FP32 master copy, which then refreshes the FP16 grid and
// BF16 weight mirrors.  vMF sampling follows Jakob 2012 ("Numerically stable
// sampling of the von Mises Fisher distribution on S^2").

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../ptlayout.h"
#include "rend_rays.h"
#include "guide_math.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__constant__ float c_gd_res[GD_LEVELS];      // grid resolution per level
__constant__ uint32_t c_gd_dense[GD_LEVELS];  // 1 when (res+1)^3 <= GD_TABLE

namespace {

constexpr int kWave = 64;

// Output layer layout: the 4 lane groups of a sample hold kLpg lobes each
// (kSlots outputs: logit, mean xyz, log concentration per lobe), spread over
// kOutTiles 16-row MFMA tiles.
constexpr uint32_t kLpg = GD_LOBES / 4;
constexpr uint32_t kSlots = 5 * kLpg;
constexpr uint32_t kOutTiles = (kSlots + 3) / 4;
static_assert(GD_LOBES % 4 == 0 && kOutTiles <= 4 && GD_MIX <= GD_OUT,
              "the lobes' outputs fit the 64-row output layer");

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Index and weight of corner c of level l for a normalised position.
__device__ __forceinline__ void grid_corner(int l, const float p[3], int c, uint32_t &idx,
                                            float &w) {
  const float res = c_gd_res[l];
  float wt = 1.0f;
  uint32_t ix[3];
  for (int a = 0; a < 3; a++) {
    const float x = p[a] * res;
    const float f = floorf(x);
    const float fr = x - f;
    const uint32_t bit = (c >> a) & 1u;
    ix[a] = (uint32_t)f + bit;
    wt *= bit ? fr : 1.0f - fr;
  }
  if (c_gd_dense[l]) {
    const uint32_t r1 = (uint32_t)res + 1u;
    idx = ix[0] + ix[1] * r1 + ix[2] * r1 * r1;
  } else {
    idx = (ix[0] ^ (ix[1] * 2654435761u) ^ (ix[2] * 805459861u)) & (GD_TABLE - 1u);
  }
  w = wt;
}

__device__ __forceinline__ void normalise_pos(const PtArgs &a, const float p[3], float q[3]) {
  for (int k = 0; k < 3; k++) {
    float v = (p[k] - a.gmin[k]) * a.ginv[k];
    q[k] = fminf(fmaxf(v, 0.0f), 0.99999f);
  }
}

// Input row (GD_IN values) of one sample into LDS as bf16.
__device__ __forceinline__ void encode(const PtArgs &a, const __half *__restrict__ grid,
                                       const float p[3], const float n[3], __bf16 *row) {
  float q[3];
  normalise_pos(a, p, q);
  for (int l = 0; l < GD_LEVELS; l++) {
    float f0 = 0.0f, f1 = 0.0f;
    for (int c = 0; c < 8; c++) {
      uint32_t idx;
      float w;
      grid_corner(l, q, c, idx, w);
      const __half2 g = *(const __half2 *)&grid[((size_t)l * GD_TABLE + idx) * GD_FEAT];
      f0 += w * __low2float(g);
      f1 += w * __high2float(g);
    }
    row[2 * l] = (__bf16)f0;
    row[2 * l + 1] = (__bf16)f1;
  }
  row[24] = (__bf16)n[0];
  row[25] = (__bf16)n[1];
  row[26] = (__bf16)n[2];
  row[27] = (__bf16)1.0f;
  for (int k = 28; k < GD_IN; k++) row[k] = (__bf16)0.0f;
}

// Gathers 8 bf16 at base[0], base[s], ..., base[7s].
__device__ __forceinline__ bf16x8 frag(const __bf16 *base, int s) {
  bf16x8 v;
  if (s == 1) return *(const bf16x8 *)base;
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = base[j * s];
  return v;
}

// D (16x16 tile at rows r0.., cols n0..) += A[rows][k] * B[k][cols] over K,
// with A(i,k) = A[i*ars + k*acs] and B(k,n) = B[k*bks + n*bns].
__device__ __forceinline__ f32x4 tile_mm(const __bf16 *A, int ars, int acs, const __bf16 *B,
                                         int bks, int bns, int r0, int n0, int K) {
  const uint32_t l = lane_id();
  const int i = r0 + (int)(l & 15u), n = n0 + (int)(l & 15u), kq = 8 * (int)(l >> 4);
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bf16x8 av = frag(A + (size_t)i * ars + (size_t)(k0 + kq) * acs, acs);
    const bf16x8 bv = frag(B + (size_t)(k0 + kq) * bks + (size_t)n * bns, bks);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// One dense layer for the wave's 64 rows: out[m][n] = act(in[m][:] . W[n][:] + b[n]).
// W is out-major (W[n][k]); out rows hold `nout` columns.
template <bool RELU>
__device__ __forceinline__ void layer(const __bf16 *in, int kin, const __bf16 *__restrict__ W,
                                      const float *__restrict__ b, int nout, __bf16 *out,
                                      float *outf) {
  const uint32_t l = lane_id();
  for (int rt = 0; rt < kWave; rt += 16)
    for (int ct = 0; ct < nout; ct += 16) {
      const f32x4 d = tile_mm(in, kin, 1, W, 1, kin, rt, ct, kin);
      const int n = ct + (int)(l & 15u);
      const float bn = b[n];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = rt + 4 * (int)(l >> 4) + r;
        float v = d[r] + bn;
        if (RELU) v = fmaxf(v, 0.0f);
        if (out) out[m * nout + n] = (__bf16)v;
        if (outf) outf[m * nout + n] = v;
      }
    }
}

struct Net {
  const __half *grid;   // fp16 mirror [levels][table][feat]
  const __bf16 *w;      // bf16 mirror of W1..W4 (out-major, packed)
  const float *master;  // fp32 master (biases read from here)
};

__device__ __forceinline__ const __bf16 *wptr(const Net &N, int layer_id) {
  switch (layer_id) {
    case 0: return N.w;
    case 1: return N.w + GD_HID * GD_IN;
    case 2: return N.w + GD_HID * GD_IN + GD_HID * GD_HID;
    default: return N.w + GD_HID * GD_IN + 2 * GD_HID * GD_HID;
  }
}

// Forward for the wave: inputs already in act0; leaves hidden activations in
// act1..act3 and raw outputs (fp32) in outf.
__device__ __forceinline__ void forward(const Net &N, __bf16 *act0, __bf16 *act1, __bf16 *act2,
                                        __bf16 *act3, float *outf) {
  layer<true>(act0, GD_IN, wptr(N, 0), N.master + GD_OFS_B1, GD_HID, act1, nullptr);
  __syncthreads();
  layer<true>(act1, GD_HID, wptr(N, 1), N.master + GD_OFS_B2, GD_HID, act2, nullptr);
  __syncthreads();
  layer<true>(act2, GD_HID, wptr(N, 2), N.master + GD_OFS_B3, GD_HID, act3, nullptr);
  __syncthreads();
  layer<false>(act3, GD_HID, wptr(N, 3), N.master + GD_OFS_B4, GD_OUT, nullptr, outf);
  __syncthreads();
}

// Geometric hit point and facing normal of queue entry i (as in k_pt_shade).
struct RTri {
  float v0[3], p0, v1[3], p1, v2[3], p2;
};

__device__ __forceinline__ bool hit_point(const rray &r, const rhit &h,
                                          const dinst *__restrict__ insts,
                                          const RTri *__restrict__ tris, float p[3], float n[3]) {
  if (h.id == 0xffffffffu) return false;
  const dinst &in = insts[h.id & 0xfffu];
  const RTri &T = tris[in.triofs + (h.id >> 12)];
  const float e0x = T.v1[0] - T.v0[0], e0y = T.v1[1] - T.v0[1], e0z = T.v1[2] - T.v0[2];
  const float e1x = T.v2[0] - T.v0[0], e1y = T.v2[1] - T.v0[1], e1z = T.v2[2] - T.v0[2];
  const float gx0 = e0y * e1z - e0z * e1y, gy0 = e0z * e1x - e0x * e1z, gz0 = e0x * e1y - e0y * e1x;
  const float *m = in.inv;
  float gx = m[0] * gx0 + m[4] * gy0 + m[8] * gz0;
  float gy = m[1] * gx0 + m[5] * gy0 + m[9] * gz0;
  float gz = m[2] * gx0 + m[6] * gy0 + m[10] * gz0;
  const float il = 1.0f / sqrtf(gx * gx + gy * gy + gz * gz);
  gx *= il; gy *= il; gz *= il;
  if (gx * r.dx + gy * r.dy + gz * r.dz > 0.0f) { gx = -gx; gy = -gy; gz = -gz; }
  p[0] = r.ox + r.dx * h.t; p[1] = r.oy + r.dy * h.t; p[2] = r.oz + r.dz * h.t;
  n[0] = gx; n[1] = gy; n[2] = gz;
  return true;
}

// LDS per wave: act0 (64 x 32) + act1..3 (64 x 64) bf16, outputs 64 x 64 fp32.
struct WaveLds {
  __attribute__((aligned(16))) __bf16 act0[kWave * GD_IN];
  __attribute__((aligned(16))) __bf16 act1[kWave * GD_HID];
  __attribute__((aligned(16))) __bf16 act2[kWave * GD_HID];
  __attribute__((aligned(16))) __bf16 act3[kWave * GD_HID];
  __attribute__((aligned(16))) float out[kWave * GD_OUT];
};

// Incident-radiance targets of the frame's records from the final path
// radiance; accumulates the sum of training weights into wsum.
__global__ void k_guide_targets(GdRecord *__restrict__ rec, const uint32_t *__restrict__ counts,
                                const PtPath *__restrict__ paths, float *__restrict__ wsum,
                                float *__restrict__ dense, uint32_t cap, uint32_t mode) {
  const uint32_t n = min(counts[PT_CNT_TRAIN], cap);
  float acc = 0.0f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    GdRecord &r = rec[i];
    const PtPath &p = paths[r.path];
    const float Lf = 0.2126f * p.L[0] + 0.7152f * p.L[1] + 0.0722f * p.L[2];
    // mode 0: everything the path gathered after this vertex; mode 1: the
    // next vertex's direct light only (lower variance, one bounce)
    const float gathered = mode ? r.l1 : fmaxf(Lf - r.lbefore, 0.0f);
    float li = r.thr > 0.0f ? gathered / r.thr : 0.0f;
    if (!(li < 1e6f)) li = 0.0f;
    const float c = fmaxf(r.nrm[0] * r.dir[0] + r.nrm[1] * r.dir[1] + r.nrm[2] * r.dir[2], 0.0f);
    r.target = r.pdf > 0.0f ? li * c / r.pdf : 0.0f;
    if (!(r.target < 1e30f)) r.target = 0.0f;
    if (dense) dense[i] = r.target;
    acc += r.target;
  }
  // one atomic per block (per-wave atomics would all queue on wsum)
  __shared__ float s_part[4];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63u) == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
    if (t > 0.0f) atomicAdd(wsum, t);
  }
}

// Records drawn in proportion to their targets: inclusive prefix sums of the
// targets in 64-bit fixed point (target * 2^(60 - e) with 2^e above the
// batch's target sum, so sums stay below 2^62 and are exact in any order,
// hence monotone), in three passes: sums of 4096-record tiles, one block
// scanning the tile sums, then each tile scanned with its offset.  A training
// step finds record j for u in [cdf[j-1], cdf[j]) by binary search.
constexpr uint32_t kScanPer = 16;                   // records per thread
constexpr uint32_t kScanTile = 256 * kScanPer;      // records per 256-thread block

__device__ __forceinline__ float cdf_scale(float wsum) {
  int e;
  frexpf(fmaxf(wsum, 1e-30f), &e);  // wsum < 2^e
  return ldexpf(1.0f, 60 - e);
}

__device__ __forceinline__ uint64_t cdf_quant(float t, float scale) {
  return t > 0.0f ? (uint64_t)(t * scale) : 0ull;
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int o) {
  const uint32_t lo = __shfl_up((uint32_t)v, o), hi = __shfl_up((uint32_t)(v >> 32), o);
  return ((uint64_t)hi << 32) | lo;
}

// Exclusive scan over a 256-thread block; total receives the block's sum.
__device__ __forceinline__ uint64_t block_excl_scan256(uint64_t v, uint64_t *s_w,
                                                       uint64_t &total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = shfl_up64(x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  uint64_t before = 0;
  for (uint32_t k = 0; k < w; k++) before += s_w[k];
  total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();
  return before + x - v;
}

__global__ __launch_bounds__(256) void k_guide_tile_sums(const float *__restrict__ dense,
                                                         const uint32_t *__restrict__ counts,
                                                         uint32_t cap,
                                                         const float *__restrict__ wsum,
                                                         uint64_t *__restrict__ tsum) {
  __shared__ uint64_t s_w[4];
  const uint32_t n = min(counts[PT_CNT_TRAIN], cap);
  const uint32_t base = blockIdx.x * kScanTile;
  if (base >= n) return;
  const float sc = cdf_scale(wsum[0]);
  uint64_t v = 0;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += 256) {
    const uint32_t i = base + k;
    if (i < n) v += cdf_quant(dense[i], sc);
  }
  uint64_t total;
  block_excl_scan256(v, s_w, total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_guide_scan_sums(uint64_t *__restrict__ tsum,
                                                         const uint32_t *__restrict__ counts,
                                                         uint32_t cap) {
  __shared__ uint64_t s_w[4];
  const uint32_t n = min(counts[PT_CNT_TRAIN], cap);
  const uint32_t nt = (n + kScanTile - 1) / kScanTile;
  uint64_t carry = 0;
  for (uint32_t b = 0; b < nt; b += 256) {
    const uint32_t i = b + threadIdx.x;
    const uint64_t v = i < nt ? tsum[i] : 0ull;
    uint64_t total;
    const uint64_t ex = block_excl_scan256(v, s_w, total);
    if (i < nt) tsum[i] = carry + ex;  // exclusive offset of tile i
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_guide_tile_scan(const float *__restrict__ dense,
                                                         const uint64_t *__restrict__ tsum,
                                                         const uint32_t *__restrict__ counts,
                                                         uint32_t cap,
                                                         const float *__restrict__ wsum,
                                                         uint64_t *__restrict__ cdf) {
  __shared__ uint64_t s_w[4];
  const uint32_t n = min(counts[PT_CNT_TRAIN], cap);
  const uint32_t base = blockIdx.x * kScanTile;
  if (base >= n) return;
  const float sc = cdf_scale(wsum[0]);
  const uint32_t i0 = base + threadIdx.x * kScanPer;
  uint64_t run = 0;
  for (uint32_t k = 0; k < kScanPer; k++)
    if (i0 + k < n) run += cdf_quant(dense[i0 + k], sc);
  uint64_t total;
  uint64_t acc = block_excl_scan256(run, s_w, total) + tsum[blockIdx.x];
  for (uint32_t k = 0; k < kScanPer; k++)
    if (i0 + k < n) {
      acc += cdf_quant(dense[i0 + k], sc);
      cdf[i0 + k] = acc;
    }
}


// ---- register-resident inference ------------------------------------------
//
// Each layer is evaluated transposed, Z^T (units x samples) = W (units x K)
// X^T (K x samples), 16 samples per wave: the MFMA B operand of lane l holds
// 8 consecutive inputs (k = 8 (l/16) .. +7) of sample l % 16 and its D
// result holds units 4 (l/16) + r (r = 0..3) of each 16-unit tile for the
// same sample.  Choosing the next layer's column order so that input k of
// block kb is unit 32 kb + 16 (j/4) + 4 (l/16) + j%4 (j = k % 8) turns the D
// registers of tiles 2kb and 2kb+1 directly into that block's B operand:
// activations never leave the registers.  The output layer's rows are
// ordered so that lane group g holds lobes kLpg g .. kLpg g + kLpg - 1
// (kSlots outputs).
// The permuted bf16 copy of the weights lives at w16 + GD_NW16.

// Position of hidden unit u in the permuted column order.
__host__ __device__ __forceinline__ uint32_t perm_col(uint32_t u) {
  return 32u * (u >> 5) + 8u * ((u & 15u) >> 2) + 4u * ((u & 31u) >> 4) + (u & 3u);
}

// Permuted row of output o (o < GD_MIX used): lane group g = o / kSlots,
// slot s = o % kSlots in tile s / 4, register s % 4.
__host__ __device__ __forceinline__ uint32_t perm_out_row(uint32_t o) {
  const uint32_t g = o / kSlots, sl = o % kSlots;
  return 16u * (sl >> 2) + 4u * g + (sl & 3u);
}

// Index of master parameter i in the permuted mirror, or -1.
__device__ __forceinline__ int perm_dst(uint32_t i) {
  if (i >= GD_OFS_W1 && i < GD_OFS_B1) return (int)(i - GD_OFS_W1);
  if (i >= GD_OFS_W2 && i < GD_OFS_B2) {
    const uint32_t k = i - GD_OFS_W2;
    return (int)(GD_HID * GD_IN + (k / GD_HID) * GD_HID + perm_col(k % GD_HID));
  }
  if (i >= GD_OFS_W3 && i < GD_OFS_B3) {
    const uint32_t k = i - GD_OFS_W3;
    return (int)(GD_HID * GD_IN + GD_HID * GD_HID + (k / GD_HID) * GD_HID + perm_col(k % GD_HID));
  }
  if (i >= GD_OFS_W4 && i < GD_OFS_B4) {
    const uint32_t k = i - GD_OFS_W4, o = k / GD_HID;
    if (o >= GD_MIX) return -1;
    return (int)(GD_HID * GD_IN + 2 * GD_HID * GD_HID + perm_out_row(o) * GD_HID +
                 perm_col(k % GD_HID));
  }
  return -1;
}

// bias + ReLU of tiles 2kb, 2kb+1 packed as the next layer's B operand.
__device__ __forceinline__ bf16x8 relu_pack(const f32x4 &d0, const f32x4 &d1, const float *b,
                                            uint32_t t0, uint32_t g) {
  bf16x8 x;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    x[r] = (__bf16)fmaxf(d0[r] + b[16u * t0 + 4u * g + (uint32_t)r], 0.0f);
    x[4 + r] = (__bf16)fmaxf(d1[r] + b[16u * (t0 + 1u) + 4u * g + (uint32_t)r], 0.0f);
  }
  return x;
}

// One hidden layer with K = 64: 4 output tiles from two B blocks.
__device__ __forceinline__ void layer64(const __bf16 *__restrict__ W, const bf16x8 xb[2],
                                        uint32_t s, uint32_t g, f32x4 h[4]) {
#pragma unroll
  for (int t = 0; t < 4; t++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      const bf16x8 av = *(const bf16x8 *)(W + (16u * t + s) * GD_HID + 32u * kb + 8u * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, xb[kb], acc, 0, 0, 0);
    }
    h[t] = acc;
  }
}

// ---- training kernel ------------------------------------------------------
//
// Same register layout as k_guide_mix: a wave owns 16 samples, lane l holds
// sample s = l % 16 and k-group g = l / 16; activations stay in MFMA D/B
// fragments.  Backward products W^T dY use a third bf16 mirror of the
// weights, transposed and with the output units in the permuted K order
// (w16 + 2 GD_NW16: W1T [32][64], W2T [64][64], W3T [64][64], W4T [64][64]).
// Weight gradients sum over the wave's 16 samples with
// v_mfma_f32_16x16x16_bf16 after a 16 x 16 transpose through LDS, and are
// accumulated per block in LDS (fp32) before one global flush.

__device__ __forceinline__ int permT_dst(uint32_t i) {
  if (i >= GD_OFS_W1 && i < GD_OFS_B1) {
    const uint32_t k = i - GD_OFS_W1;
    return (int)((k % GD_IN) * GD_HID + perm_col(k / GD_IN));
  }
  if (i >= GD_OFS_W2 && i < GD_OFS_B2) {
    const uint32_t k = i - GD_OFS_W2;
    return (int)(GD_IN * GD_HID + (k % GD_HID) * GD_HID + perm_col(k / GD_HID));
  }
  if (i >= GD_OFS_W3 && i < GD_OFS_B3) {
    const uint32_t k = i - GD_OFS_W3;
    return (int)(GD_IN * GD_HID + GD_HID * GD_HID + (k % GD_HID) * GD_HID + perm_col(k / GD_HID));
  }
  if (i >= GD_OFS_W4 && i < GD_OFS_B4) {
    const uint32_t k = i - GD_OFS_W4, o = k / GD_HID;
    if (o >= GD_MIX) return -1;
    return (int)(GD_IN * GD_HID + 2 * GD_HID * GD_HID + (k % GD_HID) * GD_HID +
                 perm_col(perm_out_row(o)));
  }
  return -1;
}

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short s16x4;
typedef short __attribute__((ext_vector_type(2))) s16x2;
typedef __attribute__((address_space(1))) s16x2 gs16x2;

// Hash-grid gradients: the two features of an entry are one word of two
// bf16 values at word l * GD_TABLE + idx of the gradient buffer's grid part,
// added with one packed atomic (half the atomic requests of two fp32 adds;
// the requests, which the memory-side atomic units serve, bound a training
// step).  k_guide_adam reads and clears them.
__device__ __forceinline__ void grid_grad_add(float *grad, size_t entry, float v0, float v1) {
  s16x2 v;
  v.x = __builtin_bit_cast(short, (__bf16)v0);
  v.y = __builtin_bit_cast(short, (__bf16)v1);
  __builtin_amdgcn_global_atomic_fadd_v2bf16((gs16x2 *)((uint32_t *)grad + entry), v);
}

// One wave per training block: the hash-grid gradient atomics of a step
// (the bulk of its time, latency-bound per CU) spread over as many CUs as
// there are 16-sample tiles, and the block's weight gradients need no LDS
// atomics.  Each block leaves its weight/bias gradients and loss in its own
// slot of a partial-sum buffer, summed by k_guide_reduce in a fixed order.
constexpr int kTrainWaves = 1;
constexpr uint32_t kTrainMaxBlocks = 512;
constexpr uint32_t kDwW1 = 0, kDwW2 = GD_HID * GD_IN, kDwW3 = kDwW2 + GD_HID * GD_HID,
                   kDwW4 = kDwW3 + GD_HID * GD_HID, kDwN = kDwW4 + GD_MIX * GD_HID;
constexpr uint32_t kDbB1 = 0, kDbB2 = GD_HID, kDbB3 = 2 * GD_HID, kDbB4 = 3 * GD_HID,
                   kDbN = 3 * GD_HID + GD_MIX;
// partial-sum slot of a block: dW, db, loss (floats, padded to 64)
constexpr uint32_t kPartN = kDwN + kDbN + 1, kPartStride = (kPartN + 63u) & ~63u;

struct TrainLds {
  float dw[kDwN];                                        // weight gradients (fp32)
  float db[kDbN];                                        // bias gradients
  __attribute__((aligned(16))) __bf16 ty[kTrainWaves][GD_HID * 16];  // dY rows x 16 samples
  __attribute__((aligned(16))) __bf16 tx[kTrainWaves][GD_HID * 16];  // X rows x 16 samples
};

// Orders the wave's LDS accesses (lanes of one wave exchange data).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hidden unit of K position 32 kb + 8 g + j of a packed B fragment.
__device__ __forceinline__ uint32_t kpos_unit(uint32_t kb, uint32_t g, uint32_t j) {
  return 16u * (2u * kb + (j >> 2)) + 4u * g + (j & 3u);
}

// D fragments of tiles 2 kb, 2 kb + 1 packed as a B operand (no bias).
__device__ __forceinline__ bf16x8 pack_b(const f32x4 &d0, const f32x4 &d1) {
  bf16x8 x;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    x[r] = (__bf16)d0[r];
    x[4 + r] = (__bf16)d1[r];
  }
  return x;
}

// X rows (units of a packed B operand pair) into the wave's transpose tile.
__device__ __forceinline__ void put_x(__bf16 *tx, const bf16x8 xb[2], uint32_t s, uint32_t g) {
#pragma unroll
  for (uint32_t kb = 0; kb < 2; kb++)
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) tx[kpos_unit(kb, g, j) * 16u + s] = xb[kb][j];
}

// dY rows (D fragments of 4 tiles) into the wave's transpose tile.
__device__ __forceinline__ void put_y(__bf16 *ty, const f32x4 d[4], uint32_t s, uint32_t g) {
#pragma unroll
  for (uint32_t t = 0; t < 4; t++)
#pragma unroll
    for (uint32_t r = 0; r < 4; r++) ty[(16u * t + 4u * g + r) * 16u + s] = (__bf16)d[t][r];
}

// dW[n][k] += sum over the 16 samples of Y[n][s] X[k][s] (LDS accumulator
// of the block's one wave, row stride ldk; rows n >= nvalid are skipped).
// The lanes of one MFMA result hold distinct (n, k), so plain read-add-write
// is exact.
__device__ __forceinline__ void dw_accum(const __bf16 *ty, const __bf16 *tx, int nrows,
                                         int nvalid, int krows, float *dw, int ldk) {
  const uint32_t l = lane_id(), c = l & 15u, g = l >> 4;
  for (int nt = 0; nt < nrows; nt += 16) {
    const s16x4 av = *(const s16x4 *)(ty + (nt + (int)c) * 16 + 4 * (int)g);
    for (int kt = 0; kt < krows; kt += 16) {
      const s16x4 bv = *(const s16x4 *)(tx + (kt + (int)c) * 16 + 4 * (int)g);
      const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
      const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, z, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int n = nt + 4 * (int)g + r;
        if (n < nvalid) dw[n * ldk + kt + (int)c] += d[r];
      }
    }
  }
}

// Bias gradient: sum of v over the 16 samples of the lane's k-group.
__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

// Mask by the ReLU of the forward activations xb (packed B layout) and
// return as D fragments + packed B operands for the next product.
__device__ __forceinline__ void relu_back(f32x4 d[4], const bf16x8 xb[2]) {
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (!((float)xb[t >> 1][4 * (t & 1) + r] > 0.0f)) d[t][r] = 0.0f;
}

// D (4 tiles of 16 units) = W^T dY for a 64-unit K with the transposed mirror.
__device__ __forceinline__ void back64(const __bf16 *__restrict__ WT, const bf16x8 yb[2],
                                       uint32_t s, uint32_t g, int ntiles, f32x4 *d) {
  for (int t = 0; t < ntiles; t++) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      const bf16x8 av = *(const bf16x8 *)(WT + (16u * t + s) * GD_HID + 32u * kb + 8u * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, yb[kb], acc, 0, 0, 0);
    }
    d[t] = acc;
  }
}

// Forward + backward over n = min(chunk, records) samples: gradients of the
// mean of -(target / mean target) log q(dir) (uniformly drawn records) or of
// the mean of -log q(dir) (records drawn in proportion to target) into grad;
// q is the vMF mixture p, or with a.train_bsdf = b > 0 the density
// (1 - b) p + b cos/pi that one-sample MIS with the BSDF samples.
__global__ __launch_bounds__(64) void k_guide_train(PtArgs a, const __half *__restrict__ grid,
                                                     const __bf16 *__restrict__ wp,
                                                     const __bf16 *__restrict__ wt,
                                                     const float *__restrict__ master,
                                                     const GdRecord *__restrict__ rec,
                                                     const uint32_t *__restrict__ counts,
                                                     uint32_t cap, uint32_t first,
                                                     uint32_t chunk,
                                                     const float *__restrict__ wsum,
                                                     const uint64_t *__restrict__ cdf,
                                                     float *__restrict__ grad,
                                                     float *__restrict__ part) {
  __shared__ TrainLds S;
  for (uint32_t k = threadIdx.x; k < kDwN; k += blockDim.x) S.dw[k] = 0.0f;
  for (uint32_t k = threadIdx.x; k < kDbN; k += blockDim.x) S.db[k] = 0.0f;
  __syncthreads();
  const uint32_t lane = lane_id(), s = lane & 15u, g = lane >> 4;
  const uint32_t wv = threadIdx.x >> 6;
  __bf16 *ty = S.ty[wv], *tx = S.tx[wv];
  const uint32_t total = counts ? min(counts[PT_CNT_TRAIN], cap) : cap;
  const uint32_t n = min(chunk, total);
  const float wmean = wsum[0] / (float)(total ? total : 1u);
  const float inv = wmean > 0.0f ? 1.0f / (wmean * (float)(n ? n : 1u)) : 0.0f;
  const uint64_t ctot = (cdf && total) ? cdf[total - 1] : 0ull;
  const double cstep = (double)ctot / (double)(n ? n : 1u);
  const double xi = (double)(first >> 8) * (1.0 / 16777216.0);
  float lacc = 0.0f;
  const uint32_t ntiles = (n + 15u) / 16u;
  const uint32_t wave = blockIdx.x * kTrainWaves + wv, nw = gridDim.x * kTrainWaves;
  for (uint32_t tile = wave; tile < ntiles; tile += nw) {
    // weight fragments are re-read per tile (from the caches); the opaque
    // zero offset stops the compiler from hoisting all of them into
    // registers
    uint32_t wofs = 0;
    asm volatile("" : "+s"(wofs));
    const __bf16 *W1 = wp + wofs, *W2 = W1 + GD_HID * GD_IN, *W3 = W2 + GD_HID * GD_HID,
                 *W4 = W3 + GD_HID * GD_HID;
    const __bf16 *W1T = wt + wofs, *W2T = W1T + GD_IN * GD_HID, *W3T = W2T + GD_HID * GD_HID,
                 *W4T = W3T + GD_HID * GD_HID;
    const float *mp = master + wofs;
    const uint32_t i = tile * 16u + s;
    bool valid = i < n;
    GdRecord r;
    float wgt = 0.0f;
    if (valid && cdf) {
      // systematic draw in proportion to target (offset xi in each stratum)
      const uint64_t u = (uint64_t)(((double)i + xi) * cstep);
      uint32_t lo = 0, hi = total;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1; else hi = mid;
      }
      r = rec[lo < total ? lo : total - 1];
      valid = ctot > 0 && r.target > 0.0f;
      wgt = valid ? 1.0f / (float)n : 0.0f;
    } else if (valid) {
      r = rec[((uint64_t)i * 2654435761ull + first) % total];
      wgt = r.target * inv;
    }
    if (!valid || !(wgt > 0.0f) || !(wgt < 1e30f)) {
      for (int k = 0; k < 3; k++) { r.pos[k] = 0.0f; r.nrm[k] = 0.0f; r.dir[k] = 0.0f; }
      r.nrm[2] = 1.0f; r.dir[2] = 1.0f;
      wgt = 0.0f;
      valid = false;
    }
    float q[3];
    normalise_pos(a, r.pos, q);
    // ---- forward (as k_guide_mix), keeping every layer's input
    bf16x8 x0;
    if (g < 3u) {
#pragma unroll
      for (int lv = 0; lv < 4; lv++) {
        const int l = 4 * (int)g + lv;
        float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; c++) {
          uint32_t idx;
          float w;
          grid_corner(l, q, c, idx, w);
          const __half2 gv = *(const __half2 *)&grid[((size_t)l * GD_TABLE + idx) * GD_FEAT];
          f0 += w * __low2float(gv);
          f1 += w * __high2float(gv);
        }
        x0[2 * lv] = (__bf16)f0;
        x0[2 * lv + 1] = (__bf16)f1;
      }
    } else {
      x0[0] = (__bf16)r.nrm[0];
      x0[1] = (__bf16)r.nrm[1];
      x0[2] = (__bf16)r.nrm[2];
      x0[3] = (__bf16)1.0f;
      x0[4] = x0[5] = x0[6] = x0[7] = (__bf16)0.0f;
    }
    f32x4 h[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const bf16x8 av = *(const bf16x8 *)(W1 + (16u * t + s) * GD_IN + 8u * g);
      const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
      h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, x0, z, 0, 0, 0);
    }
    bf16x8 x1[2], x2[2], x3[2];
    x1[0] = relu_pack(h[0], h[1], mp + GD_OFS_B1, 0u, g);
    x1[1] = relu_pack(h[2], h[3], mp + GD_OFS_B1, 2u, g);
    layer64(W2, x1, s, g, h);
    x2[0] = relu_pack(h[0], h[1], mp + GD_OFS_B2, 0u, g);
    x2[1] = relu_pack(h[2], h[3], mp + GD_OFS_B2, 2u, g);
    layer64(W3, x2, s, g, h);
    x3[0] = relu_pack(h[0], h[1], mp + GD_OFS_B3, 0u, g);
    x3[1] = relu_pack(h[2], h[3], mp + GD_OFS_B3, 2u, g);
    float o[4 * kOutTiles];
#pragma unroll
    for (int t = 0; t < (int)kOutTiles; t++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {
        const bf16x8 av = *(const bf16x8 *)(W4 + (16u * t + s) * GD_HID + 32u * kb + 8u * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, x3[kb], acc, 0, 0, 0);
      }
#pragma unroll
      for (int rr = 0; rr < 4; rr++) o[4 * t + rr] = acc[rr];
    }
#pragma unroll
    for (int sl = 0; sl < (int)kSlots; sl++) o[sl] += mp[GD_OFS_B4 + kSlots * g + (uint32_t)sl];
    // ---- loss gradient for this lane's lobes kLpg g .. (slots 5 lb .. 5 lb + 4)
    float mx = o[0];
#pragma unroll
    for (int lb = 1; lb < (int)kLpg; lb++) mx = fmaxf(mx, o[5 * lb]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float ex[kLpg], esum = 0.0f;
#pragma unroll
    for (int lb = 0; lb < (int)kLpg; lb++) {
      ex[lb] = expf(o[5 * lb] - mx);
      esum += ex[lb];
    }
    esum += __shfl_xor(esum, 16);
    esum += __shfl_xor(esum, 32);
    float pi[kLpg], mu[kLpg][3], kap[kLpg], cth[kLpg], vv[kLpg], il[kLpg];
    bool dirok[kLpg];
#pragma unroll
    for (int lb = 0; lb < (int)kLpg; lb++) {
      const float *ol = &o[5 * lb];
      const float x = ol[1], y = ol[2], z = ol[3];
      const float l2 = x * x + y * y + z * z;
      dirok[lb] = l2 >= 1e-12f;
      il[lb] = dirok[lb] ? 1.0f / sqrtf(l2) : 0.0f;
      mu[lb][0] = dirok[lb] ? x * il[lb] : 0.0f;
      mu[lb][1] = dirok[lb] ? y * il[lb] : 0.0f;
      mu[lb][2] = dirok[lb] ? z * il[lb] : 1.0f;
      kap[lb] = expf(fminf(fmaxf(ol[4], GD_LOGK_MIN), GD_LOGK_MAX));
      pi[lb] = ex[lb] / esum;
      cth[lb] = mu[lb][0] * r.dir[0] + mu[lb][1] * r.dir[1] + mu[lb][2] * r.dir[2];
      vv[lb] = gd_vmf(kap[lb], cth[lb]);
    }
    float pd = 0.0f;
#pragma unroll
    for (int lb = 0; lb < (int)kLpg; lb++) pd += pi[lb] * vv[lb];
    pd += __shfl_xor(pd, 16);
    pd += __shfl_xor(pd, 32);
    pd = fmaxf(pd, 1e-20f);
    // loss -w log q with q = (1 - b) pd + b cos/pi: its gradient is that of
    // -w log pd scaled by (1 - b) pd / q
    float wl = wgt, qd = pd;
    if (a.train_bsdf > 0.0f) {
      const float cb = fmaxf(r.nrm[0] * r.dir[0] + r.nrm[1] * r.dir[1] + r.nrm[2] * r.dir[2],
                             0.0f) * 0.318309886f;
      qd = fmaxf((1.0f - a.train_bsdf) * pd + a.train_bsdf * cb, 1e-20f);
      wl = wgt * ((1.0f - a.train_bsdf) * pd / qd);
    }
    float go[4 * kOutTiles];
#pragma unroll
    for (int k = 0; k < 4 * (int)kOutTiles; k++) go[k] = 0.0f;
    if (wgt > 0.0f) {
#pragma unroll
      for (int lb = 0; lb < (int)kLpg; lb++) {
        const float rsp = pi[lb] * vv[lb] / pd;  // responsibility
        go[5 * lb] = -wl * (rsp - pi[lb]);
        if (dirok[lb]) {
          const float sc = -wl * rsp * kap[lb] * il[lb];
#pragma unroll
          for (int k = 0; k < 3; k++) go[5 * lb + 1 + k] = sc * (r.dir[k] - cth[lb] * mu[lb][k]);
        }
        const float lk = o[5 * lb + 4];
        if (lk > GD_LOGK_MIN && lk < GD_LOGK_MAX) {
          const float dl = 1.0f / kap[lb] - 2.0f / expm1f(2.0f * kap[lb]) + (cth[lb] - 1.0f);
          go[5 * lb + 4] = -wl * rsp * kap[lb] * dl;
        }
      }
      if (g == 0) lacc -= wgt * logf(qd);
    }
    // ---- output layer: dW4, db4, dX3
#pragma unroll
    for (int sl = 0; sl < (int)kSlots; sl++)
      ty[(kSlots * g + (uint32_t)sl) * 16u + s] = (__bf16)go[sl];
    put_x(tx, x3, s, g);
    wave_lds_sync();
    dw_accum(ty, tx, 16 * (int)kOutTiles, GD_MIX, GD_HID, S.dw + kDwW4, GD_HID);
    wave_lds_sync();
#pragma unroll
    for (int sl = 0; sl < (int)kSlots; sl++) {
      const float v = sum16(go[sl]);
      if (s == 0) S.db[kDbB4 + kSlots * g + (uint32_t)sl] += v;
    }
    bf16x8 yb[2];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      yb[0][j] = (__bf16)go[j];
      yb[1][j] = (__bf16)(j < (int)kSlots - 8 ? go[8 + j] : 0.0f);
    }
    f32x4 d[4];
    back64(W4T, yb, s, g, 4, d);
    relu_back(d, x3);
    // ---- hidden layers 3 and 2: dW, db, dX
#pragma unroll
    for (int layer = 3; layer >= 2; layer--) {
      const bf16x8 *xin = layer == 3 ? x2 : x1;
      const bf16x8 *xmask = layer == 3 ? x2 : x1;
      put_y(ty, d, s, g);
      put_x(tx, xin, s, g);
      wave_lds_sync();
      dw_accum(ty, tx, GD_HID, GD_HID, GD_HID, S.dw + (layer == 3 ? kDwW3 : kDwW2), GD_HID);
      wave_lds_sync();
      const uint32_t bo = layer == 3 ? kDbB3 : kDbB2;
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const float v = sum16(d[t][rr]);
          if (s == 0) S.db[bo + 16u * t + 4u * g + rr] += v;
        }
      yb[0] = pack_b(d[0], d[1]);
      yb[1] = pack_b(d[2], d[3]);
      back64(layer == 3 ? W3T : W2T, yb, s, g, 4, d);
      relu_back(d, xmask);
    }
    // ---- layer 1: dW1, db1, dX0 (inputs; the first 24 are grid features)
    put_y(ty, d, s, g);
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) tx[(8u * g + j) * 16u + s] = x0[j];
    wave_lds_sync();
    dw_accum(ty, tx, GD_HID, GD_HID, GD_IN, S.dw + kDwW1, GD_IN);
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const float v = sum16(d[t][rr]);
        if (s == 0) S.db[kDbB1 + 16u * t + 4u * g + rr] += v;
      }
    yb[0] = pack_b(d[0], d[1]);
    yb[1] = pack_b(d[2], d[3]);
    f32x4 dx[2];
    back64(W1T, yb, s, g, 2, dx);
    // input rows 16 t + 4 g + rr: pairs (rr 0,1) and (rr 2,3) are the two
    // features of level (16 t + 4 g + rr) / 2.  The 16 lanes of a k-group
    // hold the same levels for 16 samples; where all of them (invalid
    // samples count as agreeing) fall in one cell of a level -- the common
    // case on coarse levels, where per-sample atomics would all contend for
    // the same few addresses -- their corner gradients are summed across
    // the group and added once.
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int pr = 0; pr < 2; pr++) {
        const uint32_t jin = 16u * t + 4u * g + 2u * pr;
        if (jin >= 2u * GD_LEVELS) continue;  // k-group uniform
        const float d0 = valid ? dx[t][2 * pr] : 0.0f, d1 = valid ? dx[t][2 * pr + 1] : 0.0f;
        const int l = (int)(jin >> 1);
        const float res = c_gd_res[l];
        const uint32_t r1 = (uint32_t)res + 1u;
        const uint32_t key = (uint32_t)floorf(q[0] * res) +
                             r1 * ((uint32_t)floorf(q[1] * res) + r1 * (uint32_t)floorf(q[2] * res));
        const uint32_t key0 = __shfl(key, (int)(lane & 48u));
        const uint64_t agree = __ballot(!valid || key == key0);
        const bool shared = ((agree >> (lane & 48u)) & 0xffffull) == 0xffffull;
        for (int c = 0; c < 8; c++) {
          uint32_t idx;
          float w;
          grid_corner(l, q, c, idx, w);
          if (shared) {
            const float v0 = sum16(w * d0), v1 = sum16(w * d1);
            if (s == 0 && (v0 != 0.0f || v1 != 0.0f)) {
              // lane 16 g holds a valid sample's corner unless the whole
              // group is invalid (then v0 = v1 = 0)
              grid_grad_add(grad, (size_t)l * GD_TABLE + idx, v0, v1);
            }
          } else if (d0 != 0.0f || d1 != 0.0f) {
            grid_grad_add(grad, (size_t)l * GD_TABLE + idx, w * d0, w * d1);
          }
        }
      }
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) lacc += __shfl_down(lacc, o2);
  __syncthreads();
  // the block's weight and bias gradients and loss into its partial slot
  float *__restrict__ slot = part + (size_t)blockIdx.x * kPartStride;
  for (uint32_t k = threadIdx.x; k < kDwN + kDbN; k += blockDim.x)
    slot[k] = k < kDwN ? S.dw[k] : S.db[k - kDwN];
  if (threadIdx.x == 0) slot[kDwN + kDbN] = lacc;
}

// Sums the nb training blocks' partial slots (block order, so the MLP
// gradients of a step do not depend on scheduling) into the MLP part of
// grad, which it overwrites, and adds the loss to *loss.
// A 256-thread block sums 16 entries: 16 groups of 16 threads each add
// every 16th slot, then the group sums are added in group order.
__global__ __launch_bounds__(256) void k_guide_reduce(const float *__restrict__ part, uint32_t nb,
                                                      float *__restrict__ grad,
                                                      float *__restrict__ loss) {
  __shared__ float acc[16][17];
  const uint32_t c = threadIdx.x & 15u, grp = threadIdx.x >> 4;
  const uint32_t k = blockIdx.x * 16u + c;
  float v = 0.0f;
  if (k < kPartN) {
    uint32_t b = grp;
    for (; b + 48u < nb; b += 64u) {
      const float v0 = part[(size_t)b * kPartStride + k];
      const float v1 = part[(size_t)(b + 16u) * kPartStride + k];
      const float v2 = part[(size_t)(b + 32u) * kPartStride + k];
      const float v3 = part[(size_t)(b + 48u) * kPartStride + k];
      v += v0;
      v += v1;
      v += v2;
      v += v3;
    }
    for (; b < nb; b += 16u) v += part[(size_t)b * kPartStride + k];
  }
  acc[grp][c] = v;
  __syncthreads();
  if (grp != 0 || k >= kPartN) return;
  v = 0.0f;
  for (uint32_t j = 0; j < 16u; j++) v += acc[j][c];
  if (k == kDwN + kDbN) {
    if (loss) loss[0] += v;
    return;
  }
  uint32_t dst;
  if (k < kDwW2) dst = GD_OFS_W1 + k;
  else if (k < kDwW3) dst = GD_OFS_W2 + (k - kDwW2);
  else if (k < kDwW4) dst = GD_OFS_W3 + (k - kDwW3);
  else if (k < kDwN) dst = GD_OFS_W4 + (k - kDwW4);
  else if (k < kDwN + kDbB2) dst = GD_OFS_B1 + (k - kDwN);
  else if (k < kDwN + kDbB3) dst = GD_OFS_B2 + (k - kDwN - kDbB2);
  else if (k < kDwN + kDbB4) dst = GD_OFS_B3 + (k - kDwN - kDbB3);
  else dst = GD_OFS_B4 + (k - kDwN - kDbB4);
  grad[dst] = v;
}

// Mixture records (GD_MIX floats) of queued hit points (pn == nullptr) or of
// n position/normal pairs pn[6 i ..].
__global__ __launch_bounds__(256, 4) void k_guide_mix(PtArgs a, const __half *__restrict__ grid,
                                                   const __bf16 *__restrict__ wp,
                                                   const float *__restrict__ master,
                                                   const rray *__restrict__ rays,
                                                   const rhit *__restrict__ hits,
                                                   const uint32_t *__restrict__ counts,
                                                   const dinst *__restrict__ insts,
                                                   const RTri *__restrict__ tris,
                                                   const float *__restrict__ pn, uint32_t npn,
                                                   uint32_t share, float *__restrict__ mix) {
  // queue entries 0, share, 2 share, ... -> records 0, 1, 2, ...
  const uint32_t nq = pn ? npn : counts[PT_CNT_CUR];
  const uint32_t n = (nq + share - 1u) / share;
  const uint32_t lane = lane_id();
  const uint32_t s = lane & 15u, g = lane >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t base = wave * 16u; base < n; base += nwaves * 16u) {
    // The weight fragments are loop invariant; an opaque zero offset keeps
    // the compiler from hoisting all 26 of them into registers (they are
    // re-read from the caches per 16 samples instead of spilling).
    uint32_t wofs = 0;
    asm volatile("" : "+s"(wofs));
    const __bf16 *W1 = wp + wofs, *W2 = W1 + GD_HID * GD_IN, *W3 = W2 + GD_HID * GD_HID,
                 *W4 = W3 + GD_HID * GD_HID;
    const uint32_t i = base + s;
    float p[3] = {0, 0, 0}, nn[3] = {0, 0, 1};
    bool ok = false;
    if (i < n) {
      if (pn) {
        for (int k = 0; k < 3; k++) { p[k] = pn[6 * i + k]; nn[k] = pn[6 * i + 3 + k]; }
        ok = true;
      } else {
        const uint32_t qi = i * share;
        ok = hit_point(rays[qi], hits[qi], insts, tris, p, nn);
      }
    }
    // layer-1 B operand: inputs 8g .. 8g+7 of sample s
    bf16x8 x0;
    if (g < 3u) {
      float q[3];
      normalise_pos(a, p, q);
#pragma unroll
      for (int lv = 0; lv < 4; lv++) {
        const int l = 4 * (int)g + lv;
        float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; c++) {
          uint32_t idx;
          float w;
          grid_corner(l, q, c, idx, w);
          const __half2 gv = *(const __half2 *)&grid[((size_t)l * GD_TABLE + idx) * GD_FEAT];
          f0 += w * __low2float(gv);
          f1 += w * __high2float(gv);
        }
        x0[2 * lv] = (__bf16)f0;
        x0[2 * lv + 1] = (__bf16)f1;
      }
    } else {
      x0[0] = (__bf16)nn[0];
      x0[1] = (__bf16)nn[1];
      x0[2] = (__bf16)nn[2];
      x0[3] = (__bf16)1.0f;
      x0[4] = x0[5] = x0[6] = x0[7] = (__bf16)0.0f;
    }
    // layer 1 (K = 32)
    f32x4 h[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const bf16x8 av = *(const bf16x8 *)(W1 + (16u * t + s) * GD_IN + 8u * g);
      const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
      h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, x0, z, 0, 0, 0);
    }
    bf16x8 xb[2];
    xb[0] = relu_pack(h[0], h[1], master + GD_OFS_B1, 0u, g);
    xb[1] = relu_pack(h[2], h[3], master + GD_OFS_B1, 2u, g);
    layer64(W2, xb, s, g, h);
    xb[0] = relu_pack(h[0], h[1], master + GD_OFS_B2, 0u, g);
    xb[1] = relu_pack(h[2], h[3], master + GD_OFS_B2, 2u, g);
    layer64(W3, xb, s, g, h);
    xb[0] = relu_pack(h[0], h[1], master + GD_OFS_B3, 0u, g);
    xb[1] = relu_pack(h[2], h[3], master + GD_OFS_B3, 2u, g);
    // output layer: tiles 0..kOutTiles-1 hold this lane group's kSlots outputs
    float o[4 * kOutTiles];
#pragma unroll
    for (int t = 0; t < (int)kOutTiles; t++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {
        const bf16x8 av = *(const bf16x8 *)(W4 + (16u * t + s) * GD_HID + 32u * kb + 8u * g);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, xb[kb], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) o[4 * t + r] = acc[r];
    }
#pragma unroll
    for (int sl = 0; sl < (int)kSlots; sl++)
      o[sl] += master[GD_OFS_B4 + kSlots * g + (uint32_t)sl];
    // softmax over the lobes: the sample's 4 lanes are s, s+16, s+32, s+48
    float mx = o[0];
#pragma unroll
    for (int lb = 1; lb < (int)kLpg; lb++) mx = fmaxf(mx, o[5 * lb]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float ex[kLpg], sum = 0.0f;
#pragma unroll
    for (int lb = 0; lb < (int)kLpg; lb++) {
      ex[lb] = expf(o[5 * lb] - mx);
      sum += ex[lb];
    }
    sum += __shfl_xor(sum, 16);
    sum += __shfl_xor(sum, 32);
    float rec[kSlots];
#pragma unroll
    for (int lb = 0; lb < (int)kLpg; lb++) {
      const float *ol = &o[5 * lb];
      float x = ol[1], y = ol[2], z = ol[3];
      const float l2 = x * x + y * y + z * z;
      if (l2 < 1e-12f) { x = 0.0f; y = 0.0f; z = 1.0f; }
      else { const float il = 1.0f / sqrtf(l2); x *= il; y *= il; z *= il; }
      const float lk = fminf(fmaxf(ol[4], GD_LOGK_MIN), GD_LOGK_MAX);
      rec[5 * lb] = ex[lb] / sum;
      rec[5 * lb + 1] = x;
      rec[5 * lb + 2] = y;
      rec[5 * lb + 3] = z;
      rec[5 * lb + 4] = expf(lk);
    }
    if (!ok)  // empty proposal
      for (int lb = 0; lb < (int)kLpg; lb++) rec[5 * lb] = -1.0f;
    if (i < n) {
      float *dst = &mix[(size_t)i * GD_MIX + kSlots * g];
#pragma unroll
      for (int k = 0; k < (int)kSlots; k++) dst[k] = rec[k];
    }
  }
}

// Adam on the fp32 master parameters; refreshes the fp16/bf16 mirrors and
// clears the gradient.  The grid part of the gradient holds one bf16 pair
// per entry (grid_grad_add) in its first GD_N_GRID / 2 words; the MLP part
// is fp32.  Hash-grid features a step did not touch (gradient exactly zero)
// keep their parameters and moments: their master copy and fp16 mirror are
// neither read nor written, so a step costs one read of the packed grid
// gradient plus the touched entries.
__device__ __forceinline__ float adam_one(float *master, float *m1, float *m2, uint32_t i,
                                          float g, float lr, float c1, float c2) {
  const float b1 = 0.9f, b2 = 0.99f, eps = 1e-10f;
  const float a = b1 * m1[i] + (1.0f - b1) * g;
  const float b = b2 * m2[i] + (1.0f - b2) * g * g;
  m1[i] = a;
  m2[i] = b;
  const float p = master[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  master[i] = p;
  return p;
}

__global__ void k_guide_adam(float *__restrict__ master, float *__restrict__ grad,
                             float *__restrict__ m1, float *__restrict__ m2,
                             __half *__restrict__ grid16, __bf16 *__restrict__ w16,
                             uint32_t step, float lr_grid, float lr_mlp) {
  const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.99f, (float)step);
  constexpr uint32_t kWords = GD_N_GRID / 2u;
  uint32_t *__restrict__ gw = (uint32_t *)grad;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kWords + GD_NMLP;
       i += gridDim.x * blockDim.x) {
    if (i < kWords) {
      const uint32_t w = gw[i];
      if (!w) continue;
      gw[i] = 0u;
      const float g0 = __uint_as_float(w << 16), g1 = __uint_as_float(w & 0xffff0000u);
      if (g0 != 0.0f)
        grid16[2u * i] = __float2half(adam_one(master, m1, m2, 2u * i, g0, lr_grid, c1, c2));
      if (g1 != 0.0f)
        grid16[2u * i + 1u] =
            __float2half(adam_one(master, m1, m2, 2u * i + 1u, g1, lr_grid, c1, c2));
    } else {
      const uint32_t p = GD_N_GRID + (i - kWords);
      adam_one(master, m1, m2, p, grad[p], lr_mlp, c1, c2);
      grad[p] = 0.0f;
    }
  }
  // bf16 mirror of the weights (not biases), packed W1 W2 W3 W4
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < (uint32_t)GD_NMLP;
       j += gridDim.x * blockDim.x) {
    const uint32_t i = GD_N_GRID + j;
    int dst = -1;
    if (i >= GD_OFS_W1 && i < GD_OFS_B1) dst = (int)(i - GD_OFS_W1);
    else if (i >= GD_OFS_W2 && i < GD_OFS_B2) dst = GD_HID * GD_IN + (int)(i - GD_OFS_W2);
    else if (i >= GD_OFS_W3 && i < GD_OFS_B3) dst = GD_HID * GD_IN + GD_HID * GD_HID + (int)(i - GD_OFS_W3);
    else if (i >= GD_OFS_W4 && i < GD_OFS_B4) dst = GD_HID * GD_IN + 2 * GD_HID * GD_HID + (int)(i - GD_OFS_W4);
    if (dst >= 0) w16[dst] = (__bf16)master[i];
    const int pd = perm_dst(i);
    if (pd >= 0) w16[GD_NW16 + pd] = (__bf16)master[i];
    const int td = permT_dst(i);
    if (td >= 0) w16[2 * GD_NW16 + td] = (__bf16)master[i];
  }
}

// fp16 mirror of the whole hash grid from the master copy (after an upload).
__global__ void k_guide_grid_mirror(const float *__restrict__ master, __half *__restrict__ grid16) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < GD_N_GRID;
       i += gridDim.x * blockDim.x)
    grid16[i] = __float2half(master[i]);
}

// Standalone forward for tests: positions/normals (6 floats each) -> raw outputs.
__global__ __launch_bounds__(64) void k_guide_forward(PtArgs a, Net N, const float *__restrict__ pn,
                                                      uint32_t n, float *__restrict__ out) {
  __shared__ WaveLds S;
  const uint32_t lane = lane_id();
  for (uint32_t base = blockIdx.x * kWave; base < n; base += gridDim.x * kWave) {
    const uint32_t i = base + lane;
    float p[3] = {0, 0, 0}, nn[3] = {0, 0, 1};
    if (i < n) {
      for (int k = 0; k < 3; k++) { p[k] = pn[6 * i + k]; nn[k] = pn[6 * i + 3 + k]; }
    }
    encode(a, N.grid, p, nn, &S.act0[lane * GD_IN]);
    __syncthreads();
    forward(N, S.act0, S.act1, S.act2, S.act3, S.out);
    if (i < n)
      for (int k = 0; k < GD_OUT; k++) out[(size_t)i * GD_OUT + k] = S.out[lane * GD_OUT + k];
    __syncthreads();
  }
}

// MFMA lane-layout check for tests: D = A (16x32) * B (32x16), row-major.
__global__ void k_mfma_check(const float *A, const float *B, float *D) {
  __shared__ __attribute__((aligned(16))) __bf16 a[16 * 32];
  __shared__ __attribute__((aligned(16))) __bf16 b[32 * 16];
  const uint32_t l = lane_id();
  for (int k = l; k < 16 * 32; k += 64) { a[k] = (__bf16)A[k]; b[k] = (__bf16)B[k]; }
  __syncthreads();
  const f32x4 d = tile_mm(a, 32, 1, b, 16, 1, 0, 0, 32);
  for (int r = 0; r < 4; r++) D[(4 * (l >> 4) + r) * 16 + (l & 15u)] = d[r];
}

}  // namespace

// ---- C-ABI --------------------------------------------------------------------

extern "C" int rk_guide_setup(void *stream) {
  float res[GD_LEVELS];
  uint32_t dense[GD_LEVELS];
  const double b = exp((log((double)GD_NMAX) - log((double)GD_NMIN)) / (GD_LEVELS - 1));
  for (int l = 0; l < GD_LEVELS; l++) {
    res[l] = (float)floor(GD_NMIN * pow(b, l));
    const double r1 = res[l] + 1.0;
    dense[l] = r1 * r1 * r1 <= (double)GD_TABLE;
  }
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_gd_res), res, sizeof(res), 0,
                                        hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e == hipSuccess)
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_gd_dense), dense, sizeof(dense), 0,
                               hipMemcpyHostToDevice, (hipStream_t)stream);
  return (int)e;
}

static Net mknet(const void *grid16, const void *w16, const float *master) {
  Net N;
  N.grid = (const __half *)grid16;
  N.w = (const __bf16 *)w16;
  N.master = master;
  return N;
}

extern "C" int rk_guide_mix(const PtArgs *a, const void *grid16, const void *w16,
                            const float *master, const void *rays, const void *hits,
                            const unsigned *counts, const void *insts, const void *tris,
                            const float *pn, unsigned npn, unsigned share, float *mix,
                            unsigned grid, void *stream) {
  hipLaunchKernelGGL(k_guide_mix, dim3(grid), dim3(256), 0, (hipStream_t)stream, *a,
                     (const __half *)grid16, (const __bf16 *)w16 + GD_NW16, master,
                     (const rray *)rays, (const rhit *)hits, counts, (const dinst *)insts,
                     (const RTri *)tris, pn, npn, share ? share : 1u, mix);
  return (int)hipGetLastError();
}

extern "C" int rk_guide_targets(void *rec, const unsigned *counts, const void *paths,
                                float *wsum, float *dense, unsigned cap, unsigned mode,
                                void *stream) {
  hipLaunchKernelGGL(k_guide_targets, dim3(256), dim3(256), 0, (hipStream_t)stream,
                     (GdRecord *)rec, counts, (const PtPath *)paths, wsum, dense, cap, mode);
  return (int)hipGetLastError();
}

extern "C" unsigned rk_guide_cdf_tiles(unsigned cap) {
  return (cap + kScanTile - 1) / kScanTile;
}

// Fixed-point inclusive prefix sums cdf[i] of dense[0, min(counts[train],
// cap)) scaled by the sum wsum[0]; tsum holds rk_guide_cdf_tiles(cap) words.
extern "C" int rk_guide_cdf(const float *dense, const float *wsum, unsigned long long *tsum,
                            unsigned long long *cdf, const unsigned *counts, unsigned cap,
                            void *stream) {
  const unsigned nt = (cap + kScanTile - 1) / kScanTile;
  if (!nt) return 0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_guide_tile_sums, dim3(nt), dim3(256), 0, st, dense, counts, cap, wsum,
                     (uint64_t *)tsum);
  hipLaunchKernelGGL(k_guide_scan_sums, dim3(1), dim3(256), 0, st, (uint64_t *)tsum, counts,
                     cap);
  hipLaunchKernelGGL(k_guide_tile_scan, dim3(nt), dim3(256), 0, st, dense,
                     (const uint64_t *)tsum, counts, cap, wsum, (uint64_t *)cdf);
  return (int)hipGetLastError();
}

// `grid` bounds the number of 256-thread blocks (each holds 69 KB of LDS).
extern "C" unsigned long long rk_guide_part_bytes(void) {
  return (unsigned long long)kTrainMaxBlocks * kPartStride * sizeof(float);
}

// part: rk_guide_part_bytes() of device memory.
extern "C" int rk_guide_train(const PtArgs *a, const void *grid16, const void *w16,
                              const float *master, const void *rec, const unsigned *counts,
                              unsigned cap, unsigned first, unsigned chunk, const float *wsum,
                              const unsigned long long *cdf, float *grad, float *loss,
                              float *part, unsigned grid, void *stream) {
  if (!chunk) return 0;
  const unsigned need = (chunk + 16u * kTrainWaves - 1) / (16u * kTrainWaves);
  unsigned nb = grid < need ? (grid ? grid : 1u) : need;
  if (nb > kTrainMaxBlocks) nb = kTrainMaxBlocks;
  const __bf16 *w = (const __bf16 *)w16;
  hipLaunchKernelGGL(k_guide_train, dim3(nb), dim3(64 * kTrainWaves), 0, (hipStream_t)stream,
                     *a, (const __half *)grid16, w + GD_NW16, w + 2 * GD_NW16, master,
                     (const GdRecord *)rec, counts, cap, first, chunk, wsum,
                     (const uint64_t *)cdf, grad, part);
  hipLaunchKernelGGL(k_guide_reduce, dim3((kPartN + 15u) / 16u), dim3(256), 0,
                     (hipStream_t)stream, part, nb, grad, loss);
  return (int)hipGetLastError();
}

extern "C" int rk_guide_adam(float *master, float *grad, float *m1, float *m2, void *grid16,
                             void *w16, unsigned step, float lr_grid, float lr_mlp,
                             void *stream) {
  hipLaunchKernelGGL(k_guide_adam, dim3(4096), dim3(256), 0, (hipStream_t)stream, master, grad,
                     m1, m2, (__half *)grid16, (__bf16 *)w16, step, lr_grid, lr_mlp);
  return (int)hipGetLastError();
}

extern "C" int rk_guide_grid_mirror(const float *master, void *grid16, void *stream) {
  hipLaunchKernelGGL(k_guide_grid_mirror, dim3(4096), dim3(256), 0, (hipStream_t)stream, master,
                     (__half *)grid16);
  return (int)hipGetLastError();
}

extern "C" int rk_guide_forward(const PtArgs *a, const void *grid16, const void *w16,
                                const float *master, const float *pn, unsigned n, float *out,
                                void *stream) {
  if (!n) return 0;
  hipLaunchKernelGGL(k_guide_forward, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                     *a, mknet(grid16, w16, master), pn, n, out);
  return (int)hipGetLastError();
}

extern "C" int rk_mfma_check(const float *A, const float *B, float *D, void *stream) {
  hipLaunchKernelGGL(k_mfma_check, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
  return (int)hipGetLastError();
}

// The sampling routine of the guided vertices, evaluated on the host.
extern "C" int rend_hip_guide_sample_host(const float *mix, float u0, float u1, float u2,
                                          float dir[3]) {
  if (!mix || !dir) return -3;  // REND_EINVAL
  gd_sample(mix, u0, u1, u2, dir[0], dir[1], dir[2]);
  return 0;
}
