// von Mises-Fisher mixture used by the learned proposal: parameters from the
// network's raw outputs, density, sampling and the gradient of
// -w log p(dir) with respect to the raw outputs.
//
// Lobe j of the raw output vector o: o[5j] logit, o[5j+1..3] unnormalised mean
// direction, o[5j+4] log concentration (clamped to [GD_LOGK_MIN, GD_LOGK_MAX]).
// Mixture record (GD_MIX floats): per lobe weight, mean xyz, kappa.
// vMF density on S^2: k / (2 pi (1 - exp(-2k))) exp(k (mu.w - 1)).
#ifndef GUIDE_MATH_H
#define GUIDE_MATH_H

#include "../ptlayout.h"

#define GD_LOGK_MIN -2.0f
#define GD_LOGK_MAX 7.0f

__device__ __forceinline__ float gd_vmf(float k, float cosang) {
  // k / (2 pi (1 - e^{-2k})) written with expm1 for small k
  const float norm = k / (6.28318531f * -expm1f(-2.0f * k));
  return norm * expf(k * (cosang - 1.0f));
}

__device__ __forceinline__ void gd_mixture(const float *o, float *mix) {
  float mx = -3.0e38f;
  for (int j = 0; j < GD_LOBES; j++) mx = fmaxf(mx, o[5 * j]);
  float s = 0.0f;
  float e[GD_LOBES];
  for (int j = 0; j < GD_LOBES; j++) {
    e[j] = expf(o[5 * j] - mx);
    s += e[j];
  }
  for (int j = 0; j < GD_LOBES; j++) {
    float x = o[5 * j + 1], y = o[5 * j + 2], z = o[5 * j + 3];
    const float l2 = x * x + y * y + z * z;
    if (l2 < 1e-12f) { x = 0.0f; y = 0.0f; z = 1.0f; }
    else { const float il = 1.0f / sqrtf(l2); x *= il; y *= il; z *= il; }
    const float lk = fminf(fmaxf(o[5 * j + 4], GD_LOGK_MIN), GD_LOGK_MAX);
    mix[5 * j] = e[j] / s;
    mix[5 * j + 1] = x;
    mix[5 * j + 2] = y;
    mix[5 * j + 3] = z;
    mix[5 * j + 4] = expf(lk);
  }
}

__device__ __forceinline__ float gd_pdf(const float *mix, float dx, float dy, float dz) {
  float p = 0.0f;
  for (int j = 0; j < GD_LOBES; j++) {
    const float c = mix[5 * j + 1] * dx + mix[5 * j + 2] * dy + mix[5 * j + 3] * dz;
    p += mix[5 * j] * gd_vmf(mix[5 * j + 4], c);
  }
  return p;
}

// Lobe chosen by u0; direction about its mean (Jakob 2012).  u1 is in [0, 1),
// so v = 1 - u1 is in (0, 1] and log(v + (1 - v) e^{-2k}) is finite even when
// e^{-2k} underflows to 0 (k > ~52); v has the distribution of u1.
__host__ __device__ __forceinline__ void gd_sample(const float *mix, float u0, float u1,
                                                   float u2, float &dx, float &dy,
                                                   float &dz) {
  int j = 0;
  float acc = mix[0];
  while (j < GD_LOBES - 1 && u0 >= acc) {
    j++;
    acc += mix[5 * j];
  }
  const float mx = mix[5 * j + 1], my = mix[5 * j + 2], mz = mix[5 * j + 3];
  const float k = mix[5 * j + 4];
  // w = 1 + log(v + (1 - v) e^{-2k}) / k, v = 1 - u1 in (0, 1]
  const float v = 1.0f - u1;
  const float w = 1.0f + logf(v + (1.0f - v) * expf(-2.0f * k)) / k;
  const float sw = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float sn, cs;
  sincosf(6.28318531f * u2, &sn, &cs);
  const float sgn = copysignf(1.0f, mz);
  const float a = -1.0f / (sgn + mz);
  const float b = mx * my * a;
  const float t1x = 1.0f + sgn * mx * mx * a, t1y = sgn * b, t1z = -sgn * mx;
  const float t2x = b, t2y = sgn + my * my * a, t2z = -my;
  dx = w * mx + sw * (cs * t1x + sn * t2x);
  dy = w * my + sw * (cs * t1y + sn * t2y);
  dz = w * mz + sw * (cs * t1z + sn * t2z);
}

// g = d(-wgt log p(dir)) / d o ; returns -wgt log p (loss contribution).
__device__ __forceinline__ float gd_loss_grad(const float *o, const float *dir, float wgt,
                                              float *g) {
  for (int k = 0; k < GD_OUT; k++) g[k] = 0.0f;
  if (!(wgt > 0.0f)) return 0.0f;
  float mix[GD_MIX];
  gd_mixture(o, mix);
  float v[GD_LOBES], c[GD_LOBES], p = 0.0f;
  for (int j = 0; j < GD_LOBES; j++) {
    c[j] = mix[5 * j + 1] * dir[0] + mix[5 * j + 2] * dir[1] + mix[5 * j + 3] * dir[2];
    v[j] = gd_vmf(mix[5 * j + 4], c[j]);
    p += mix[5 * j] * v[j];
  }
  p = fmaxf(p, 1e-20f);
  for (int j = 0; j < GD_LOBES; j++) {
    const float pi = mix[5 * j], k = mix[5 * j + 4];
    const float r = pi * v[j] / p;                 // responsibility
    g[5 * j] = -wgt * (r - pi);                     // softmax logits
    const float x = o[5 * j + 1], y = o[5 * j + 2], z = o[5 * j + 3];
    const float l2 = x * x + y * y + z * z;
    if (l2 >= 1e-12f) {
      const float il = 1.0f / sqrtf(l2);
      const float s = -wgt * r * k * il;           // d log v / d mu = k dir
      g[5 * j + 1] = s * (dir[0] - c[j] * mix[5 * j + 1]);
      g[5 * j + 2] = s * (dir[1] - c[j] * mix[5 * j + 2]);
      g[5 * j + 3] = s * (dir[2] - c[j] * mix[5 * j + 3]);
    }
    const float lk = o[5 * j + 4];
    if (lk > GD_LOGK_MIN && lk < GD_LOGK_MAX) {
      // d log v / d k = 1/k - 2 / (e^{2k} - 1) + (mu.dir - 1); dk/d(log k) = k
      const float dl = 1.0f / k - 2.0f / expm1f(2.0f * k) + (c[j] - 1.0f);
      g[5 * j + 4] = -wgt * r * k * dl;
    }
  }
  return -wgt * logf(p);
}

#endif
