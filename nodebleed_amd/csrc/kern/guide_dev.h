// Device side of the learned proposal as used by k_pt_shade: one-sample MIS
// over the mixture alpha * vMF mixture + (1 - alpha) * cosine BSDF sampling.
#ifndef GUIDE_DEV_H
#define GUIDE_DEV_H

#include "../ptlayout.h"
#include "guide_math.h"

// Samples a direction about shading normal n (ONB b1, b2) from the mixture;
// wgt = (cos / pi) / p(dir) is the Lambertian throughput factor divided by
// the albedo; pdf = p(dir).  Returns false when the direction is below the
// shading hemisphere (zero BSDF).
__device__ __forceinline__ bool guide_sample(const float *mix, float alpha, float nx, float ny,
                                             float nz, float b1x, float b1y, float b1z,
                                             float b2x, float b2y, float b2z, float u0,
                                             float u1, float u2, float &dx, float &dy,
                                             float &dz, float &wgt, float &pdf) {
  // a proposal whose first weight is negative is empty (its point missed):
  // BSDF sampling alone
  if (mix[0] < 0.0f) alpha = 0.0f;
  if (u0 < alpha) {
    gd_sample(mix, u0 / alpha, u1, u2, dx, dy, dz);
  } else {
    const float uu = (u0 - alpha) / (1.0f - alpha);
    float sn, cs;
    sincosf(6.28318531f * uu, &sn, &cs);
    const float su = sqrtf(u1), lx = cs * su, ly = sn * su, lz = sqrtf(1.0f - u1);
    dx = b1x * lx + b2x * ly + nx * lz;
    dy = b1y * lx + b2y * ly + ny * lz;
    dz = b1z * lx + b2z * ly + nz * lz;
  }
  const float c = nx * dx + ny * dy + nz * dz;
  if (!(c > 0.0f)) {
    wgt = 0.0f;
    pdf = 0.0f;
    return false;
  }
  const float pc = c * 0.318309886f;
  const float p = alpha * gd_pdf(mix, dx, dy, dz) + (1.0f - alpha) * pc;
  pdf = p;
  wgt = p > 0.0f ? pc / p : 0.0f;
  return p > 0.0f;
}

#endif
