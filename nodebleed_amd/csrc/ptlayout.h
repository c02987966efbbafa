/*
 * Device records of the wavefront path tracer, shared by the host C code
 * (host/pt.c) and the kernels (kern/pt.hip, kern/guide.hip).  Plain C.
 */
#ifndef PTLAYOUT_H
#define PTLAYOUT_H

#include <stdint.h>

#include "dlayout.h"

/* One view of a batch: camera, resolution, samples per pixel, the index of
 * its first path in the batch (prefix sum of w*h*spp), its accumulator
 * offset in pixels, the sample index of its first sample this frame, and
 * its random stream: a hash of the camera and resolution, so a view's
 * samples do not depend on its place in the batch or on the other views. */
struct PtView {
	struct CamArgs  cam;
	uint32_t        w, h, spp, first;
	uint32_t        acc_ofs, sample_base, pitch, stream;
};

/* Launch constants of a frame. */
struct PtArgs {
	uint32_t  nviews, nlights, maxdepth, rr_depth;
	float     bg[3];
	uint32_t  seed;
	float     guide_mix;      /* probability of sampling the learned proposal */
	float     gmin[3];        /* guide grid domain: scene box */
	float     ginv[3];        /* 1 / extent */
	uint32_t  train_stride;   /* record vertices of every n-th path (0 = off) */
	uint32_t  train_cap;      /* training records per batch */
	uint32_t  guide_depth;    /* vertices at depth < guide_depth are guided */
	uint32_t  guide_share;    /* depth 0: one proposal per this many paths */
	float     train_bsdf;     /* training density: (1 - b) vMF + b cos/pi (0: vMF alone) */
};

/* Path state, 64 bytes. */
struct PtPath {
	float     thr[3];
	float     L[3];
	uint32_t  pix;        /* accumulator pixel (view offset + pixel) */
	uint32_t  key;        /* counter-based RNG key of the path */
	uint32_t  depth;
	float     pdf_prev;   /* pdf of the last sampled direction */
	uint32_t  rec;        /* training record of the current vertex, or ~0 */
	uint32_t  rec_prev;   /* training record of the previous vertex, or ~0 */
	uint32_t  pad[4];
};

/* Emissive triangle in world space, 64 bytes; pick = probability of
 * choosing it (area * luminance / total). */
struct PtLight {
	float     v0[3], e0[3], e1[3], n[3], E[3];
	float     area, pick;
	uint32_t  pad;
};

/* ---- learned path-guiding proposal ---------------------------------------
 * Hash-grid encoding of the shading position (GD_LEVELS levels of
 * GD_TABLE entries, GD_FEAT features each) + the shading normal + a
 * constant 1 -> GD_IN inputs; 3 hidden ReLU layers of GD_HID; GD_OUT
 * outputs of which the first 5 * GD_LOBES parameterise the vMF mixture
 * (logit, mean direction xyz, log concentration per lobe). */
#define GD_LEVELS  12
#define GD_TABLE   (1u << 19)
#define GD_FEAT    2
#define GD_IN      32
#define GD_HID     64
#define GD_OUT     64
#define GD_LOBES   8
/* training records per optimiser step unless set otherwise */
#define GD_DEFAULT_CHUNK 4096u
/* records drawn in proportion to their target (1) or uniformly (0) */
#define GD_DEFAULT_IMPORTANCE 1
/* guided vertices: depth < GD_DEFAULT_DEPTH (1: primary hits only) */
#define GD_DEFAULT_DEPTH 1
/* training target: 0 = all radiance the path gathers after the vertex,
   1 = the next vertex's direct light (default: lower-variance targets; the
   equal-error ratio of the occluded-lights scene 0.490 against 0.502, means
   of four and six runs) */
#define GD_DEFAULT_TARGET 1
/* default training objective: 0 fits the vMF mixture alone, 1 the MIS density */
#define GD_DEFAULT_OBJECTIVE 0
/* warm-up: Adam steps per frame for the first frames after enable */
#define GD_DEFAULT_WARM_FRAMES 16u
#define GD_DEFAULT_WARM_STEPS  64u
/* the MLP's learning rate relative to the hash grid's */
#define GD_DEFAULT_MLP_LR_SCALE 0.1f
#define GD_NMIN    16.0f
#define GD_NMAX    1024.0f

/* Offsets (in floats) of the fp32 master parameters.  Weights are stored
 * out-major: W[n][k] multiplies input k into output n. */
#define GD_N_GRID  ((unsigned)(GD_LEVELS * GD_TABLE * GD_FEAT))
#define GD_OFS_W1  GD_N_GRID
#define GD_OFS_B1  (GD_OFS_W1 + GD_HID * GD_IN)
#define GD_OFS_W2  (GD_OFS_B1 + GD_HID)
#define GD_OFS_B2  (GD_OFS_W2 + GD_HID * GD_HID)
#define GD_OFS_W3  (GD_OFS_B2 + GD_HID)
#define GD_OFS_B3  (GD_OFS_W3 + GD_HID * GD_HID)
#define GD_OFS_W4  (GD_OFS_B3 + GD_HID)
#define GD_OFS_B4  (GD_OFS_W4 + GD_OUT * GD_HID)
#define GD_NPARAM  (GD_OFS_B4 + GD_OUT)
#define GD_NMLP    (GD_NPARAM - GD_N_GRID)
/* bf16 weight mirror: W1..W4 packed in natural order, then the same in the
 * permuted order of the register-resident inference kernel */
#define GD_NW16    (GD_HID * GD_IN + 2 * GD_HID * GD_HID + GD_OUT * GD_HID)
/* bf16 weight mirrors: plain, permuted (inference), transposed (backward) */
#define GD_NW16T   (GD_IN * GD_HID + 3 * GD_HID * GD_HID)
#define GD_W16_ELEMS (2 * GD_NW16 + GD_NW16T)

/* Per shading point mixture: for each lobe weight, mean xyz, kappa. */
#define GD_MIX     (5 * GD_LOBES)

/* Training record written where an indirect direction is sampled. */
struct GdRecord {
	float     pos[3];
	float     nrm[3];
	float     dir[3];
	float     pdf;        /* density the direction was sampled with */
	float     lbefore;    /* path luminance before this vertex's continuation */
	float     thr;        /* throughput luminance after the continuation */
	uint32_t  path;
	float     target;     /* incident radiance estimate, filled after the frame */
	float     l1;         /* luminance of the next vertex's direct light */
	uint32_t  pad;
};

#define PT_CNT_CUR     0
#define PT_CNT_NEXT    1
#define PT_CNT_SHADOW  2
#define PT_CNT_TRAIN   3
#define PT_NCOUNTS     8

#endif
