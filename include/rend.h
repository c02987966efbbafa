#ifndef REND_H
#define REND_H

/*
 * librend: drop-in replacement for the reference renderer's public header
 * (rend.h:1-164).  Every struct keeps the reference's field names, sizes and
 * offsets, because callers (the scene bridge in main.c:53-195) write these
 * arrays directly.  The render entry points keep their signatures; the work
 * behind them runs on one MI355X through HIP (see rend_hip.h for the
 * additional, out-of-band entry points).
 *
 * x86 vector types are kept where the reference exposes them (bnode8, leaf4,
 * rcam) so that caller code such as bcmin8(n->minx8) (main.c:135) compiles
 * unchanged.
 */

#include <immintrin.h>
#include "vec3.h"

#define SPP 1                    /* rend.h:7 */

#define BRANCH_MAX     8         /* rend.h:9  children per node */
#define BLAS_LEAF_MAX  4         /* rend.h:10 triangles per BLAS leaf */
#define TLAS_LEAF_MAX  1         /* rend.h:11 instances per TLAS leaf */

#define NODE_LEAF  0x80000000    /* rend.h:13 bit 31 of a child word */

/*
 * 8-wide BVH node, 256 bytes (rend.h:17-30).  children8[i]: NODE_LEAF|byte
 * offset of a leaf4 (BLAS), NODE_LEAF|instance id (TLAS), or the byte offset
 * of a child bnode8, all relative to the root of the tree.  perm8[i] holds,
 * for each of the 8 ray-direction octants o, the child index at traversal
 * rank i in bits [3o, 3o+3).
 */
struct bnode8 {
	__m256   minx8;
	__m256   maxx8;
	__m256   miny8;
	__m256   maxy8;
	__m256   minz8;
	__m256   maxz8;
	__m256i  children8;
	__m256i  perm8;
};

/* Four triangles with precomputed edges, 192 bytes (rend.h:32-45). */
struct leaf4 {
	__m128        v0x4;
	__m128        v0y4;
	__m128        v0z4;
	__m128        e0x4;   /* v1 - v0 */
	__m128        e0y4;
	__m128        e0z4;
	__m128        e1x4;   /* v2 - v0 */
	__m128        e1y4;
	__m128        e1z4;
	__m128i       id4;    /* triangle index local to the BLAS */
	unsigned int  tricnt;
	unsigned int  pad[7];
};

struct aabb {  /* 32 bytes, rend.h:47-52 */
	struct vec3  min;
	float        pad0;
	struct vec3  max;
	float        pad1;
};

struct rmtl {  /* 32 bytes, rend.h:54-62 */
	struct vec3   col;
	float         metallic;
	float         roughness;
	float         ior;
	unsigned int  flags;   /* EMISSIVE (0x1000): col is the emitted radiance */
	unsigned int  pad0;
};

struct rtri {  /* 48 bytes, rend.h:64-71 */
	struct vec3  v0;
	float        pad0;
	struct vec3  v1;
	float        pad1;
	struct vec3  v2;
	float        pad2;
};

struct rnrm {  /* 48 bytes, rend.h:73-80 */
	struct vec3   n0;
	unsigned int  mtlid;
	struct vec3   n1;
	unsigned int  pad0;
	struct vec3   n2;
	unsigned int  pad1;
};

struct rinst {  /* 64 bytes, rend.h:82-88 */
	float         globinv[12];  /* inverse world transform, rows of 3x4 */
	unsigned int  flags;
	unsigned int  triofs;
	unsigned int  tricnt;
	unsigned int  pad0;
};

struct rcam {  /* rend.h:89-117 */
	struct vec3  eye;
	float        tanvfov;
	struct vec3  ri;
	float        focdist;
	struct vec3  up;
	float        tanfangle;
	struct vec3  fwd;
	float        aspect;
	__m256       eyex8;
	__m256       eyey8;
	__m256       eyez8;
	__m256       rix8;
	__m256       riy8;
	__m256       riz8;
	__m256       upx8;
	__m256       upy8;
	__m256       upz8;
	__m256       fwdx8;
	__m256       fwdy8;
	__m256       fwdz8;
	__m256       rw8;
	__m256       rh8;
	__m256       aspect8;
	__m256       fdist8;
	__m256       fovfdist8;
	__m256       focangle8;
	__m256       focrad8;
};

struct rdata {  /* rend.h:119-147 */
	struct rmtl    *mtls;

	struct rtri    *tris;
	struct rnrm    *nrms;

	unsigned int   instcnt;
	struct rinst   *insts;
	struct aabb    *aabbs;

	struct bnode8  *bnodes;   /* all BLAS arenas, then the TLAS */
	unsigned int   tlasofs;   /* TLAS root index in bnodes (2 * maxtris) */

	struct rcam    cam;

	struct vec3    bgcol;

	unsigned int   width;
	unsigned int   height;
	struct vec3    *acc;      /* accumulator, owned by the renderer */
	unsigned int   *buf;      /* ARGB8888, owned by the application */

	int            blknum;    /* block claim counter, reset per frame */

	unsigned int   rays;
	unsigned int   spp;
	unsigned int   samples;
};

#ifdef __cplusplus
extern "C" {
#endif

/* rend.h:149-150; rend.c:117-169 */
void  rend_staticinit(unsigned int blkszx, unsigned int blkszy);
void  rend_staticrelease(void);

/* rend.h:152-154; rend.c:3000-3031 */
void  rend_init(struct rdata *rd, unsigned int maxmtls, unsigned int maxtris,
                unsigned int maxinsts, unsigned int spp);
void  rend_release(struct rdata *rd);

/* rend.h:156-157; rend.c:3033-3090.  Builds on the host, then repacks and
 * uploads the BVH to the GPU. */
void  rend_prepstatic(struct rdata *rd);
void  rend_prepdynamic(struct rdata *rd);

/* rend.h:159-160; rend.c:3092-3107 */
void  rend_resaccum(struct rdata *rd, unsigned int w, unsigned int h);
void  rend_clraccum(struct rdata *rd);

/* rend.h:162; rend.c:3603-3708.  A thrd_start_t: the first caller of a frame
 * claims every block and renders the frame on the GPU, later callers of the
 * same frame return at once.  Always returns 0. */
int   rend_render(void *rd);

/* rend.c:171-187, called by the reference's upd_rinsts (main.c:114-115,143). */
void  mulpos(float *ox, float *oy, float *oz,
             float m[16], float x, float y, float z);
void  muldir(float *ox, float *oy, float *oz,
             float m[16], float x, float y, float z);

#ifdef __cplusplus
}
#endif

#endif
