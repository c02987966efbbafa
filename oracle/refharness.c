/*
 * ORACLE (test infrastructure only).  Headless harness around the
 * reference's own renderer sources, compiled from /root/reference by
 * oracle/Makefile into oracle/_ref/libref.so (never copied into this
 * repository).  It exposes the reference's BVH builds and single-ray
 * traversal over plain buffers so tests can compare this repository's
 * builder bytes and the oracle's hits with the real thing:
 *   ref_scene: rend_init + rend_prepstatic + upd_rinsts-equivalent AABBs +
 *              rend_prepdynamic (rend.c:3000-3090, main.c:117-157)
 *   ref_trace: intersect_tlas per ray (rend.c:1976), t starts at tmax
 *   ref_occluded: intersect_any_tlas per ray (rend.c:2883)
 * The instance AABB step restates upd_rinsts (main.c:117-157) because main.c
 * needs SDL; it calls the reference's own mulpos.
 */
#include <float.h>
#include <pthread.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "rend.h"
#include "util.h"
#include "vec3.h"

struct rray { float ox, oy, oz, tmax, dx, dy, dz, pad; };
struct rhit { float t, u, v; unsigned int id; };

void intersect_tlas(float *t, float *u, float *v, unsigned int *id,
                    float ox, float oy, float oz, float dx, float dy,
                    float dz, struct bnode8 *nodes, struct rinst *insts,
                    unsigned int tlasofs);
bool intersect_any_tlas(float tfar, float ox, float oy, float oz, float dx,
                        float dy, float dz, struct bnode8 *nodes,
                        struct rinst *insts, unsigned int tlasofs);
void mulpos(float *ox, float *oy, float *oz, float m[16], float x, float y,
            float z);

static int g_init;

/* globs: 16 floats (world matrix) per instance, globinv already in insts. */
struct rdata *ref_scene(const struct rtri *tris, unsigned ntris,
                        const struct rinst *insts, unsigned ninst,
                        const float *globs)
{
	if (!g_init) {
		rend_staticinit(64, 32);
		g_init = 1;
	}
	struct rdata *rd = aligned_alloc(64, (sizeof(*rd) + 63) & ~(size_t)63);
	memset(rd, 0, sizeof(*rd));
	rend_init(rd, 1, ntris, ninst, 1);
	memcpy(rd->tris, tris, (size_t)ntris * sizeof(*tris));
	memcpy(rd->insts, insts, (size_t)ninst * sizeof(*insts));
	rd->instcnt = ninst;
	rend_prepstatic(rd);
	for (unsigned j = 0; j < ninst; j++) {
		struct rinst *ri = &rd->insts[j];
		float m[16];
		memcpy(m, &globs[16 * j], sizeof(m));
		struct bnode8 *n = &rd->bnodes[ri->triofs << 1];
		float nmix = bcmin8(n->minx8)[0];
		float nmiy = bcmin8(n->miny8)[0];
		float nmiz = bcmin8(n->minz8)[0];
		float nmax = bcmax8(n->maxx8)[0];
		float nmay = bcmax8(n->maxy8)[0];
		float nmaz = bcmax8(n->maxz8)[0];
		struct vec3 mi = {FLT_MAX, FLT_MAX, FLT_MAX};
		struct vec3 ma = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
		for (unsigned char i = 0; i < 8; i++) {
			struct vec3 v;
			mulpos(&v.x, &v.y, &v.z, m, i & 1 ? nmix : nmax,
			  i & 2 ? nmiy : nmay, i & 4 ? nmiz : nmaz);
			mi = vec3_min(mi, v);
			ma = vec3_max(ma, v);
		}
		rd->aabbs[j] = (struct aabb){.min = mi, .max = ma};
	}
	rend_prepdynamic(rd);
	return rd;
}

void ref_scene_free(struct rdata *rd)
{
	rend_release(rd);
	free(rd);
}

struct bnode8 *ref_bnodes(struct rdata *rd) { return rd->bnodes; }
unsigned ref_tlasofs(struct rdata *rd) { return rd->tlasofs; }
struct aabb *ref_aabbs(struct rdata *rd) { return rd->aabbs; }

struct job {
	struct rdata *rd;
	const struct rray *rays;
	struct rhit *hits;
	unsigned char *occ;
	unsigned n, next;
};

static void *worker(void *arg)
{
	struct job *j = arg;
	for (;;) {
		unsigned b = __atomic_fetch_add(&j->next, 64u, __ATOMIC_SEQ_CST);
		if (b >= j->n)
			break;
		unsigned e = b + 64 < j->n ? b + 64 : j->n;
		for (unsigned i = b; i < e; i++) {
			const struct rray *r = &j->rays[i];
			if (j->hits) {
				float t = r->tmax, u = 0.0f, v = 0.0f;
				unsigned id = 0xffffffffu;
				intersect_tlas(&t, &u, &v, &id, r->ox, r->oy, r->oz, r->dx,
				  r->dy, r->dz, j->rd->bnodes, j->rd->insts, j->rd->tlasofs);
				j->hits[i] = (struct rhit){t, u, v, id};
			} else {
				j->occ[i] = intersect_any_tlas(r->tmax, r->ox, r->oy, r->oz,
				  r->dx, r->dy, r->dz, j->rd->bnodes, j->rd->insts,
				  j->rd->tlasofs);
			}
		}
	}
	return NULL;
}

static void run(struct job *j, int threads)
{
	pthread_t th[256];
	if (threads < 1)
		threads = 1;
	if (threads > 256)
		threads = 256;
	int started = 0;
	for (int i = 1; i < threads; i++)
		if (!pthread_create(&th[i], NULL, worker, j))
			started = i;
		else
			break;
	worker(j);
	for (int i = 1; i <= started; i++)
		pthread_join(th[i], NULL);
}

void ref_trace(struct rdata *rd, const struct rray *rays, struct rhit *hits,
               unsigned n, int threads)
{
	struct job j = {rd, rays, hits, NULL, n, 0};
	run(&j, threads);
}

void ref_occluded(struct rdata *rd, const struct rray *rays,
                  unsigned char *occ, unsigned n, int threads)
{
	struct job j = {rd, rays, NULL, occ, n, 0};
	run(&j, threads);
}

float ref_rcpps(float x)
{
	return _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
}

/* The reference's primary rays of block blk (rend_render's walk,
 * rend.c:3603-3660): rand8_init seeded as there, packets in Morton order
 * through its own make_camray_pckt; rd supplies width and the camera. */
void make_camray_pckt(__m256 *ox8, __m256 *oy8, __m256 *oz8,
                      __m256 *dx8, __m256 *dy8, __m256 *dz8,
                      unsigned int x, unsigned int y,
                      __m256 xofs8, __m256 yofs8,
                      struct rcam *c, struct rngstate8 *rngstate);

void ref_camrays(struct rdata *rd, unsigned bszx, unsigned bszy, unsigned samples,
                 unsigned blk, struct rray *out)
{
	rend_staticinit(bszx, bszy);
	unsigned blkcntx = rd->width / bszx, npk = bszx * bszy / 8;
	struct rngstate8 rng;
	rand8_init(&rng, ((int)blk + 13) * 1571, samples * 23);
	__m256 xo = _mm256_set_ps(3, 2, 1, 0, 3, 2, 1, 0);
	__m256 yo = _mm256_set_ps(1, 1, 1, 1, 0, 0, 0, 0);
	for (unsigned k = 0; k < npk; k++) {
		__m256 v[6];
		make_camray_pckt(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
		  (blk % blkcntx) * bszx + mortdecx(k) * 4,
		  (blk / blkcntx) * bszy + mortdecy(k) * 2, xo, yo, &rd->cam, &rng);
		float f[6][8];
		for (int a = 0; a < 6; a++)
			_mm256_storeu_ps(f[a], v[a]);
		for (int l = 0; l < 8; l++)
			out[k * 8 + l] = (struct rray){f[0][l], f[1][l], f[2][l], FLT_MAX,
			  f[3][l], f[4][l], f[5][l], 0.0f};
	}
	rend_staticrelease();
}

float ref_rsqrtps(float x)
{
	return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
}
