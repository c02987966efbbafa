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
 *   ref_render / ref_render_single: whole frames of rend_render
 *              (rend.c:3603-3708) and of its single-ray shading path
 *   ref_gltf_*: a glTF scene through the reference's importer and scene
 *              graph (gltfimp.c imp_gltf_f, scene.c scene_updanims /
 *              scene_updtransforms / scene_updcams) into an rdata, set up
 *              and advanced as main.c's init() and update() do
 *              (main.c:208-278)
 * The instance AABB step restates upd_rinsts (main.c:117-157), and the glTF
 * path restates cpy_rdata and upd_rcam (main.c:53-110, 159-195), because
 * main.c needs SDL; they call the reference's own mulpos, rend_* and scene_*
 * functions.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "gltfimp.h"
#include "mat4.h"
#include "rend.h"
#include "scene.h"
#include "types.h"
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
	free(rd->buf);
	rd->buf = NULL;
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

/* ---- the reference's own frames (rend_render, rend.c:3603-3708) ----------
 *
 * ref_frame_setup copies what main.c's cpy_rdata/upd_rcam give rend_render
 * (materials, normals, camera, background) from an rdata that this
 * repository's scene builder filled (identical struct layouts), allocates the
 * accumulator (rend_resaccum) and an application colour buffer.
 * ref_render runs `frames` frames of the reference's active rend_render on
 * `threads` threads with main.c's frame protocol (main.c:407-426).
 * ref_render_single walks the same blocks and packets with the same random
 * streams but traces every packet with the reference's single-ray
 * trace_rays (rend.c:3404) before accum_pckt (rend.c:3565): the frame
 * rend_render would produce if no packet were coherent. */
void trace_rays(struct vec3 *col, __m256 ox8, __m256 oy8, __m256 oz8,
                __m256 dx8, __m256 dy8, __m256 dz8, struct rdata *rd,
                unsigned int *rays);
void accum_pckt(struct vec3 *acc, unsigned int *buf, struct vec3 *col,
                unsigned int px, unsigned int py, float *xofs, float *yofs,
                unsigned int w, float invsamples);
int rend_render(void *d);

void ref_frame_setup(struct rdata *rd, const struct rdata *src, unsigned nmtls,
                     unsigned ntris)
{
	free(rd->mtls);
	rd->mtls = aligned_alloc(64, ((size_t)(nmtls ? nmtls : 1) * sizeof(*rd->mtls) + 63) &
	  ~(size_t)63);
	memcpy(rd->mtls, src->mtls, (size_t)nmtls * sizeof(*rd->mtls));
	memcpy(rd->nrms, src->nrms, (size_t)ntris * sizeof(*rd->nrms));
	rd->cam = src->cam;
	rd->bgcol = src->bgcol;
	rd->spp = src->spp;
	rend_resaccum(rd, src->width, src->height);
	free(rd->buf);
	rd->buf = calloc((size_t)src->width * src->height, sizeof(unsigned));
	rd->samples = 0;
}

static void *frame_thread(void *rd)
{
	rend_render(rd);
	return NULL;
}

static void copy_out(const struct rdata *rd, unsigned *buf, float *acc)
{
	size_t n = (size_t)rd->width * rd->height;
	memcpy(buf, rd->buf, n * sizeof(unsigned));
	memcpy(acc, rd->acc, n * sizeof(struct vec3));
}

unsigned ref_render(struct rdata *rd, unsigned bszx, unsigned bszy, unsigned frames,
                    int threads, unsigned *buf, float *acc)
{
	rend_staticinit(bszx, bszy);
	rend_clraccum(rd);
	memset(rd->buf, 0, (size_t)rd->width * rd->height * sizeof(unsigned));
	unsigned rays = 0;
	pthread_t th[64];
	if (threads < 1)
		threads = 1;
	if (threads > 64)
		threads = 64;
	for (unsigned f = 0; f < frames; f++) {
		rd->blknum = 0;
		rd->rays = 0;
		int started = 0;
		for (int i = 0; i < threads; i++)
			if (!pthread_create(&th[i], NULL, frame_thread, rd))
				started++;
			else
				break;
		if (!started)
			rend_render(rd);
		for (int i = 0; i < started; i++)
			pthread_join(th[i], NULL);
		rays += rd->rays;
		rd->samples += rd->spp;
	}
	copy_out(rd, buf, acc);
	rend_staticrelease();
	return rays;
}

unsigned ref_render_single(struct rdata *rd, unsigned bszx, unsigned bszy, unsigned frames,
                           unsigned *buf, float *acc)
{
	rend_staticinit(bszx, bszy);
	rend_clraccum(rd);
	memset(rd->buf, 0, (size_t)rd->width * rd->height * sizeof(unsigned));
	float xofs[8] = {0, 1, 2, 3, 0, 1, 2, 3}, yofs[8] = {0, 0, 0, 0, 1, 1, 1, 1};
	__m256 xo = _mm256_set_ps(3, 2, 1, 0, 3, 2, 1, 0);
	__m256 yo = _mm256_set_ps(1, 1, 1, 1, 0, 0, 0, 0);
	unsigned blkcntx = rd->width / bszx, blkcnt = blkcntx * (rd->height / bszy);
	unsigned npk = bszx * bszy / 8, rays = 0;
	for (unsigned f = 0; f < frames; f++) {
		float invs = 1.0f / (rd->spp + rd->samples);
		for (unsigned blk = 0; blk < blkcnt; blk++) {
			struct rngstate8 rng;
			rand8_init(&rng, ((int)blk + 13) * 1571, rd->samples * 23);
			for (unsigned k = 0; k < npk; k++) {
				unsigned px = (blk % blkcntx) * bszx + mortdecx(k) * 4;
				unsigned py = (blk / blkcntx) * bszy + mortdecy(k) * 2;
				__m256 v[6];
				make_camray_pckt(&v[0], &v[1], &v[2], &v[3], &v[4], &v[5], px, py,
				  xo, yo, &rd->cam, &rng);
				struct vec3 col[8];
				trace_rays(col, v[0], v[1], v[2], v[3], v[4], v[5], rd, &rays);
				accum_pckt(rd->acc, rd->buf, col, px, py, xofs, yofs, rd->width, invs);
			}
		}
		rd->samples += rd->spp;
	}
	copy_out(rd, buf, acc);
	rend_staticrelease();
	return rays;
}

/* ---- glTF scenes through the reference's scene graph ----------------------- */

struct ref_gltf {
	struct scene  s;
	struct rdata *rd;
	unsigned      ntris;
};

/* World AABB of instance `inst` from its BLAS root and world matrix m
 * (upd_rinsts, main.c:117-157). */
static void inst_box(struct rdata *rd, unsigned inst, float *m)
{
	struct rinst *ri = &rd->insts[inst];
	struct bnode8 *n = &rd->bnodes[ri->triofs << 1];
	float lo[3] = {bcmin8(n->minx8)[0], bcmin8(n->miny8)[0], bcmin8(n->minz8)[0]};
	float hi[3] = {bcmax8(n->maxx8)[0], bcmax8(n->maxy8)[0], bcmax8(n->maxz8)[0]};
	struct vec3 mi = {FLT_MAX, FLT_MAX, FLT_MAX};
	struct vec3 ma = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (unsigned char c = 0; c < 8; c++) {
		struct vec3 v;
		mulpos(&v.x, &v.y, &v.z, m, c & 1 ? lo[0] : hi[0], c & 2 ? lo[1] : hi[1],
		  c & 4 ? lo[2] : hi[2]);
		mi = vec3_min(mi, v);
		ma = vec3_max(ma, v);
	}
	rd->aabbs[inst] = (struct aabb){.min = mi, .max = ma};
}

static void gltf_insts(struct ref_gltf *g)
{
	struct scene *s = &g->s;
	for (unsigned j = 0; j < s->nodecnt; j++) {
		struct obj *o = &s->objs[j];
		if (!hasflags(o->flags, MESH))
			continue;
		float inv[16];
		mat4_inv(inv, s->tfmats[j].glob);
		memcpy(g->rd->insts[o->instid].globinv, inv, sizeof(g->rd->insts[0].globinv));
		inst_box(g->rd, (unsigned)o->instid, s->tfmats[j].glob);
	}
}

/* rcam of the scene's current camera (upd_rcam, main.c:159-195). */
static void gltf_rcam(struct ref_gltf *g)
{
	struct cam *c = &g->s.cams[g->s.currcam];
	struct rcam *rc = &g->rd->cam;
	float w = g->rd->width, h = g->rd->height;
	float aspect = w / h;
	float tv = tanf(0.5f * c->vfov * PI / 180.0f);
	float tf = tanf(0.5f * c->focangle * PI / 180.0f);
	memset(rc, 0, sizeof(*rc));
	rc->eye = c->eye;
	rc->tanvfov = tv;
	rc->ri = c->ri;
	rc->focdist = c->focdist;
	rc->up = c->up;
	rc->tanfangle = tf;
	rc->fwd = c->fwd;
	rc->aspect = aspect;
	rc->eyex8 = _mm256_set1_ps(c->eye.x);
	rc->eyey8 = _mm256_set1_ps(c->eye.y);
	rc->eyez8 = _mm256_set1_ps(c->eye.z);
	rc->rix8 = _mm256_set1_ps(c->ri.x);
	rc->riy8 = _mm256_set1_ps(c->ri.y);
	rc->riz8 = _mm256_set1_ps(c->ri.z);
	rc->upx8 = _mm256_set1_ps(c->up.x);
	rc->upy8 = _mm256_set1_ps(c->up.y);
	rc->upz8 = _mm256_set1_ps(c->up.z);
	rc->fwdx8 = _mm256_set1_ps(c->fwd.x);
	rc->fwdy8 = _mm256_set1_ps(c->fwd.y);
	rc->fwdz8 = _mm256_set1_ps(c->fwd.z);
	rc->rw8 = _mm256_set1_ps(1.0f / w);
	rc->rh8 = _mm256_set1_ps(1.0f / h);
	rc->aspect8 = _mm256_set1_ps(aspect);
	rc->fdist8 = _mm256_set1_ps(c->focdist);
	rc->fovfdist8 = _mm256_set1_ps(2.0f * tv * c->focdist);
	rc->focangle8 = _mm256_set1_ps(tf);
	rc->focrad8 = _mm256_set1_ps(tf * c->focdist);
}

/* Import and set up as main.c's init() (main.c:208-251); bg is the
 * background colour the application sets. */
struct ref_gltf *ref_gltf_open(const char *path, unsigned w, unsigned h, unsigned spp,
                               float bgr, float bgg, float bgb)
{
	if (!g_init) {
		rend_staticinit(64, 32);
		g_init = 1;
	}
	struct ref_gltf *g = calloc(1, sizeof(*g));
	struct scene *s = &g->s;
	imp_gltf_f(s, (char *)path);
	s->bgcol = (struct vec3){bgr, bgg, bgb};
	unsigned ninst = 0;
	for (unsigned i = 0; i < s->meshcnt; i++)
		g->ntris += s->meshes[i].icnt / 3;
	for (unsigned i = 0; i < s->nodecnt; i++)
		ninst += hasflags(s->objs[i].flags, MESH);
	g->rd = aligned_alloc(64, (sizeof(*g->rd) + 63) & ~(size_t)63);
	memset(g->rd, 0, sizeof(*g->rd));
	struct rdata *rd = g->rd;
	rend_init(rd, s->mtlmax, g->ntris, ninst, spp);
	rend_resaccum(rd, w, h);
	rd->buf = calloc((size_t)w * h, sizeof(unsigned));

	/* cpy_rdata: triangles mesh by mesh and material run by run */
	unsigned *triofs = calloc(s->meshcnt ? s->meshcnt : 1, sizeof(unsigned));
	struct rtri *rt = rd->tris;
	struct rnrm *rn = rd->nrms;
	unsigned ofs = 0;
	for (unsigned k = 0; k < s->meshcnt; k++) {
		struct mesh *m = &s->meshes[k];
		const unsigned *ip = m->inds;
		for (unsigned j = 0; j < m->mcnt; j++)
			for (unsigned i = 0; i < m->mtls[j].tricnt; i++, ip += 3, rt++, rn++) {
				memcpy(&rt->v0, &m->vrts[ip[0]], sizeof(rt->v0));
				memcpy(&rt->v1, &m->vrts[ip[1]], sizeof(rt->v1));
				memcpy(&rt->v2, &m->vrts[ip[2]], sizeof(rt->v2));
				memcpy(&rn->n0, &m->nrms[ip[0]], sizeof(rn->n0));
				memcpy(&rn->n1, &m->nrms[ip[1]], sizeof(rn->n1));
				memcpy(&rn->n2, &m->nrms[ip[2]], sizeof(rn->n2));
				rn->mtlid = m->mtls[j].mtlid;
			}
		triofs[k] = ofs;
		ofs += m->icnt / 3;
	}
	/* struct mtl (28 bytes) into struct rmtl (32): the pad is left out */
	for (unsigned i = 0; i < s->mtlcnt; i++) {
		memset(&rd->mtls[i], 0, sizeof(rd->mtls[i]));
		memcpy(&rd->mtls[i], &s->mtls[i], sizeof(s->mtls[i]));
	}
	unsigned cnt = 0;
	for (unsigned i = 0; i < s->nodecnt; i++) {
		struct obj *o = &s->objs[i];
		if (!hasflags(o->flags, MESH))
			continue;
		rd->insts[cnt] = (struct rinst){.flags = o->flags, .triofs = triofs[o->objid],
		  .tricnt = s->meshes[o->objid].icnt / 3};
		o->instid = (int)cnt++;
	}
	rd->instcnt = cnt;
	rd->bgcol = s->bgcol;
	free(triofs);

	rend_prepstatic(rd);
	scene_updanims(s, 0.0f);
	scene_updtransforms(s);
	scene_updcams(s);
	return g;
}

/* One update() at `time` (main.c:253-278); returns 1 when instances and the
 * TLAS were refreshed, | 2 when the camera was. */
int ref_gltf_update(struct ref_gltf *g, float time)
{
	struct scene *s = &g->s;
	int r = 0;
	scene_updanims(s, time);
	if (hasflags(s->dirty, TRANSFORM)) {
		scene_updtransforms(s);
		gltf_insts(g);
		rend_prepdynamic(g->rd);
		r |= 1;
	}
	if (hasflags(s->dirty, CAM)) {
		scene_updcams(s);
		gltf_rcam(g);
		r |= 2;
	}
	clrflags(&s->dirty, TRANSFORM | CAM);
	return r;
}

struct rdata *ref_gltf_rdata(struct ref_gltf *g) { return g->rd; }

/* A new accumulation (main.c:275-278 clears it after every change). */
void ref_gltf_restart(struct ref_gltf *g)
{
	g->rd->samples = 0;
	g->rd->blknum = 0;
	g->rd->rays = 0;
}

/* mtls, tris, nodes, tracks, samplers of the imported scene. */
void ref_gltf_counts(struct ref_gltf *g, unsigned out[5])
{
	out[0] = g->s.mtlcnt;
	out[1] = g->ntris;
	out[2] = g->s.nodecnt;
	out[3] = g->s.trackcnt;
	out[4] = g->s.samplercnt;
}

void ref_gltf_close(struct ref_gltf *g)
{
	free(g->rd->buf);
	g->rd->buf = NULL;
	rend_release(g->rd);
	free(g->rd);
	scene_release(&g->s);
	free(g);
}
