/*
 * Wavefront path tracing on the host side: light list, path/queue buffers,
 * the per-frame launch sequence (generate, then per segment: trace, shade,
 * occlusion, shadow accumulation, queue advance), batched views and the
 * launch-shape autotuner.
 *
 * The estimator follows the reference's trace3 (rend.c:3205-3291); the
 * kernels are in kern/pt.hip.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_hip.h"
#include "librend.h"
#include "../ptlayout.h"

int rk_pt_gen(const struct PtArgs *a, const void *views, void *paths, void *rays,
              void *qpath, unsigned *counts, unsigned first, unsigned n,
              void *stream);
int rk_pt_shade(const struct PtArgs *a, void *paths, const void *rays,
                const void *hits, const void *qpath, void *nrays, void *nqpath,
                void *srays, void *spath, float *scontrib, unsigned *counts,
                const void *insts, const void *tris, const void *nrms,
                const void *mtls, const void *lights, const float *lcdf,
                const float *guide, void *train, unsigned depth, unsigned grid,
                void *stream);
int rk_pt_shadow(void *paths, const void *spath, const float *scontrib,
                 const void *occ, const unsigned *counts, void *rec,
                 unsigned grid, void *stream);
int rk_pt_advance(unsigned *counts, void *stream);
int rk_pt_accum(const void *paths, unsigned n, float *acc, void *stream);
int rk_pt_resolve(const float *acc, unsigned *buf, unsigned npx, float invs,
                  void *stream);

struct rl_pt {
	size_t    cap;                       /* allocated paths */
	size_t    batch;                     /* paths per launch batch */
	void     *paths, *rays[2], *q[2], *hits, *srays, *spath, *occ;
	float    *scontrib;
	unsigned *counts;
	void     *views;
	unsigned  views_cap;
	void     *lights;
	float    *lcdf;
	unsigned  nlights;
	int       lights_ok;
	unsigned long long lights_gen;  /* scene upload generation they follow */
};

static struct pt_params {
	unsigned maxdepth, rr_depth, batch, shade_mult;
	float    guide_mix;
} P = {4, 3, 0, 4, 0.0f};

int rend_hip_pt_config(unsigned int maxdepth, unsigned int rr_depth,
                       unsigned int batch)
{
	if (maxdepth < 1 || maxdepth > 64)
		return rl_fail(REND_EINVAL, "maxdepth %u", maxdepth);
	P.maxdepth = maxdepth;
	P.rr_depth = rr_depth ? rr_depth : 1000;
	P.batch = batch;
	return 0;
}

static void dfree(void **p)
{
	if (*p)
		hipFree(*p);
	*p = NULL;
}

static void graph_drop(void);

void rl_pt_release(struct rl_scene *s)
{
	struct rl_pt *t = s->pt;
	if (!t)
		return;
	/* the cached frame graph may point at this scene's buffers */
	graph_drop();
	dfree(&t->paths);
	dfree(&t->rays[0]);
	dfree(&t->rays[1]);
	dfree(&t->q[0]);
	dfree(&t->q[1]);
	dfree(&t->hits);
	dfree(&t->srays);
	dfree(&t->spath);
	dfree(&t->occ);
	dfree((void **)&t->scontrib);
	dfree((void **)&t->counts);
	dfree(&t->views);
	dfree(&t->lights);
	dfree((void **)&t->lcdf);
	free(t);
	s->pt = NULL;
}

static int pt_buffers(struct rl_scene *s, size_t cap)
{
	if (!s->pt) {
		s->pt = calloc(1, sizeof(*s->pt));
		if (!s->pt)
			return rl_fail(REND_ENOMEM, "path state");
	}
	struct rl_pt *t = s->pt;
	if (cap <= t->cap)
		return 0;
	graph_drop();
	dfree(&t->paths);
	dfree(&t->rays[0]);
	dfree(&t->rays[1]);
	dfree(&t->q[0]);
	dfree(&t->q[1]);
	dfree(&t->hits);
	dfree(&t->srays);
	dfree(&t->spath);
	dfree(&t->occ);
	dfree((void **)&t->scontrib);
	t->cap = 0;
	hipError_t e;
	if ((e = hipMalloc(&t->paths, cap * sizeof(struct PtPath))) ||
	    (e = hipMalloc(&t->rays[0], cap * sizeof(struct rray))) ||
	    (e = hipMalloc(&t->rays[1], cap * sizeof(struct rray))) ||
	    (e = hipMalloc(&t->q[0], cap * sizeof(unsigned))) ||
	    (e = hipMalloc(&t->q[1], cap * sizeof(unsigned))) ||
	    (e = hipMalloc(&t->hits, cap * sizeof(struct rhit))) ||
	    (e = hipMalloc(&t->srays, cap * sizeof(struct rray))) ||
	    (e = hipMalloc(&t->spath, cap * sizeof(unsigned))) ||
	    (e = hipMalloc(&t->occ, cap)) ||
	    (e = hipMalloc((void **)&t->scontrib, cap * 3 * sizeof(float))))
		return rl_hipfail(e, "path buffers");
	if (!t->counts &&
	    (e = hipMalloc((void **)&t->counts, PT_NCOUNTS * sizeof(unsigned))))
		return rl_hipfail(e, "path counters");
	t->cap = cap;
	return 0;
}

/* Emissive triangles in world space with an area * luminance CDF.  World
 * vertices come from inverting the 3x4 globinv in double.  Rebuilt after
 * every upload of scene data (rend_prepstatic, rend_prepdynamic): moving an
 * emissive instance moves its triangles here too. */
static int pt_lights(struct rl_scene *s)
{
	struct rl_pt *t = s->pt;
	if (t->lights_ok && t->lights_gen == s->gen)
		return 0;
	graph_drop();
	struct rdata *rd = s->rd;
	unsigned cap = 64, n = 0;
	struct PtLight *L = malloc(cap * sizeof(*L));
	if (!L)
		return rl_fail(REND_ENOMEM, "lights");
	for (unsigned i = 0; i < rd->instcnt; i++) {
		const struct rinst *ri = &rd->insts[i];
		const float *g = ri->globinv;
		double a[9] = {g[0], g[1], g[2], g[4], g[5], g[6], g[8], g[9], g[10]};
		double det = a[0] * (a[4] * a[8] - a[5] * a[7]) -
		  a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
		if (det == 0.0)
			continue;
		double m[9] = {
		  (a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
		  (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
		  (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
		  (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
		  (a[0] * a[4] - a[1] * a[3]) / det};
		double tr[3] = {g[3], g[7], g[11]};
		double off[3] = {-(m[0] * tr[0] + m[1] * tr[1] + m[2] * tr[2]),
		  -(m[3] * tr[0] + m[4] * tr[1] + m[5] * tr[2]),
		  -(m[6] * tr[0] + m[7] * tr[1] + m[8] * tr[2])};
		for (unsigned k = 0; k < ri->tricnt; k++) {
			const struct rnrm *rn = &rd->nrms[ri->triofs + k];
			const struct rmtl *mt = &rd->mtls[rn->mtlid];
			if (!(mt->flags & RFLAG_EMISSIVE))
				continue;
			const struct rtri *tr3 = &rd->tris[ri->triofs + k];
			const struct vec3 *v[3] = {&tr3->v0, &tr3->v1, &tr3->v2};
			double w[3][3];
			for (int q = 0; q < 3; q++) {
				double p[3] = {v[q]->x, v[q]->y, v[q]->z};
				for (int r = 0; r < 3; r++)
					w[q][r] = m[3 * r] * p[0] + m[3 * r + 1] * p[1] +
					  m[3 * r + 2] * p[2] + off[r];
			}
			double e0[3], e1[3], c[3];
			for (int r = 0; r < 3; r++) {
				e0[r] = w[1][r] - w[0][r];
				e1[r] = w[2][r] - w[0][r];
			}
			c[0] = e0[1] * e1[2] - e0[2] * e1[1];
			c[1] = e0[2] * e1[0] - e0[0] * e1[2];
			c[2] = e0[0] * e1[1] - e0[1] * e1[0];
			double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
			if (len <= 0.0)
				continue;
			if (n == cap) {
				cap *= 2;
				struct PtLight *nl = realloc(L, cap * sizeof(*L));
				if (!nl) {
					free(L);
					return rl_fail(REND_ENOMEM, "lights");
				}
				L = nl;
			}
			struct PtLight *l = &L[n++];
			memset(l, 0, sizeof(*l));
			for (int r = 0; r < 3; r++) {
				l->v0[r] = (float)w[0][r];
				l->e0[r] = (float)e0[r];
				l->e1[r] = (float)e1[r];
				l->n[r] = (float)(c[r] / len);
			}
			l->E[0] = mt->col.x;
			l->E[1] = mt->col.y;
			l->E[2] = mt->col.z;
			l->area = (float)(0.5 * len);
			l->pick = (float)(0.5 * len * (0.2126 * mt->col.x +
			  0.7152 * mt->col.y + 0.0722 * mt->col.z));
		}
	}
	double tot = 0.0;
	for (unsigned i = 0; i < n; i++)
		tot += L[i].pick;
	float *cdf = malloc((n ? n : 1) * sizeof(float));
	if (!cdf) {
		free(L);
		return rl_fail(REND_ENOMEM, "light cdf");
	}
	double acc = 0.0;
	for (unsigned i = 0; i < n; i++) {
		acc += L[i].pick;
		L[i].pick = (float)(L[i].pick / tot);
		cdf[i] = (float)(acc / tot);
	}
	if (n)
		cdf[n - 1] = 1.0f;
	hipError_t e = hipSuccess;
	dfree(&t->lights);
	dfree((void **)&t->lcdf);
	if (n) {
		if ((e = hipMalloc(&t->lights, n * sizeof(*L))) ||
		    (e = hipMalloc((void **)&t->lcdf, n * sizeof(float))) ||
		    (e = hipMemcpy(t->lights, L, n * sizeof(*L), hipMemcpyHostToDevice)) ||
		    (e = hipMemcpy(t->lcdf, cdf, n * sizeof(float), hipMemcpyHostToDevice))) {
			free(L);
			free(cdf);
			return rl_hipfail(e, "light upload");
		}
	}
	free(L);
	free(cdf);
	t->nlights = n;
	t->lights_ok = 1;
	t->lights_gen = s->gen;
	return 0;
}

/* The path tracer's traversal divides by the determinant exactly (IEEE
 * 1/det), whatever rend_hip_set_rcp selects for the reference-exact entry
 * points: a hit position p = o + t d feeds the next segment's origin, and
 * with the host's 12-bit RCPSS reciprocal t carries a relative error of up
 * to 3.7e-4, which at the instanced scene's distances (~100 units) puts the
 * origin up to 0.04 below the surface, 40 times the 1e-3 spawn offset. */
#define PT_RCP_IEEE 1

static unsigned shade_grid(void)
{
	return (unsigned)G.cus * P.shade_mult;
}

int rk_fill_f32(float *p, unsigned long long n, float v, void *stream);
static void pt_args(struct rl_scene *s, struct PtArgs *a, unsigned nviews,
                    unsigned seed);

/* Launch sequence of one frame over `total` paths of the views already on
 * the device at t->views: per batch of at most t->cap paths, generate, then
 * per segment trace / shade / occlusion / shadow accumulation / advance,
 * then accumulate.  Every count lives on the device, so the sequence has no
 * host synchronisation and can be captured as a graph. */
static int pt_launches(struct rl_scene *s, const struct PtArgs *a,
                       unsigned long long total, float *acc, hipStream_t st)
{
	struct rl_pt *t = s->pt;
	size_t cap = t->batch;
	int r;
	unsigned cgrid, agrid;
	if ((r = rl_trace_grid(1, 0, PT_RCP_IEEE, &cgrid)) ||
	    (r = rl_trace_grid(0, 0, PT_RCP_IEEE, &agrid)))
		return r;
	if ((r = rk_fill_f32((float *)&t->counts[4], 2, 0.0f, st)))
		return rl_hipfail(r, "ray counter");
	const int guided = s->guide && s->guide->enabled;
	if (guided && (r = rl_guide_buffers(s, cap)))
		return r;
	float *mix = guided && a->guide_mix > 0.0f ? s->guide->mix : NULL;
	void *rec = guided && a->train_stride ? s->guide->rec : NULL;
	/* the frame's training steps are spread over its batches */
	const unsigned long long nb = (total + cap - 1) / cap;
	const unsigned steps = rec ? rl_guide_frame_steps(s->guide) : 0;
	unsigned long long b = 0;
	for (unsigned long long first = 0; first < total; first += cap, b++) {
		unsigned n = (unsigned)(total - first < cap ? total - first : cap);
		if ((r = rk_pt_gen(a, t->views, t->paths, t->rays[0], t->q[0], t->counts,
		  (unsigned)first, n, st)))
			return rl_hipfail(r, "path generation");
		int cur = 0;
		for (unsigned d = 0; d < a->maxdepth; d++) {
			unsigned slot = G.ctl_next++ % RL_CTL_SLOTS;
			r = rk_trace_launch(s->arena, s->tlas, s->insts, t->rays[cur], t->hits,
			  NULL, n, G.ctl + RK_CTL_WORDS * slot, G.spill, NULL, 1, 0, PT_RCP_IEEE, G.variant, cgrid,
			  G.refill, G.max_iter, &t->counts[PT_CNT_CUR], G.diag, rl_trace_flags(RK_TRACE_NUDGE), st);
			if (r)
				return rl_hipfail(r, "extend trace");
			if (mix && d < a->guide_depth && (r = rl_guide_eval(s, a, t->rays[cur], t->hits,
			  t->counts, d == 0 ? a->guide_share : 1u, st)))
				return r;
			if ((r = rk_pt_shade(a, t->paths, t->rays[cur], t->hits, t->q[cur],
			  t->rays[cur ^ 1], t->q[cur ^ 1], t->srays, t->spath, t->scontrib,
			  t->counts, s->insts, s->tris, s->nrms, s->mtls, t->lights, t->lcdf,
			  mix, rec, d, shade_grid(), st)))
				return rl_hipfail(r, "shade");
			slot = G.ctl_next++ % RL_CTL_SLOTS;
			r = rk_trace_launch(s->arena, s->tlas, s->insts, t->srays, NULL, t->occ,
			  n, G.ctl + RK_CTL_WORDS * slot, G.spill, NULL, 0, 0, PT_RCP_IEEE, G.variant, agrid, G.refill,
			  G.max_iter, &t->counts[PT_CNT_SHADOW], G.diag, rl_trace_flags(RK_TRACE_NUDGE), st);
			if (r)
				return rl_hipfail(r, "shadow trace");
			G.last_slot = slot;
			if ((r = rk_pt_shadow(t->paths, t->spath, t->scontrib, t->occ, t->counts,
			  rec, shade_grid(), st)))
				return rl_hipfail(r, "shadow accumulate");
			if ((r = rk_pt_advance(t->counts, st)))
				return rl_hipfail(r, "advance");
			cur ^= 1;
		}
		if ((r = rk_pt_accum(t->paths, n, acc, st)))
			return rl_hipfail(r, "accumulate");
		const unsigned sb = (unsigned)((steps * (b + 1)) / nb - (steps * b) / nb);
		if (rec && sb && (r = rl_guide_train(s, a, t->paths, t->counts, sb, st)))
			return r;
	}
	return 0;
}

/* FNV-1a over the launch-relevant state: a graph is replayed only while
 * this key is unchanged (sample_base lives in device memory and is not part
 * of it). */
static unsigned long long pt_key(const struct rl_scene *s, const struct PtArgs *a,
                                 const struct PtView *v, unsigned nv, float *acc,
                                 hipStream_t st)
{
	unsigned long long h = 1469598103934665603ull;
#define MIX(p, n) do { const unsigned char *b_ = (const unsigned char *)(p); \
	for (size_t i_ = 0; i_ < (n); i_++) { h ^= b_[i_]; h *= 1099511628211ull; } } while (0)
	MIX(a, sizeof(*a));
	for (unsigned i = 0; i < nv; i++) {
		struct PtView c = v[i];
		c.sample_base = 0;
		MIX(&c, sizeof(c));
	}
	/* every device pointer the captured kernels receive */
	const struct rl_pt *t = s->pt;
	const void *ptrs[] = {s, acc, st, s->arena, s->insts, s->tris, s->nrms, s->mtls,
	  t->paths, t->views, t->rays[0], t->rays[1], t->q[0], t->q[1], t->hits, t->srays,
	  t->spath, t->occ, t->scontrib, t->counts, t->lights, t->lcdf, G.ctl, G.diag,
	  G.spill};
	MIX(ptrs, sizeof(ptrs));
	unsigned long long misc[] = {s->tlas, t->batch, G.refill, (unsigned long long)PT_RCP_IEEE,
	  P.shade_mult, (unsigned long long)G.variant, G.grid_mult, s->gen, t->lights_gen,
	  G.spill_bytes, t->cap, G.step_batch};
	MIX(misc, sizeof(misc));
#undef MIX
	return h;
}

struct pt_graph {
	hipGraphExec_t      exec;
	unsigned long long  key;
	int                 valid;
};
static struct pt_graph g_graph;
static int g_use_graph = 1;

int rend_hip_pt_graphs(int on)
{
	g_use_graph = on != 0;
	return 0;
}

static void graph_drop(void)
{
	if (g_graph.valid)
		hipGraphExecDestroy(g_graph.exec);
	g_graph.valid = 0;
}

/* Uploads the view table and runs (or replays) the frame. */
static int pt_frame(struct rl_scene *s, const struct PtView *views, unsigned nv,
                    float *acc, hipStream_t st, unsigned long long *rays)
{
	unsigned long long total = 0;
	for (unsigned i = 0; i < nv; i++)
		total += (unsigned long long)views[i].w * views[i].h * views[i].spp;
	size_t cap = P.batch ? P.batch :
	  (size_t)(total < (1ull << 22) ? total : (1ull << 22));
	int r;
	if ((r = pt_buffers(s, cap ? cap : 1)) || (r = pt_lights(s)))
		return r;
	struct rl_pt *t = s->pt;
	t->batch = cap ? cap : 1;
	if (nv > t->views_cap) {
		dfree(&t->views);
		hipError_t e = hipMalloc(&t->views, nv * sizeof(struct PtView));
		if (e)
			return rl_hipfail(e, "views");
		t->views_cap = nv;
	}
	if ((r = rl_stream_order(st)))
		return r;
	hipError_t e = hipMemcpyAsync(t->views, views, nv * sizeof(*views),
	  hipMemcpyHostToDevice, st);
	if (e)
		return rl_hipfail(e, "view upload");
	struct PtArgs a;
	pt_args(s, &a, nv, 0x68bc21ebu);
	/* one view whose pixels' samples are consecutive paths of one batch:
	 * its depth-0 proposal is evaluated once per pixel, or once per run of
	 * `share` pixels of a row */
	if (a.guide_mix > 0.0f && s->guide && s->guide->share && nv == 1) {
		const unsigned k = views[0].w % s->guide->share ? 1u : s->guide->share;
		const unsigned long long run = (unsigned long long)views[0].spp * k;
		if (run > 1 && t->batch % run == 0)
			a.guide_share = (unsigned)run;
	}
	/* allocations (spill area) happen before the key and any capture */
	unsigned g0;
	if ((r = rl_trace_grid(1, 0, PT_RCP_IEEE, &g0)) || (r = rl_trace_grid(0, 0, PT_RCP_IEEE, &g0)))
		return r;
	unsigned long long key = pt_key(s, &a, views, nv, acc, st);
	/* Guided frames change the network between calls and run host-scheduled
	 * training: launched directly. */
	if (g_use_graph && !(s->guide && s->guide->enabled)) {
		if (!g_graph.valid || g_graph.key != key) {
			graph_drop();
			hipGraph_t gr;
			e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
			if (e) {
				return rl_hipfail(e, "graph capture");
			}
			r = pt_launches(s, &a, total, acc, st);
			hipError_t e2 = hipStreamEndCapture(st, &gr);
			if (r || e2) {
				if (!e2)
					hipGraphDestroy(gr);
				return r ? r : rl_hipfail(e2, "graph capture end");
			}
			e = hipGraphInstantiate(&g_graph.exec, gr, NULL, NULL, 0);
			hipGraphDestroy(gr);
			if (e) {
				return rl_hipfail(e, "graph instantiate");
			}
			g_graph.key = key;
			g_graph.valid = 1;
		}
		e = hipGraphLaunch(g_graph.exec, st);
		if (e) {
			return rl_hipfail(e, "graph launch");
		}
	} else if ((r = pt_launches(s, &a, total, acc, st))) {
		return r;
	}
	if ((r = rl_stream_mark(st)))
		return r;
	if (rays) {
		unsigned long long n = 0;
		e = hipMemcpyAsync(&n, &t->counts[4], sizeof(n), hipMemcpyDeviceToHost, st);
		if (!e)
			e = hipStreamSynchronize(st);
		if (e)
			return rl_hipfail(e, "ray count");
		*rays = n;
	}
	return 0;
}

/* Camera constants of a rend_view (upd_rcam, main.c:159-195, with the
 * camera basis of calc_cam, scene.c:401-409). */
static void view_cam(const struct rend_view *v, struct CamArgs *c)
{
	double f[3] = {(double)v->target[0] - v->eye[0], (double)v->target[1] - v->eye[1],
	  (double)v->target[2] - v->eye[2]};
	double l = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
	for (int k = 0; k < 3; k++)
		f[k] /= l;
	double ri[3] = {-f[2], 0.0, f[0]};  /* fwd x (0,1,0) */
	l = sqrt(ri[0] * ri[0] + ri[2] * ri[2]);
	for (int k = 0; k < 3; k++)
		ri[k] /= l;
	double up[3] = {ri[1] * f[2] - ri[2] * f[1], ri[2] * f[0] - ri[0] * f[2],
	  ri[0] * f[1] - ri[1] * f[0]};
	float tanv = tanf((float)(0.5f * v->vfov * 3.14159265 / 180.0f));
	float tanf_ = tanf((float)(0.5f * v->focangle * 3.14159265 / 180.0f));
	for (int k = 0; k < 3; k++) {
		c->eye[k] = v->eye[k];
		c->fwd[k] = (float)f[k];
		c->ri[k] = (float)ri[k];
		c->up[k] = (float)up[k];
	}
	c->rw = 1.0f / (float)v->w;
	c->rh = 1.0f / (float)v->h;
	c->aspect = (float)v->w / (float)v->h;
	c->fdist = v->focdist;
	c->fovfdist = 2.0f * tanv * v->focdist;
	c->focrad = tanf_ * v->focdist;
	c->dof = tanf_ > 0.0f;
}

/* FNV-1a over a view's camera and resolution (not its spp): the view's
 * random stream. */
static uint32_t view_stream(const struct rend_view *v)
{
	struct rend_view c = *v;
	c.spp = 0;
	uint32_t h = 2166136261u;
	const unsigned char *b = (const unsigned char *)&c;
	for (size_t i = 0; i < sizeof(c); i++) {
		h ^= b[i];
		h *= 16777619u;
	}
	return h;
}

static int views_table(const struct rend_view *in, unsigned nv, unsigned sample_base,
                       struct PtView **out)
{
	struct PtView *v = calloc(nv ? nv : 1, sizeof(*v));
	if (!v)
		return rl_fail(REND_ENOMEM, "view table");
	unsigned long long first = 0, acc = 0;
	for (unsigned i = 0; i < nv; i++) {
		if (!in[i].w || !in[i].h || !in[i].spp) {
			free(v);
			return rl_fail(REND_EINVAL, "view %u: empty", i);
		}
		view_cam(&in[i], &v[i].cam);
		v[i].w = in[i].w;
		v[i].h = in[i].h;
		v[i].spp = in[i].spp;
		v[i].first = (unsigned)first;
		v[i].acc_ofs = (unsigned)acc;
		v[i].sample_base = sample_base;
		v[i].pitch = in[i].w;
		v[i].stream = view_stream(&in[i]);
		first += (unsigned long long)in[i].w * in[i].h * in[i].spp;
		acc += (unsigned long long)in[i].w * in[i].h;
	}
	if (first >= (1ull << 32) || acc >= (1ull << 30)) {
		free(v);
		return rl_fail(REND_EINVAL, "batch too large (%llu paths)", first);
	}
	*out = v;
	return 0;
}

int rend_hip_render_views(struct rdata *rd, const struct rend_view *views,
                          unsigned int nviews, unsigned int sample_base,
                          float *d_acc, void *stream, unsigned long long *rays)
{
	struct rl_scene *s = rl_scene(rd);
	int r = rl_device();
	if (r)
		return r;
	if (!s || !s->tlas_built)
		return rl_fail(REND_ESTATE, "scene not prepared");
	if (!nviews || !views || !d_acc)
		return rl_fail(REND_EINVAL, "no views");
	if ((r = rl_sync_scene(s)))
		return r;
	struct PtView *v;
	if ((r = views_table(views, nviews, sample_base, &v)))
		return r;
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	r = pt_frame(s, v, nviews, d_acc, st, rays);
	free(v);
	return r;
}

int rend_hip_autotune(struct rdata *rd, const struct rend_view *views,
                      unsigned int nviews, float *d_acc, float out[5])
{
	static const unsigned refills[] = {16, 32, 48};
	unsigned long long total = 0;
	for (unsigned i = 0; i < nviews; i++)
		total += (unsigned long long)views[i].w * views[i].h * views[i].spp;
	unsigned batches[5] = {1u << 18, 1u << 20, 1u << 21, 1u << 22, 0};
	struct pt_params keep = P;
	unsigned keep_refill = G.refill;
	int keep_variant = G.variant;
	float best = 1e30f, dflt = -1.0f;
	unsigned bb = P.batch, br = G.refill;
	int bv = G.variant;
	void *e0 = rend_hip_event_create(), *e1 = rend_hip_event_create();
	if (!e0 || !e1)
		return rl_fail(REND_ENODEV, "events");
	int r = 0;
	/* one timed call (after a capturing call) per candidate */
#define TIME_ONE(ms_) do { \
		if ((r = rend_hip_render_views(rd, views, nviews, 0, d_acc, NULL, NULL))) \
			break; \
		rend_hip_event_record(e0, NULL); \
		r = rend_hip_render_views(rd, views, nviews, 0, d_acc, NULL, NULL); \
		rend_hip_event_record(e1, NULL); \
		ms_ = rend_hip_event_elapsed(e0, e1); \
	} while (0)
	/* traversal variant at the default batch and refill */
	P.batch = 0;
	G.refill = 48;
	for (int v = 0; v < rk_trace_variants() && !r; v++) {
		float ms = -1.0f;
		G.variant = v;
		TIME_ONE(ms);
		if (!r && v == 0)
			dflt = ms;
		if (!r && ms > 0.0f && ms < best) {
			best = ms;
			bv = v;
			bb = 0;
			br = 48;
		}
	}
	/* then batch x refill with that variant */
	G.variant = bv;
	for (unsigned bi = 0; bi < 5 && !r; bi++) {
		if (batches[bi] && batches[bi] >= total * 2 && bi > 0)
			continue;
		for (unsigned ri = 0; ri < 3 && !r; ri++) {
			if (batches[bi] == 0 && refills[ri] == 48)
				continue;  /* timed above */
			float ms = -1.0f;
			P.batch = batches[bi];
			G.refill = refills[ri];
			TIME_ONE(ms);
			if (!r && ms > 0.0f && ms < best) {
				best = ms;
				bb = P.batch;
				br = G.refill;
			}
		}
	}
#undef TIME_ONE
	rend_hip_event_destroy(e0);
	rend_hip_event_destroy(e1);
	P = keep;
	if (r) {
		G.refill = keep_refill;
		G.variant = keep_variant;
		return r;
	}
	P.batch = bb;
	G.refill = br;
	G.variant = bv;
	graph_drop();
	if (out) {
		out[0] = (float)bb;
		out[1] = (float)br;
		out[2] = best;
		out[3] = dflt;
		out[4] = (float)bv;
	}
	return 0;
}

static void pt_args(struct rl_scene *s, struct PtArgs *a, unsigned nviews,
                    unsigned seed)
{
	struct rdata *rd = s->rd;
	memset(a, 0, sizeof(*a));
	a->nviews = nviews;
	a->nlights = s->pt->nlights;
	a->maxdepth = P.maxdepth;
	a->rr_depth = P.rr_depth;
	a->bg[0] = rd->bgcol.x;
	a->bg[1] = rd->bgcol.y;
	a->bg[2] = rd->bgcol.z;
	a->seed = seed;
	a->guide_mix = P.guide_mix;
	rl_guide_args(s, a, s->pt->batch);
}

int rl_render_pt(struct rl_scene *s, unsigned rw, unsigned rh, unsigned *rays)
{
	struct rdata *rd = s->rd;
	hipStream_t st = G.stream;
	int r = rl_acc_ready(s, st);
	if (r)
		return r;
	struct PtView v;
	memset(&v, 0, sizeof(v));
	rl_camargs(&rd->cam, &v.cam);
	v.w = rw;
	v.h = rh;
	v.spp = rd->spp ? rd->spp : 1;
	v.sample_base = rd->samples;
	v.pitch = rd->width;
	unsigned long long n = 0;
	if ((r = pt_frame(s, &v, 1, s->acc, st, &n)))
		return r;
	float inv = 1.0f / (float)(rd->spp + rd->samples);
	if ((r = rk_pt_resolve(s->acc, s->buf, rd->width * rd->height, inv, st)))
		return rl_hipfail(r, "resolve");
	*rays = (unsigned)(n < 0xffffffffull ? n : 0xffffffffull);
	return 0;
}
