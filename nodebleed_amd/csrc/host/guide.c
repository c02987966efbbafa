/*
 * Learned path guiding, host side: parameter/optimiser buffers, network
 * initialisation, the per-batch training schedule used by the wavefront
 * path tracer, and test entry points on host buffers.  The kernels are in
 * kern/guide.hip; the model is described there.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_hip.h"
#include "librend.h"
#include "../ptlayout.h"

int rk_guide_setup(void *stream);
int rk_guide_mix(const struct PtArgs *a, const void *grid16, const void *w16,
                 const float *master, const void *rays, const void *hits,
                 const unsigned *counts, const void *insts, const void *tris,
                 const float *pn, unsigned npn, unsigned share, float *mix,
                 unsigned grid, void *stream);
int rk_guide_targets(void *rec, const unsigned *counts, const void *paths,
                     float *wsum, float *dense, unsigned cap, unsigned mode,
                     void *stream);
unsigned rk_guide_cdf_tiles(unsigned cap);
int rk_guide_cdf(const float *dense, const float *wsum, unsigned long long *tsum,
                 unsigned long long *cdf, const unsigned *counts, unsigned cap,
                 void *stream);
int rk_guide_train(const struct PtArgs *a, const void *grid16, const void *w16,
                   const float *master, const void *rec, const unsigned *counts,
                   unsigned cap, unsigned first, unsigned chunk,
                   const float *wsum, const unsigned long long *cdf, float *grad,
                   float *loss, float *part, unsigned grid, void *stream);
unsigned long long rk_guide_part_bytes(void);
int rk_guide_adam(float *master, float *grad, float *m1, float *m2,
                  void *grid16, void *w16, unsigned step, float lr_grid,
                  float lr_mlp, void *stream);
int rk_guide_grid_mirror(const float *master, void *grid16, void *stream);
int rk_guide_forward(const struct PtArgs *a, const void *grid16,
                     const void *w16, const float *master, const float *pn,
                     unsigned n, float *out, void *stream);
int rk_mfma_check(const float *A, const float *B, float *D, void *stream);
int rk_fill_f32(float *p, unsigned long long n, float v, void *stream);


static void dfree(void **p)
{
	if (*p)
		hipFree(*p);
	*p = NULL;
}

void rl_guide_release(struct rl_scene *s)
{
	struct rl_guide *g = s->guide;
	if (!g)
		return;
	dfree((void **)&g->master);
	dfree((void **)&g->grad);
	dfree((void **)&g->part);
	dfree((void **)&g->m1);
	dfree((void **)&g->m2);
	dfree(&g->grid16);
	dfree(&g->w16);
	dfree((void **)&g->mix);
	dfree(&g->rec);
	dfree((void **)&g->dense);
	dfree((void **)&g->cdf);
	dfree((void **)&g->tsum);
	dfree((void **)&g->scal);
	free(g);
	s->guide = NULL;
}

/* Deterministic initial parameters: small uniform grid features, Xavier-
 * uniform weights, zero biases except the output's log-concentration bias. */
static void init_params(float *p, uint32_t seed)
{
	uint64_t s = 0x9e3779b97f4a7c15ull ^ seed;
#define RND() ((s ^= s << 13, s ^= s >> 7, s ^= s << 17), \
	(double)(s >> 11) * (1.0 / 9007199254740992.0))
	for (size_t i = 0; i < GD_N_GRID; i++)
		p[i] = (float)((RND() * 2.0 - 1.0) * 1e-4);
	const unsigned ofs[4] = {GD_OFS_W1, GD_OFS_W2, GD_OFS_W3, GD_OFS_W4};
	const unsigned fin[4] = {GD_IN, GD_HID, GD_HID, GD_HID};
	const unsigned fout[4] = {GD_HID, GD_HID, GD_HID, GD_OUT};
	for (int l = 0; l < 4; l++) {
		double a = sqrt(6.0 / (fin[l] + fout[l]));
		for (unsigned i = 0; i < fin[l] * fout[l]; i++)
			p[ofs[l] + i] = (float)((RND() * 2.0 - 1.0) * a);
	}
	memset(&p[GD_OFS_B1], 0, GD_HID * sizeof(float));
	memset(&p[GD_OFS_B2], 0, GD_HID * sizeof(float));
	memset(&p[GD_OFS_B3], 0, GD_HID * sizeof(float));
	memset(&p[GD_OFS_B4], 0, GD_OUT * sizeof(float));
	for (int j = 0; j < GD_LOBES; j++)
		p[GD_OFS_B4 + 5 * j + 4] = 1.0f;  /* kappa ~ e */
#undef RND
}

/* Uploads parameters and refreshes the fp16/bf16 mirrors. */
static int upload_params(struct rl_guide *g, const float *p, hipStream_t st)
{
	hipError_t e = hipMemcpyAsync(g->master, p, GD_NPARAM * sizeof(float),
	  hipMemcpyHostToDevice, st);
	if (e)
		return rl_hipfail(e, "guide parameters");
	/* zero moments, then one Adam call with lr 0 and zero gradient writes
	 * the MLP mirrors (it leaves untouched grid entries alone) and a copy
	 * kernel the grid's */
	if (hipMemsetAsync(g->m1, 0, GD_NPARAM * sizeof(float), st) ||
	    hipMemsetAsync(g->m2, 0, GD_NPARAM * sizeof(float), st) ||
	    hipMemsetAsync(g->grad, 0, GD_NPARAM * sizeof(float), st) ||
	    hipMemsetAsync(g->w16, 0, GD_W16_ELEMS * 2, st))
		return rl_fail(REND_ENODEV, "guide state clear");
	int r = rk_guide_adam(g->master, g->grad, g->m1, g->m2, g->grid16, g->w16, 1,
	  0.0f, 0.0f, st);
	if (!r)
		r = rk_guide_grid_mirror(g->master, g->grid16, st);
	if (r)
		return rl_hipfail(r, "guide mirrors");
	g->step = 0;
	return 0;
}

static struct rl_guide *guide_alloc(struct rl_scene *s)
{
	if (s->guide)
		return s->guide;
	struct rl_guide *g = calloc(1, sizeof(*g));
	if (!g)
		return NULL;
	hipError_t e;
	if ((e = hipMalloc((void **)&g->master, GD_NPARAM * sizeof(float))) ||
	    (e = hipMalloc((void **)&g->grad, GD_NPARAM * sizeof(float))) ||
	    (e = hipMalloc((void **)&g->part, rk_guide_part_bytes())) ||
	    (e = hipMalloc((void **)&g->m1, GD_NPARAM * sizeof(float))) ||
	    (e = hipMalloc((void **)&g->m2, GD_NPARAM * sizeof(float))) ||
	    (e = hipMalloc(&g->grid16, GD_N_GRID * 2)) ||
	    (e = hipMalloc(&g->w16, GD_W16_ELEMS * 2)) ||
	    (e = hipMalloc((void **)&g->scal, 8 * sizeof(float)))) {
		s->guide = g;
		rl_guide_release(s);
		rl_hipfail(e, "guide buffers");
		return NULL;
	}
	s->guide = g;
	return g;
}

static void scene_box(struct rdata *rd, float mn[3], float inv[3])
{
	float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (unsigned i = 0; i < rd->instcnt; i++) {
		const float *a = &rd->aabbs[i].min.x, *b = &rd->aabbs[i].max.x;
		for (int k = 0; k < 3; k++) {
			lo[k] = a[k] < lo[k] ? a[k] : lo[k];
			hi[k] = b[k] > hi[k] ? b[k] : hi[k];
		}
	}
	for (int k = 0; k < 3; k++) {
		float ext = hi[k] - lo[k];
		float pad = 0.01f * ext + 1e-4f;
		mn[k] = lo[k] - pad;
		inv[k] = 1.0f / (ext + 2.0f * pad);
	}
}

int rend_hip_guide_enable(struct rdata *rd, float mix, unsigned int train_stride,
                          unsigned int steps, float lr, unsigned int seed)
{
	int r = rl_device();
	if (r)
		return r;
	struct rl_scene *s = rl_scene(rd);
	if (!s)
		return rl_fail(REND_ESTATE, "scene not prepared");
	if (mix < 0.0f || mix >= 1.0f)
		return rl_fail(REND_EINVAL, "guide mix %f (0 <= mix < 1)", (double)mix);
	struct rl_guide *g = guide_alloc(s);
	if (!g)
		return rl_fail(REND_ENOMEM, "guide");
	if ((r = rk_guide_setup(G.stream)))
		return rl_hipfail(r, "guide constants");
	float *p = malloc(GD_NPARAM * sizeof(float));
	if (!p)
		return rl_fail(REND_ENOMEM, "guide init");
	init_params(p, seed);
	r = upload_params(g, p, G.stream);
	free(p);
	if (r)
		return r;
	scene_box(rd, g->gmin, g->ginv);
	g->mix_prob = mix;
	g->stride = train_stride ? train_stride : 4;
	g->steps = steps;
	g->lr_grid = lr > 0.0f ? lr : 1e-2f;
	g->lr_mlp = GD_DEFAULT_MLP_LR_SCALE * g->lr_grid;
	g->chunk = GD_DEFAULT_CHUNK;
	g->importance = GD_DEFAULT_IMPORTANCE;
	g->depth = GD_DEFAULT_DEPTH;
	g->share = 1;
	g->target = GD_DEFAULT_TARGET;
	g->objective = GD_DEFAULT_OBJECTIVE;
	g->warm_frames = GD_DEFAULT_WARM_FRAMES;
	g->warm_steps = GD_DEFAULT_WARM_STEPS;
	g->frames = 0;
	g->enabled = 1;
	hipError_t e = hipStreamSynchronize(G.stream);
	return e ? rl_hipfail(e, "guide enable") : 0;
}

int rend_hip_guide_disable(struct rdata *rd)
{
	struct rl_scene *s = rl_scene(rd);
	if (s && s->guide)
		s->guide->enabled = 0;
	return 0;
}

int rend_hip_guide_set(struct rdata *rd, float mix, unsigned int steps)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	if (mix < 0.0f || mix >= 1.0f)
		return rl_fail(REND_EINVAL, "guide mix %f", (double)mix);
	s->guide->mix_prob = mix;
	s->guide->steps = steps;
	return 0;
}

int rend_hip_guide_training(struct rdata *rd, unsigned int records_per_step,
                            float lr_grid, float lr_mlp, int importance)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	if (records_per_step < 64 || records_per_step > (1u << 22))
		return rl_fail(REND_EINVAL, "records per step %u (64..4M)", records_per_step);
	if (!(lr_grid > 0.0f) || !(lr_mlp > 0.0f) || lr_grid > 1.0f || lr_mlp > 1.0f)
		return rl_fail(REND_EINVAL, "learning rates %g %g", (double)lr_grid,
		  (double)lr_mlp);
	struct rl_guide *g = s->guide;
	g->chunk = records_per_step;
	g->lr_grid = lr_grid;
	g->lr_mlp = lr_mlp;
	g->importance = importance != 0;
	return 0;
}

int rend_hip_guide_target(struct rdata *rd, int next_vertex_direct)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	s->guide->target = next_vertex_direct != 0;
	return 0;
}

int rend_hip_guide_objective(struct rdata *rd, int mis_density)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	s->guide->objective = mis_density != 0;
	return 0;
}

int rend_hip_guide_warmup(struct rdata *rd, unsigned int frames, unsigned int steps)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	if (steps > 4096)
		return rl_fail(REND_EINVAL, "warm-up steps %u (<= 4096)", steps);
	s->guide->warm_frames = frames;
	s->guide->warm_steps = steps;
	return 0;
}

/* Adam steps of the next frame: while training is on, the warm-up count for
 * the first warm_frames frames after enable (when it is larger), the
 * per-frame count after that. */
unsigned rl_guide_frame_steps(struct rl_guide *g)
{
	if (!g->steps)
		return 0;
	const unsigned n = g->frames < g->warm_frames && g->warm_steps > g->steps ?
	  g->warm_steps : g->steps;
	g->frames++;
	return n;
}

/* BSDF share of the density the training loss fits. */
static float train_bsdf(const struct rl_guide *g)
{
	return g->objective ? 1.0f - g->mix_prob : 0.0f;
}

int rend_hip_guide_sampling(struct rdata *rd, unsigned int depth, int share)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	if (depth < 1)
		return rl_fail(REND_EINVAL, "guided depth %u (>= 1)", depth);
	s->guide->depth = depth;
	/* pixels per shared depth-0 proposal: 0 off, 1 per pixel, k runs of k
	 * consecutive pixels of a row */
	s->guide->share = share > 0 ? (unsigned)share : 0u;
	return 0;
}

int rend_hip_guide_info(struct rdata *rd, float out[4])
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	float h[8];
	hipError_t e = hipMemcpy(h, s->guide->scal, sizeof(h), hipMemcpyDeviceToHost);
	if (e)
		return rl_hipfail(e, "guide info");
	out[0] = h[1];                  /* loss of the last training step */
	out[1] = (float)s->guide->last_records;
	out[2] = (float)s->guide->step;
	out[3] = (float)s->guide->enabled;
	return 0;
}

/* Training records a batch of cap paths can write: one per guided vertex
 * (depth < g->depth) of every stride-th path. */
static size_t rec_capacity(const struct rl_guide *g, size_t cap)
{
	return (cap / g->stride + 1) * (g->depth ? g->depth : 1) + 64;
}

/* Fills the guide fields of the frame constants. */
void rl_guide_args(struct rl_scene *s, struct PtArgs *a, size_t cap)
{
	struct rl_guide *g = s->guide;
	if (!g || !g->enabled)
		return;
	a->guide_mix = g->step > 0 ? g->mix_prob : 0.0f;
	memcpy(a->gmin, g->gmin, sizeof(a->gmin));
	memcpy(a->ginv, g->ginv, sizeof(a->ginv));
	a->train_stride = g->steps ? g->stride : 0;
	a->train_cap = (unsigned)rec_capacity(g, cap);
	a->guide_depth = g->depth;
	a->guide_share = 1;
	a->train_bsdf = train_bsdf(g);
}

/* Per-batch buffers: mixtures per queued path and training records (as
 * many as rl_guide_args allows for this batch size and stride). */
int rl_guide_buffers(struct rl_scene *s, size_t cap)
{
	struct rl_guide *g = s->guide;
	size_t rc = rec_capacity(g, cap);
	hipError_t e;
	if (cap > g->cap) {
		dfree((void **)&g->mix);
		g->cap = 0;
		if ((e = hipMalloc((void **)&g->mix, cap * GD_MIX * sizeof(float))))
			return rl_hipfail(e, "guide mixture buffer");
		g->cap = cap;
	}
	if (rc > g->rec_cap) {
		dfree(&g->rec);
		g->rec_cap = 0;
		if ((e = hipMalloc(&g->rec, rc * sizeof(struct GdRecord))))
			return rl_hipfail(e, "guide record buffer");
		g->rec_cap = rc;
	}
	if (rc > g->cdf_cap) {
		dfree((void **)&g->dense);
		dfree((void **)&g->cdf);
		dfree((void **)&g->tsum);
		g->cdf_cap = 0;
		if ((e = hipMalloc((void **)&g->dense, rc * sizeof(float))) ||
		    (e = hipMalloc((void **)&g->cdf, rc * sizeof(unsigned long long))) ||
		    (e = hipMalloc((void **)&g->tsum,
		      (rk_guide_cdf_tiles((unsigned)rc) + 1) * sizeof(unsigned long long))))
			return rl_hipfail(e, "guide record cdf");
		g->cdf_cap = rc;
	}
	return 0;
}

/* Mixture evaluation for the current queue (before shading). */
int rl_guide_eval(struct rl_scene *s, const struct PtArgs *a, const void *rays,
                  const void *hits, const unsigned *counts, unsigned share, hipStream_t st)
{
	struct rl_guide *g = s->guide;
	int r = rk_guide_mix(a, g->grid16, g->w16, g->master, rays, hits, counts,
	  s->insts, s->tris, NULL, 0, share, g->mix, (unsigned)G.cus * 4u, st);
	return r ? rl_hipfail(r, "guide eval") : 0;
}

/* After a batch: targets from the final path radiance, then `steps`
 * training steps of g->chunk records each (drawn in proportion to their
 * targets, or uniformly), each followed by an Adam update. */
int rl_guide_train(struct rl_scene *s, const struct PtArgs *a, const void *paths,
                   unsigned *counts, unsigned steps, hipStream_t st)
{
	struct rl_guide *g = s->guide;
	if (!steps)
		return 0;
	int r;
	unsigned rc = a->train_cap;
	if ((r = rk_fill_f32(g->scal, 2, 0.0f, st)))
		return rl_hipfail(r, "guide scalars");
	unsigned long long *cdf = g->importance ? g->cdf : NULL;
	if ((r = rk_guide_targets(g->rec, counts, paths, g->scal, cdf ? g->dense : NULL, rc,
	  (unsigned)g->target, st)))
		return rl_hipfail(r, "guide targets");
	if (cdf && (r = rk_guide_cdf(g->dense, g->scal, g->tsum, cdf, counts, rc, st)))
		return rl_hipfail(r, "guide record cdf");
	for (unsigned k = 0; k < steps; k++) {
		/* seed of the step's record sample */
		unsigned first = (g->step + 1u) * 0x9e3779b9u;
		if ((r = rk_guide_train(a, g->grid16, g->w16, g->master, g->rec, counts, rc,
		  first, g->chunk, g->scal, cdf, g->grad, g->scal + 1, g->part,
		  (unsigned)G.cus * 2u, st)))
			return rl_hipfail(r, "guide train");
		g->step++;
		if ((r = rk_guide_adam(g->master, g->grad, g->m1, g->m2, g->grid16, g->w16,
		  g->step, g->lr_grid, g->lr_mlp, st)))
			return rl_hipfail(r, "guide adam");
	}
	g->last_records = rc;
	return 0;
}

/* ---- test entry points on host buffers ----------------------------------- */

int rend_hip_guide_params(struct rdata *rd, float *params, int set)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	struct rl_guide *g = s->guide;
	if (set) {
		int r = upload_params(g, params, G.stream);
		if (r)
			return r;
	} else {
		hipError_t e = hipMemcpyAsync(params, g->master, GD_NPARAM * sizeof(float),
		  hipMemcpyDeviceToHost, G.stream);
		if (e)
			return rl_hipfail(e, "guide parameters");
	}
	hipError_t e = hipStreamSynchronize(G.stream);
	return e ? rl_hipfail(e, "guide parameters") : 0;
}

static void test_args(struct rl_guide *g, struct PtArgs *a)
{
	memset(a, 0, sizeof(*a));
	memcpy(a->gmin, g->gmin, sizeof(a->gmin));
	memcpy(a->ginv, g->ginv, sizeof(a->ginv));
	a->train_bsdf = train_bsdf(g);
}

int rend_hip_guide_forward_host(struct rdata *rd, const float *pn, unsigned int n,
                                float *out)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	struct rl_guide *g = s->guide;
	struct PtArgs a;
	test_args(g, &a);
	float *dpn = NULL, *dout = NULL;
	hipError_t e;
	int r = 0;
	if ((e = hipMalloc((void **)&dpn, (size_t)n * 6 * sizeof(float) + 4)) ||
	    (e = hipMalloc((void **)&dout, (size_t)n * GD_OUT * sizeof(float) + 4)) ||
	    (e = hipMemcpy(dpn, pn, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice))) {
		r = rl_hipfail(e, "guide forward buffers");
		goto out;
	}
	if ((r = rk_guide_forward(&a, g->grid16, g->w16, g->master, dpn, n, dout, G.stream))) {
		r = rl_hipfail(r, "guide forward");
		goto out;
	}
	if ((e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(out, dout, (size_t)n * GD_OUT * sizeof(float), hipMemcpyDeviceToHost)))
		r = rl_hipfail(e, "guide forward");
out:
	dfree((void **)&dpn);
	dfree((void **)&dout);
	return r;
}

/* Mixture records (GD_MIX floats each) of the inference kernel for n
 * position/normal pairs. */
int rend_hip_guide_mix_host(struct rdata *rd, const float *pn, unsigned int n, float *out)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	struct rl_guide *g = s->guide;
	struct PtArgs a;
	test_args(g, &a);
	float *dpn = NULL, *dout = NULL;
	hipError_t e;
	int r = 0;
	if ((e = hipMalloc((void **)&dpn, (size_t)n * 6 * sizeof(float) + 4)) ||
	    (e = hipMalloc((void **)&dout, (size_t)n * GD_MIX * sizeof(float) + 4)) ||
	    (e = hipMemcpy(dpn, pn, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice))) {
		r = rl_hipfail(e, "guide mixture buffers");
		goto out;
	}
	if ((r = rk_guide_mix(&a, g->grid16, g->w16, g->master, NULL, NULL, NULL, NULL, NULL,
	  dpn, n, 1, dout, (unsigned)G.cus * 4u, G.stream))) {
		r = rl_hipfail(r, "guide mixture");
		goto out;
	}
	if ((e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(out, dout, (size_t)n * GD_MIX * sizeof(float), hipMemcpyDeviceToHost)))
		r = rl_hipfail(e, "guide mixture");
out:
	dfree((void **)&dpn);
	dfree((void **)&dout);
	return r;
}

/* Gradient of the mean weighted loss over n records (GdRecord layout with
 * `target` set) into grad (GD_NPARAM floats; the hash-grid part, which the
 * device accumulates in bf16 pairs, expanded to fp32); loss receives the
 * loss. */
int rend_hip_guide_grad_host(struct rdata *rd, const void *records, unsigned int n,
                             float *grad, float *loss)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	struct rl_guide *g = s->guide;
	struct PtArgs a;
	test_args(g, &a);
	void *drec = NULL;
	float *dsc = NULL;
	hipError_t e;
	int r = 0;
	float sc[2] = {0.0f, 0.0f};
	const struct GdRecord *rr = records;
	for (unsigned i = 0; i < n; i++)
		sc[0] += rr[i].target;
	if ((e = hipMalloc(&drec, (size_t)n * sizeof(struct GdRecord) + 4)) ||
	    (e = hipMalloc((void **)&dsc, 2 * sizeof(float))) ||
	    (e = hipMemcpy(drec, records, (size_t)n * sizeof(struct GdRecord),
	      hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dsc, sc, sizeof(sc), hipMemcpyHostToDevice)) ||
	    (e = hipMemset(g->grad, 0, GD_NPARAM * sizeof(float)))) {
		r = rl_hipfail(e, "guide grad buffers");
		goto out;
	}
	if ((r = rk_guide_train(&a, g->grid16, g->w16, g->master, drec, NULL, n, 0, n, dsc,
	  NULL, g->grad, dsc + 1, g->part, (unsigned)((n + 15) / 16), G.stream))) {
		r = rl_hipfail(r, "guide grad");
		goto out;
	}
	if ((e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(grad, g->grad, GD_NPARAM * sizeof(float), hipMemcpyDeviceToHost)) ||
	    (e = hipMemcpy(sc, dsc, sizeof(sc), hipMemcpyDeviceToHost)) ||
	    (e = hipMemset(g->grad, 0, GD_NPARAM * sizeof(float))))
		r = rl_hipfail(e, "guide grad");
	if (!r) {
		/* the grid part holds one bf16 pair per entry in its first half:
		 * expand to fp32 features (from the end, in place) */
		for (size_t w = GD_N_GRID / 2; w-- > 0;) {
			uint32_t x;
			memcpy(&x, &grad[w], 4);
			uint32_t lo = x << 16, hi = x & 0xffff0000u;
			memcpy(&grad[2 * w], &lo, 4);
			memcpy(&grad[2 * w + 1], &hi, 4);
		}
	}
	if (loss)
		*loss = sc[1];
out:
	dfree(&drec);
	dfree((void **)&dsc);
	return r;
}

/* `steps` training steps on n host records with `target` set, as a frame's
 * training runs them (g->chunk records per step, drawn in proportion to
 * target or uniformly, then Adam); loss receives the last step's loss. */
int rend_hip_guide_train_host(struct rdata *rd, const void *records, unsigned int n,
                              unsigned int steps, float *loss)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->guide)
		return rl_fail(REND_ESTATE, "guiding not enabled");
	if (!n)
		return rl_fail(REND_EINVAL, "no records");
	struct rl_guide *g = s->guide;
	struct PtArgs a;
	test_args(g, &a);
	void *drec = NULL;
	float *dsc = NULL, *dd = NULL, *hd = NULL;
	unsigned long long *dcdf = NULL, *dts = NULL;
	unsigned *dcnt = NULL;
	hipError_t e;
	int r = 0;
	float sc[2] = {0.0f, 0.0f};
	unsigned cnt[PT_NCOUNTS] = {0};
	cnt[PT_CNT_TRAIN] = n;
	const struct GdRecord *rr = records;
	if (!(hd = malloc((size_t)n * sizeof(float))))
		return rl_fail(REND_ENOMEM, "guide train");
	for (unsigned i = 0; i < n; i++) {
		sc[0] += rr[i].target;
		hd[i] = rr[i].target;
	}
	if ((e = hipMalloc(&drec, (size_t)n * sizeof(struct GdRecord) + 4)) ||
	    (e = hipMalloc((void **)&dsc, 2 * sizeof(float))) ||
	    (e = hipMalloc((void **)&dcnt, sizeof(cnt))) ||
	    (e = hipMalloc((void **)&dd, (size_t)n * sizeof(float))) ||
	    (e = hipMalloc((void **)&dcdf, (size_t)n * sizeof(unsigned long long))) ||
	    (e = hipMalloc((void **)&dts,
	      (rk_guide_cdf_tiles(n) + 1) * sizeof(unsigned long long))) ||
	    (e = hipMemcpy(drec, records, (size_t)n * sizeof(struct GdRecord),
	      hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dcnt, cnt, sizeof(cnt), hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dd, hd, (size_t)n * sizeof(float), hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dsc, sc, sizeof(sc), hipMemcpyHostToDevice))) {
		r = rl_hipfail(e, "guide train buffers");
		goto out;
	}
	unsigned long long *cdf = g->importance ? dcdf : NULL;
	if (cdf && (r = rk_guide_cdf(dd, dsc, dts, cdf, dcnt, n, G.stream))) {
		r = rl_hipfail(r, "guide record cdf");
		goto out;
	}
	const unsigned chunk = n < g->chunk ? n : g->chunk;
	for (unsigned k = 0; k < steps; k++) {
		sc[1] = 0.0f;
		if ((e = hipMemcpyAsync(dsc, sc, sizeof(sc), hipMemcpyHostToDevice, G.stream))) {
			r = rl_hipfail(e, "guide train");
			goto out;
		}
		unsigned first = (g->step + 1u) * 0x9e3779b9u;
		if ((r = rk_guide_train(&a, g->grid16, g->w16, g->master, drec, dcnt, n, first,
		  chunk, dsc, cdf, g->grad, dsc + 1, g->part, (unsigned)G.cus * 2u, G.stream))) {
			r = rl_hipfail(r, "guide train");
			goto out;
		}
		g->step++;
		if ((r = rk_guide_adam(g->master, g->grad, g->m1, g->m2, g->grid16, g->w16,
		  g->step, g->lr_grid, g->lr_mlp, G.stream))) {
			r = rl_hipfail(r, "guide adam");
			goto out;
		}
	}
	if ((e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(sc, dsc, sizeof(sc), hipMemcpyDeviceToHost)))
		r = rl_hipfail(e, "guide train");
	if (loss)
		*loss = sc[1];
out:
	free(hd);
	dfree(&drec);
	dfree((void **)&dsc);
	dfree((void **)&dcnt);
	dfree((void **)&dd);
	dfree((void **)&dcdf);
	dfree((void **)&dts);
	return r;
}

/* Fixed-point prefix sums of n host values (scaled by 2^(60 - e), 2^e
 * above wsum) by the training-record cdf kernels. */
int rend_hip_guide_cdf_host(const float *values, unsigned int n, float wsum,
                            unsigned long long *cdf)
{
	int r = rl_device();
	if (r)
		return r;
	if (!n)
		return 0;
	float *dv = NULL, *dw = NULL;
	unsigned long long *dts = NULL, *dc = NULL;
	unsigned *dcnt = NULL;
	unsigned cnt[PT_NCOUNTS] = {0};
	cnt[PT_CNT_TRAIN] = n;
	hipError_t e;
	if ((e = hipMalloc((void **)&dv, (size_t)n * sizeof(float))) ||
	    (e = hipMalloc((void **)&dw, sizeof(float))) ||
	    (e = hipMalloc((void **)&dc, (size_t)n * sizeof(unsigned long long))) ||
	    (e = hipMalloc((void **)&dts,
	      (rk_guide_cdf_tiles(n) + 1) * sizeof(unsigned long long))) ||
	    (e = hipMalloc((void **)&dcnt, sizeof(cnt))) ||
	    (e = hipMemcpy(dv, values, (size_t)n * sizeof(float), hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dw, &wsum, sizeof(float), hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(dcnt, cnt, sizeof(cnt), hipMemcpyHostToDevice))) {
		r = rl_hipfail(e, "cdf buffers");
		goto out;
	}
	if ((r = rk_guide_cdf(dv, dw, dts, dc, dcnt, n, G.stream))) {
		r = rl_hipfail(r, "cdf");
		goto out;
	}
	if ((e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(cdf, dc, (size_t)n * sizeof(unsigned long long),
	      hipMemcpyDeviceToHost)))
		r = rl_hipfail(e, "cdf");
out:
	dfree((void **)&dv);
	dfree((void **)&dw);
	dfree((void **)&dc);
	dfree((void **)&dts);
	dfree((void **)&dcnt);
	return r;
}

int rend_hip_mfma_check(const float *A, const float *B, float *D)
{
	int r = rl_device();
	if (r)
		return r;
	float *d = NULL;
	hipError_t e;
	if ((e = hipMalloc((void **)&d, (16 * 32 + 32 * 16 + 16 * 16) * sizeof(float))) ||
	    (e = hipMemcpy(d, A, 16 * 32 * sizeof(float), hipMemcpyHostToDevice)) ||
	    (e = hipMemcpy(d + 512, B, 32 * 16 * sizeof(float), hipMemcpyHostToDevice))) {
		dfree((void **)&d);
		return rl_hipfail(e, "mfma check");
	}
	if ((r = rk_mfma_check(d, d + 512, d + 1024, G.stream)) ||
	    (e = hipStreamSynchronize(G.stream)) ||
	    (e = hipMemcpy(D, d + 1024, 256 * sizeof(float), hipMemcpyDeviceToHost)))
		r = r ? rl_hipfail(r, "mfma check") : rl_hipfail(e, "mfma check");
	dfree((void **)&d);
	return r;
}
