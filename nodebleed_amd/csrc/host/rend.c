/*
 * librend host library (C11): the reference renderer's API (rend.h:149-162)
 * over one HIP device, plus the ray-buffer entry points of rend_hip.h.
 *
 * Ownership follows the reference (rend.c:3000-3107): rend_init allocates
 * the host arrays the caller fills (mtls, tris, nrms, insts, aabbs) and the
 * host BVH arena; the renderer owns acc, the application owns buf.  The HIP
 * layer owns the device mirrors: rend_prepstatic builds every BLAS on the
 * host, repacks it into the device layout (dlayout.h) and uploads it with the
 * normals and materials; rend_prepdynamic builds the TLAS and uploads it with
 * the instance records.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_hip.h"
#include "librend.h"
#include "../dlayout.h"
#define RCP_TABLE_NO_HOST_ARRAY
#include "../rcp_table.h"

/* ---- errors -------------------------------------------------------------- */

static int g_err;
static char g_errmsg[256];

int rl_fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
	va_end(ap);
	g_err = code;
	return code;
}

int rl_hipfail(int herr, const char *what)
{
	return rl_fail(REND_ENODEV, "%s: %s", what,
	  hipGetErrorString((hipError_t)herr));
}

int rend_hip_lasterror(void) { return g_err; }
const char *rend_hip_errstr(void) { return g_errmsg; }
/* This host's RCPSS/RSQRTPS tables, captured once (approx.c). */
static pthread_once_t g_approx_once = PTHREAD_ONCE_INIT;

static void approx_capture(void)
{
	G.approx_bad = rl_approx_capture(G.rcp_host, G.rsqrt_host, G.approx_cpu,
	  sizeof(G.approx_cpu));
}

void rl_approx_ready(void)
{
	pthread_once(&g_approx_once, approx_capture);
}

const char *rend_hip_rcp_cpu(void)
{
	static char desc[256];
	if (G.rcp_mode == REND_RCP_XEON_TABLE)
		return RCP_TABLE_CPU;
	if (G.rcp_mode == REND_RCP_IEEE)
		return "IEEE (correctly rounded)";
	rl_approx_ready();
	snprintf(desc, sizeof(desc), "%s%s", G.approx_cpu,
	  G.approx_bad ? " (table rule failed: not bit-exact)" : "");
	return desc;
}

/* ---- process-global device state ----------------------------------------- */

struct rl_global G;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

int rl_device(void)
{
	if (G.ready)
		return 0;
	pthread_mutex_lock(&g_lock);
	if (G.ready) {
		pthread_mutex_unlock(&g_lock);
		return 0;
	}
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n < 1) {
		pthread_mutex_unlock(&g_lock);
		return rl_fail(REND_ENODEV, "no HIP device (%s)",
		  hipGetErrorString(e));
	}
	hipDeviceProp_t p;
	if ((e = hipSetDevice(0)) != hipSuccess ||
	    (e = hipGetDeviceProperties(&p, 0)) != hipSuccess ||
	    (e = hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking))
	      != hipSuccess) {
		pthread_mutex_unlock(&g_lock);
		return rl_hipfail(e, "device init");
	}
	/* this host's RCPSS/RSQRTPS for the reference-exact kernels */
	rl_approx_ready();
	if ((e = (hipError_t)rk_trace_set_host_rcp(G.rcp_host)) != hipSuccess) {
		pthread_mutex_unlock(&g_lock);
		return rl_hipfail(e, "host RCPSS table");
	}
	G.cus = p.multiProcessorCount;
	snprintf(G.name, sizeof(G.name), "%s", p.gcnArchName);
	G.refill = 48;
	G.max_iter = 1u << 22;
	G.grid_mult = 0;
	G.variant = 0;
	const size_t ctl_words = RL_CTL_SLOTS * RK_CTL_WORDS + RK_DIAG_WORDS;
	if ((e = hipMalloc((void **)&G.ctl, ctl_words * sizeof(unsigned))) != hipSuccess ||
	    (e = hipMemset(G.ctl, 0, ctl_words * sizeof(unsigned))) != hipSuccess ||
	    (e = hipMalloc((void **)&G.stats, 8 * sizeof(unsigned long long)))
	      != hipSuccess) {
		pthread_mutex_unlock(&g_lock);
		return rl_hipfail(e, "control words");
	}
	G.diag = G.ctl + RL_CTL_SLOTS * RK_CTL_WORDS;
	if ((e = hipEventCreateWithFlags(&G.order_ev, hipEventDisableTiming)) != hipSuccess) {
		pthread_mutex_unlock(&g_lock);
		return rl_hipfail(e, "launch order event");
	}
	/* device address of the allocation-free entry point's static spill
	 * area (device 0 is current); a failure only disables that entry */
	G.scratch_err = rk_scratch_init();
	G.ready = 1;
	pthread_mutex_unlock(&g_lock);
	return 0;
}

int rend_hip_devinfo(int *cus, char *name, unsigned int namelen)
{
	int r = rl_device();
	if (r)
		return r;
	if (cus)
		*cus = G.cus;
	if (name && namelen)
		snprintf(name, namelen, "%s", G.name);
	return 0;
}

int rend_hip_set_rcp(int mode)
{
	if (mode != REND_RCP_HOST_TABLE && mode != REND_RCP_IEEE && mode != REND_RCP_XEON_TABLE)
		return rl_fail(REND_EINVAL, "rcp mode %d", mode);
	G.rcp_ieee = mode == REND_RCP_IEEE;
	G.rcp_mode = mode;
	return 0;
}

int rend_hip_set_tlas_build(int mode)
{
	if (mode != REND_TLAS_HOST && mode != REND_TLAS_GPU)
		return rl_fail(REND_EINVAL, "TLAS build mode %d", mode);
	G.tlas_gpu = mode == REND_TLAS_GPU;
	return 0;
}

int rend_hip_set_blas_build(int mode)
{
	if (mode != REND_BLAS_HOST && mode != REND_BLAS_GPU)
		return rl_fail(REND_EINVAL, "BLAS build mode %d", mode);
	G.blas_gpu = mode == REND_BLAS_GPU;
	return 0;
}

int rend_hip_set_tuning(unsigned int grid_mult, unsigned int refill)
{
	if (refill < 1 || refill > 64)
		return rl_fail(REND_EINVAL, "refill %u", refill);
	G.grid_mult = grid_mult;
	G.refill = refill;
	return 0;
}

int rend_hip_set_trace_variant(int variant)
{
	if (variant < 0 || variant >= rk_trace_variants())
		return rl_fail(REND_EINVAL, "trace variant %d", variant);
	G.variant = variant;
	return 0;
}

int rend_hip_trace_lanes(unsigned long long out[4])
{
	memcpy(out, G.lanes, sizeof(G.lanes));
	return 0;
}

int rend_hip_set_step_batch(unsigned int lanes)
{
	if (lanes > 64)
		return rl_fail(REND_EINVAL, "step batch %u (0..64)", lanes);
	G.step_batch = lanes;
	return 0;
}

/* Launch flags of every traversal launch: bits 8..15 carry the step batch. */
unsigned rl_trace_flags(unsigned extra)
{
	return extra | (G.step_batch << 8) |
	  (G.rcp_mode == REND_RCP_XEON_TABLE ? RK_TRACE_XEON : 0u);
}

int rend_hip_get_tuning(unsigned int out[3])
{
	out[0] = G.grid_mult;
	out[1] = G.refill;
	out[2] = (unsigned)G.variant;
	return 0;
}

/* Persistent grid for a trace variant and the matching spill area. */
int rl_trace_grid(int closest, int count, int rcp_ieee, unsigned *grid)
{
	int occ = rk_trace_occupancy(closest, count, rcp_ieee, G.variant);
	if (occ < 1)
		return rl_fail(REND_ENODEV, "trace kernel occupancy query failed");
	unsigned mult = G.grid_mult ? G.grid_mult : (unsigned)occ;
	unsigned g = (unsigned)G.cus * mult;
	size_t need = (size_t)g * REND_TRACE_BLOCK *
	  (REND_STACK_MAX - REND_LDS_STACK_MIN) * 8u;
	if (need > G.spill_bytes) {
		if (G.spill)
			hipFree(G.spill);
		G.spill = NULL;
		G.spill_bytes = 0;
		hipError_t e = hipMalloc(&G.spill, need);
		if (e != hipSuccess)
			return rl_hipfail(e, "trace spill area");
		G.spill_bytes = need;
	}
	*grid = g;
	return 0;
}

/* Launch order across streams (see struct rl_global): before work that
 * uses the shared control ring or spill area is queued on st, st waits for
 * the last such work queued on another stream.  A stream under capture is
 * left alone: the capturing caller ordered the stream before it began. */
static int capturing(hipStream_t st)
{
	hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
	return hipStreamIsCapturing(st, &cs) == hipSuccess &&
	  cs != hipStreamCaptureStatusNone;
}

int rl_stream_order(hipStream_t st)
{
	if (!G.order_pending || G.order_stream == st || capturing(st))
		return 0;
	hipError_t e = hipStreamWaitEvent(st, G.order_ev, 0);
	return e == hipSuccess ? 0 : rl_hipfail(e, "launch order");
}

int rl_stream_mark(hipStream_t st)
{
	if (capturing(st))
		return 0;
	hipError_t e = hipEventRecord(G.order_ev, st);
	if (e != hipSuccess)
		return rl_hipfail(e, "launch order mark");
	G.order_stream = st;
	G.order_pending = 1;
	return 0;
}

/* ---- scene registry -------------------------------------------------------- */

#define MAX_SCENES 64
static struct rl_scene *g_scenes[MAX_SCENES];

struct rl_scene *rl_scene(const struct rdata *rd)
{
	for (int i = 0; i < MAX_SCENES; i++)
		if (g_scenes[i] && g_scenes[i]->rd == rd)
			return g_scenes[i];
	return NULL;
}

static struct rl_scene *scene_add(struct rdata *rd)
{
	struct rl_scene *s = rl_scene(rd);
	if (s)
		return s;
	pthread_mutex_lock(&g_lock);
	for (int i = 0; i < MAX_SCENES; i++)
		if (!g_scenes[i]) {
			s = calloc(1, sizeof(*s));
			if (s) {
				s->rd = rd;
				g_scenes[i] = s;
			}
			break;
		}
	pthread_mutex_unlock(&g_lock);
	return s;
}

static void dfree(void **p)
{
	if (*p)
		hipFree(*p);
	*p = NULL;
}

static void scene_del(struct rdata *rd)
{
	for (int i = 0; i < MAX_SCENES; i++)
		if (g_scenes[i] && g_scenes[i]->rd == rd) {
			struct rl_scene *s = g_scenes[i];
			dfree((void **)&s->arena);
			dfree((void **)&s->insts);
			dfree((void **)&s->aabbs_dev);
			dfree((void **)&s->tlas_info);
			if (s->tlas_stage)
				hipHostFree(s->tlas_stage);
			s->tlas_stage = NULL;
			dfree((void **)&s->nrms);
			dfree((void **)&s->mtls);
			dfree((void **)&s->tris);
			dfree((void **)&s->acc);
			dfree((void **)&s->buf);
			rl_frame_release(s);
			rl_pt_release(s);
			free(s->blas_triofs);
			free(s->blas_dev);
			free(s->blas_len);
			free(s);
			g_scenes[i] = NULL;
		}
}

/* ---- reference API: statics ----------------------------------------------- */

static unsigned g_blkszx = 64, g_blkszy = 32;

void rend_staticinit(unsigned int bszx, unsigned int bszy)
{
	g_blkszx = bszx ? bszx : 1;
	g_blkszy = bszy ? bszy : 1;
}

void rend_staticrelease(void)
{
}

unsigned rl_blkszx(void) { return g_blkszx; }
unsigned rl_blkszy(void) { return g_blkszy; }

/* rend.c:171-187: left-to-right, separately rounded. */
void mulpos(float *ox, float *oy, float *oz, float m[16], float x, float y,
            float z)
{
	*ox = m[0] * x + m[1] * y + m[2] * z + m[3];
	*oy = m[4] * x + m[5] * y + m[6] * z + m[7];
	*oz = m[8] * x + m[9] * y + m[10] * z + m[11];
}

void muldir(float *ox, float *oy, float *oz, float m[16], float x, float y,
            float z)
{
	*ox = m[0] * x + m[1] * y + m[2] * z;
	*oy = m[4] * x + m[5] * y + m[6] * z;
	*oz = m[8] * x + m[9] * y + m[10] * z;
}

/* ---- reference API: lifecycle --------------------------------------------- */

static void *zalloc64(size_t bytes)
{
	size_t sz = (bytes + 63) & ~(size_t)63;
	void *p = aligned_alloc(64, sz ? sz : 64);
	return p;
}

void rend_init(struct rdata *rd, unsigned int maxmtls, unsigned int maxtris,
               unsigned int maxinsts, unsigned int spp)
{
	rd->mtls = zalloc64((size_t)maxmtls * sizeof(*rd->mtls));
	rd->tris = zalloc64((size_t)maxtris * sizeof(*rd->tris));
	rd->nrms = zalloc64((size_t)maxtris * sizeof(*rd->nrms));
	rd->insts = zalloc64((size_t)maxinsts * sizeof(*rd->insts));
	rd->aabbs = zalloc64((size_t)maxinsts * sizeof(*rd->aabbs));
	size_t nb = 2 * ((size_t)maxtris + maxinsts) * sizeof(*rd->bnodes);
	rd->bnodes = zalloc64(nb);
	if (rd->bnodes)
		memset(rd->bnodes, 0, nb);
	rd->tlasofs = 2 * maxtris;
	rd->spp = spp;
	if (!rd->mtls || !rd->tris || !rd->nrms || !rd->insts || !rd->aabbs ||
	    !rd->bnodes) {
		rl_fail(REND_ENOMEM, "rend_init: host allocation failed");
		return;
	}
	struct rl_scene *s = scene_add(rd);
	if (!s) {
		rl_fail(REND_ENOMEM, "rend_init: scene registry full");
		return;
	}
	s->maxmtls = maxmtls;
	s->maxtris = maxtris;
	s->maxinsts = maxinsts;
}

void rend_release(struct rdata *rd)
{
	scene_del(rd);
	free(rd->acc);
	free(rd->bnodes);
	free(rd->aabbs);
	free(rd->insts);
	free(rd->nrms);
	free(rd->tris);
	free(rd->mtls);
	rd->acc = NULL;
	rd->bnodes = NULL;
	rd->aabbs = NULL;
	rd->insts = NULL;
	rd->nrms = NULL;
	rd->tris = NULL;
	rd->mtls = NULL;
}

/* ---- BVH builds ------------------------------------------------------------ */

static int blas_find(const struct rl_scene *s, unsigned triofs)
{
	for (unsigned i = 0; i < s->nblas; i++)
		if (s->blas_triofs[i] == triofs)
			return (int)i;
	return -1;
}

/* Ensures the device arena holds at least `bytes`, keeping the first `keep`
 * bytes. */
static int arena_reserve(struct rl_scene *s, size_t bytes, size_t keep)
{
	if (bytes <= s->arena_cap)
		return 0;
	/* the traversal kernel keeps a tree's position in the arena as a 32-bit
	 * byte offset */
	if (bytes > RL_ARENA_MAX)
		return rl_fail(REND_EINVAL, "scene needs a %zu-byte device arena; the limit is %llu",
		  bytes, (unsigned long long)RL_ARENA_MAX);
	size_t cap = bytes + bytes / 8 + 4096;
	if (cap > RL_ARENA_MAX)
		cap = RL_ARENA_MAX;
	unsigned char *na = NULL;
	hipError_t e = hipMalloc((void **)&na, cap);
	if (e != hipSuccess)
		return rl_hipfail(e, "device arena");
	if (keep && s->arena) {
		e = hipMemcpy(na, s->arena, keep, hipMemcpyDeviceToDevice);
		if (e != hipSuccess) {
			hipFree(na);
			return rl_hipfail(e, "device arena copy");
		}
	}
	dfree((void **)&s->arena);
	s->arena = na;
	s->arena_cap = cap;
	return 0;
}

unsigned long long rk_blas_ws_bytes(unsigned n);
int rk_blas_build(const void *tris, unsigned n, void *ws, unsigned counts[2], void *stream);
int rk_blas_reloc(const void *ws, unsigned n, const unsigned counts[2], void *out,
                  void *stream);

static int upload_tris(struct rdata *rd, struct rl_scene *s)
{
	hipError_t e;
	size_t tb = (size_t)s->maxtris * sizeof(struct rtri);
	if (tb > s->tris_cap) {
		dfree((void **)&s->tris);
		if ((e = hipMalloc((void **)&s->tris, tb)) != hipSuccess)
			return rl_hipfail(e, "triangles");
		s->tris_cap = tb;
	}
	if (tb && (e = hipMemcpy(s->tris, rd->tris, tb, hipMemcpyHostToDevice)) != hipSuccess)
		return rl_hipfail(e, "triangle upload");
	return 0;
}

/* Every distinct BLAS built on the device from the uploaded triangles
 * (kern/blas.hip), placed one after another from the arena's start. */
static int build_blas_dev(struct rdata *rd, struct rl_scene *s, size_t *total_out)
{
	int r = upload_tris(rd, s);
	if (r)
		return r;
	unsigned maxn = 0;
	for (unsigned b = 0; b < s->nblas; b++)
		for (unsigned j = 0; j < rd->instcnt; j++)
			if (rd->insts[j].triofs == s->blas_triofs[b] && rd->insts[j].tricnt > maxn)
				maxn = rd->insts[j].tricnt;
	void *ws = NULL;
	hipError_t e = hipMalloc(&ws, (size_t)rk_blas_ws_bytes(maxn));
	if (e != hipSuccess)
		return rl_hipfail(e, "BLAS build workspace");
	size_t total = 0;
	for (unsigned b = 0; b < s->nblas && !r; b++) {
		unsigned triofs = s->blas_triofs[b], tricnt = 0;
		for (unsigned j = 0; j < rd->instcnt; j++)
			if (rd->insts[j].triofs == triofs) {
				tricnt = rd->insts[j].tricnt;
				break;
			}
		unsigned cnt[2] = {0, 0};
		const unsigned char *tp = (const unsigned char *)s->tris +
		  (size_t)triofs * sizeof(struct rtri);
		int k = rk_blas_build(tp, tricnt, ws, cnt, G.stream);
		if (k) {
			r = k < 0 ? rl_fail(REND_EINVAL, "device BLAS build overflow (%u tris)", tricnt)
			          : rl_hipfail(k, "device BLAS build");
			break;
		}
		size_t sz = (size_t)cnt[0] * sizeof(struct dnode) + (size_t)cnt[1] * sizeof(struct dleaf);
		if ((size_t)cnt[0] * sizeof(struct dnode) + (size_t)cnt[1] * sizeof(struct dleaf) >=
		    ((size_t)1 << 31)) {
			r = rl_fail(REND_EINVAL, "device BLAS of %u tris exceeds 2 GB", tricnt);
			break;
		}
		if ((r = arena_reserve(s, total + sz, total)))
			break;
		if ((k = rk_blas_reloc(ws, tricnt, cnt, s->arena + total, G.stream))) {
			r = rl_hipfail(k, "device BLAS placement");
			break;
		}
		s->blas_dev[b] = total;
		s->blas_len[b] = sz;
		total += (sz + 255) & ~(size_t)255;
	}
	if (!r && (e = hipStreamSynchronize(G.stream)) != hipSuccess)
		r = rl_hipfail(e, "device BLAS build");
	hipFree(ws);
	*total_out = total;
	return r;
}

static int upload_static(struct rdata *rd, struct rl_scene *s)
{
	int r = rl_device();
	if (r)
		return r;
	/* Distinct BLAS roots referenced by instances. */
	free(s->blas_triofs);
	free(s->blas_dev);
	free(s->blas_len);
	s->blas_triofs = malloc(((size_t)rd->instcnt + 1) * sizeof(unsigned));
	s->blas_dev = malloc(((size_t)rd->instcnt + 1) * sizeof(unsigned long long));
	s->blas_len = malloc(((size_t)rd->instcnt + 1) * sizeof(unsigned long long));
	s->nblas = 0;
	if (!s->blas_triofs || !s->blas_dev || !s->blas_len)
		return rl_fail(REND_ENOMEM, "blas map");

	size_t total = 0;
	if (s->blas_dev_build) {
		for (unsigned j = 0; j < rd->instcnt; j++)
			if (blas_find(s, rd->insts[j].triofs) < 0)
				s->blas_triofs[s->nblas++] = rd->insts[j].triofs;
		if ((r = build_blas_dev(rd, s, &total)))
			return r;
		size_t tlas_res = 2 * (size_t)(s->maxinsts ? s->maxinsts : 1) *
		  sizeof(struct bnode8);
		if ((r = arena_reserve(s, total + tlas_res, total)))
			return r;
		s->blas_bytes = total;
		goto shading;
	}
	long long *sizes = malloc(((size_t)rd->instcnt + 1) * sizeof(long long));
	if (!sizes)
		return rl_fail(REND_ENOMEM, "blas sizes");
	for (unsigned j = 0; j < rd->instcnt; j++) {
		const struct rinst *ri = &rd->insts[j];
		if (blas_find(s, ri->triofs) >= 0)
			continue;
		const struct bnode8 *root = &rd->bnodes[(size_t)ri->triofs << 1];
		size_t limit = 2 * (size_t)ri->tricnt * sizeof(struct bnode8);
		if (limit < sizeof(struct bnode8))
			limit = sizeof(struct bnode8);
		long long sz = bvh8_repack_size(root, 1, limit, NULL, NULL);
		if (sz < 0) {
			free(sizes);
			return rl_fail(REND_EINVAL, "malformed BLAS at triofs %u",
			  ri->triofs);
		}
		s->blas_triofs[s->nblas] = ri->triofs;
		s->blas_dev[s->nblas] = total;
		s->blas_len[s->nblas] = (unsigned long long)sz;
		sizes[s->nblas] = sz;
		s->nblas++;
		total += ((size_t)sz + 255) & ~(size_t)255;
	}
	size_t tlas_res = 2 * (size_t)(s->maxinsts ? s->maxinsts : 1) *
	  sizeof(struct bnode8);
	r = arena_reserve(s, total + tlas_res, 0);
	if (r) {
		free(sizes);
		return r;
	}
	unsigned char *stage = NULL;
	size_t stagecap = 0;
	for (unsigned b = 0; b < s->nblas; b++) {
		unsigned triofs = s->blas_triofs[b];
		const struct bnode8 *root = &rd->bnodes[(size_t)triofs << 1];
		unsigned tricnt = 0;
		for (unsigned j = 0; j < rd->instcnt; j++)
			if (rd->insts[j].triofs == triofs) {
				tricnt = rd->insts[j].tricnt;
				break;
			}
		size_t limit = 2 * (size_t)tricnt * sizeof(struct bnode8);
		if (limit < sizeof(struct bnode8))
			limit = sizeof(struct bnode8);
		if ((size_t)sizes[b] > stagecap) {
			free(stage);
			stagecap = (size_t)sizes[b];
			stage = aligned_alloc(64, (stagecap + 63) & ~(size_t)63);
			if (!stage) {
				free(sizes);
				return rl_fail(REND_ENOMEM, "repack staging");
			}
		}
		if (bvh8_repack(stage, stagecap, root, 1, limit) != sizes[b]) {
			free(stage);
			free(sizes);
			return rl_fail(REND_EINVAL, "BLAS repack failed");
		}
		hipError_t e = hipMemcpy(s->arena + s->blas_dev[b], stage,
		  (size_t)sizes[b], hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			free(stage);
			free(sizes);
			return rl_hipfail(e, "BLAS upload");
		}
	}
	free(stage);
	free(sizes);
	s->blas_bytes = total;

shading:
	/* Normals and materials for shading. */
	size_t nb = (size_t)s->maxtris * sizeof(struct rnrm);
	size_t mb = (size_t)s->maxmtls * sizeof(struct rmtl);
	hipError_t e = hipSuccess;
	if (nb > s->nrms_cap) {
		dfree((void **)&s->nrms);
		if ((e = hipMalloc((void **)&s->nrms, nb)) != hipSuccess)
			return rl_hipfail(e, "normals");
		s->nrms_cap = nb;
	}
	if (mb > s->mtls_cap) {
		dfree((void **)&s->mtls);
		if ((e = hipMalloc((void **)&s->mtls, mb)) != hipSuccess)
			return rl_hipfail(e, "materials");
		s->mtls_cap = mb;
	}
	if (!s->blas_dev_build && (r = upload_tris(rd, s)))
		return r;
	if ((nb && (e = hipMemcpy(s->nrms, rd->nrms, nb, hipMemcpyHostToDevice))
	      != hipSuccess) ||
	    (mb && (e = hipMemcpy(s->mtls, rd->mtls, mb, hipMemcpyHostToDevice))
	      != hipSuccess))
		return rl_hipfail(e, "shading data upload");
	s->static_ok = 1;
	s->dynamic_ok = 0;
	s->gen++;
	return 0;
}

int rk_tlas_build(const void *aabbs, unsigned n, void *nodes, unsigned *info, void *stream);

/* Instances and a TLAS built on the device from the instances' world AABBs
 * (kern/tlas.hip) into the arena after the BLASes. */
static int upload_dynamic_dev(struct rdata *rd, struct rl_scene *s)
{
	unsigned n = rd->instcnt;
	if (n > 4096)
		return rl_fail(REND_EINVAL, "%u instances (device TLAS build: at most 4096)", n);
	size_t sz = (size_t)(n ? n : 1) * sizeof(struct bnode8);
	size_t at = (s->blas_bytes + 255) & ~(size_t)255;
	int r = arena_reserve(s, at + sz, s->blas_bytes);
	if (r)
		return r;
	/* pinned staging (instances, boxes, result words), kept per scene, so
	 * the uploads are plain DMA */
	size_t ib = ((size_t)n + 1) * sizeof(struct dinst);
	size_t ab = ((size_t)n + 1) * sizeof(struct aabb);
	size_t need = ib + ab + 4 * sizeof(unsigned);
	hipError_t e = hipSuccess;
	if (need > s->tlas_stage_cap) {
		if (s->tlas_stage)
			hipHostFree(s->tlas_stage);
		s->tlas_stage = NULL;
		s->tlas_stage_cap = 0;
		if ((e = hipHostMalloc(&s->tlas_stage, need, hipHostMallocDefault)) != hipSuccess)
			return rl_hipfail(e, "TLAS staging");
		s->tlas_stage_cap = need;
	}
	struct dinst *di = (struct dinst *)s->tlas_stage;
	struct aabb *sab = (struct aabb *)((unsigned char *)s->tlas_stage + ib);
	unsigned *info = (unsigned *)((unsigned char *)s->tlas_stage + ib + ab);
	for (unsigned i = 0; i < n; i++) {
		const struct rinst *ri = &rd->insts[i];
		int b = blas_find(s, ri->triofs);
		if (b < 0)
			return rl_fail(REND_ESTATE,
			  "instance %u references a BLAS not built by rend_prepstatic", i);
		memcpy(di[i].inv, ri->globinv, sizeof(di[i].inv));
		di[i].blas = s->blas_dev[b];
		di[i].triofs = ri->triofs;
		di[i].flags = ri->flags;
	}
	if (n)
		memcpy(sab, rd->aabbs, (size_t)n * sizeof(struct aabb));
	info[0] = info[1] = 0;
	if (ib > s->insts_cap) {
		dfree((void **)&s->insts);
		s->insts_cap = 0;
		if ((e = hipMalloc((void **)&s->insts, ib)) == hipSuccess)
			s->insts_cap = ib;
	}
	if (e == hipSuccess && ab > s->aabbs_cap) {
		dfree(&s->aabbs_dev);
		s->aabbs_cap = 0;
		if ((e = hipMalloc(&s->aabbs_dev, ab)) == hipSuccess)
			s->aabbs_cap = ab;
	}
	if (e == hipSuccess && !s->tlas_info)
		e = hipMalloc((void **)&s->tlas_info, 4 * sizeof(unsigned));
	if (e == hipSuccess && n)
		e = hipMemcpyAsync(s->insts, di, (size_t)n * sizeof(*di), hipMemcpyHostToDevice,
		  G.stream);
	if (e == hipSuccess && n)
		e = hipMemcpyAsync(s->aabbs_dev, sab, (size_t)n * sizeof(struct aabb),
		  hipMemcpyHostToDevice, G.stream);
	if (e != hipSuccess)
		return rl_hipfail(e, "device TLAS inputs");
	r = rk_tlas_build(s->aabbs_dev, n, s->arena + at, s->tlas_info, G.stream);
	if (!r)
		r = (int)hipMemcpyAsync(info, s->tlas_info, 2 * sizeof(unsigned),
		  hipMemcpyDeviceToHost, G.stream);
	if (!r)
		r = (int)hipStreamSynchronize(G.stream);
	if (r)
		return rl_hipfail(r, "device TLAS build");
	if (info[1] || !info[0] || (size_t)info[0] * sizeof(struct bnode8) > sz)
		return rl_fail(REND_EINVAL, "device TLAS build failed (nodes %u, code %u)",
		  info[0], info[1]);
	s->tlas = at;
	s->tlas_bytes = (size_t)info[0] * sizeof(struct bnode8);
	s->arena_used = at + s->tlas_bytes;
	s->ninsts = n;
	s->dynamic_ok = 1;
	s->gen++;
	return 0;
}

static int upload_dynamic(struct rdata *rd, struct rl_scene *s)
{
	if (!s->static_ok)
		return rl_fail(REND_ESTATE, "rend_prepdynamic before rend_prepstatic");
	if (s->tlas_dev)
		return upload_dynamic_dev(rd, s);
	const struct bnode8 *root = &rd->bnodes[rd->tlasofs];
	size_t limit = 2 * (size_t)(s->maxinsts ? s->maxinsts : 1) *
	  sizeof(struct bnode8);
	long long sz = bvh8_repack_size(root, 0, limit, NULL, NULL);
	if (sz < 0)
		return rl_fail(REND_EINVAL, "malformed TLAS");
	size_t at = (s->blas_bytes + 255) & ~(size_t)255;
	int r = arena_reserve(s, at + (size_t)sz, s->blas_bytes);
	if (r)
		return r;
	unsigned char *stage = aligned_alloc(64, ((size_t)sz + 63) & ~(size_t)63);
	struct dinst *di = malloc(((size_t)rd->instcnt + 1) * sizeof(*di));
	if (!stage || !di) {
		free(stage);
		free(di);
		return rl_fail(REND_ENOMEM, "TLAS staging");
	}
	if (bvh8_repack(stage, (size_t)sz, root, 0, limit) != sz) {
		free(stage);
		free(di);
		return rl_fail(REND_EINVAL, "TLAS repack failed");
	}
	for (unsigned i = 0; i < rd->instcnt; i++) {
		const struct rinst *ri = &rd->insts[i];
		int b = blas_find(s, ri->triofs);
		if (b < 0) {
			free(stage);
			free(di);
			return rl_fail(REND_ESTATE,
			  "instance %u references a BLAS not built by rend_prepstatic", i);
		}
		memcpy(di[i].inv, ri->globinv, sizeof(di[i].inv));
		di[i].blas = s->blas_dev[b];
		di[i].triofs = ri->triofs;
		di[i].flags = ri->flags;
	}
	hipError_t e;
	size_t ib = ((size_t)rd->instcnt + 1) * sizeof(*di);
	if (ib > s->insts_cap) {
		dfree((void **)&s->insts);
		if ((e = hipMalloc((void **)&s->insts, ib)) != hipSuccess) {
			free(stage);
			free(di);
			return rl_hipfail(e, "instances");
		}
		s->insts_cap = ib;
	}
	e = hipMemcpy(s->arena + at, stage, (size_t)sz, hipMemcpyHostToDevice);
	if (e == hipSuccess && rd->instcnt)
		e = hipMemcpy(s->insts, di, (size_t)rd->instcnt * sizeof(*di),
		  hipMemcpyHostToDevice);
	free(stage);
	free(di);
	if (e != hipSuccess)
		return rl_hipfail(e, "TLAS upload");
	s->tlas = at;
	s->arena_used = at + (size_t)sz;
	s->tlas_bytes = (size_t)sz;
	s->ninsts = rd->instcnt;
	s->dynamic_ok = 1;
	s->gen++;
	return 0;
}

/* GPU BLAS mode: the host arena keeps, per mesh, a root whose first child
 * box is the mesh's bounds and whose first child word is non-zero (the
 * reference's "already built" test, rend.c:3039); upd_rinsts reads the root's
 * bounds (main.c:117-157).  The tree itself is built on the device. */
static void stub_root(struct bnode8 *rn, struct vec3 lo, struct vec3 hi)
{
	memset(rn, 0, sizeof(*rn));
	float *b[6] = {(float *)&rn->minx8, (float *)&rn->maxx8, (float *)&rn->miny8,
	  (float *)&rn->maxy8, (float *)&rn->minz8, (float *)&rn->maxz8};
	const float v[6] = {lo.x, hi.x, lo.y, hi.y, lo.z, hi.z};
	for (int a = 0; a < 6; a++)
		for (int c = 0; c < 8; c++)
			b[a][c] = c ? ((a & 1) ? -FLT_MAX : FLT_MAX) : v[a];
	const unsigned w = DNODE_LEAF;
	memcpy(&rn->children8, &w, sizeof(w));
}

void rend_prepstatic(struct rdata *rd)
{
	const int dev_build = G.blas_gpu;
	for (unsigned j = 0; j < rd->instcnt; j++) {
		struct rinst *ri = &rd->insts[j];
		unsigned triofs = ri->triofs;
		struct bnode8 *rn = &rd->bnodes[(size_t)triofs << 1];
		if (((unsigned *)&rn->children8)[0])
			continue;  /* already built (rend.c:3039) */
		unsigned tricnt = ri->tricnt;
		const struct rtri *tp = &rd->tris[triofs];
		if (dev_build) {
			struct vec3 lo = {FLT_MAX, FLT_MAX, FLT_MAX};
			struct vec3 hi = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
			for (unsigned i = 0; i < tricnt; i++) {
				lo = rv_min(rv_min(rv_min(lo, tp[i].v0), tp[i].v1), tp[i].v2);
				hi = rv_max(rv_max(rv_max(hi, tp[i].v0), tp[i].v1), tp[i].v2);
			}
			stub_root(rn, lo, hi);
			continue;
		}
		unsigned *imap = malloc(((size_t)tricnt + 1) * sizeof(*imap));
		struct aabb *ab = malloc(((size_t)tricnt + 1) * sizeof(*ab));
		if (!imap || !ab) {
			free(imap);
			free(ab);
			rl_fail(REND_ENOMEM, "BLAS build buffers");
			return;
		}
		struct vec3 rmin = {FLT_MAX, FLT_MAX, FLT_MAX};
		struct vec3 rmax = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
		for (unsigned i = 0; i < tricnt; i++) {
			ab[i].min = ab[i].max = tp[i].v0;
			ab[i].min = rv_min(ab[i].min, tp[i].v1);
			ab[i].max = rv_max(ab[i].max, tp[i].v1);
			ab[i].min = rv_min(ab[i].min, tp[i].v2);
			ab[i].max = rv_max(ab[i].max, tp[i].v2);
			rmin = rv_min(rmin, ab[i].min);
			rmax = rv_max(rmax, ab[i].max);
			imap[i] = i;
		}
		struct bvh8_stats st;
		bvh8_build(rn, ab, imap, tp, tricnt, rmin, rmax, &st);
		struct rl_scene *s = rl_scene(rd);
		if (s)
			s->hazards |= st.hazards;
		free(ab);
		free(imap);
	}
	struct rl_scene *s = scene_add(rd);
	if (!s) {
		rl_fail(REND_ENOMEM, "scene registry full");
		return;
	}
	s->blas_dev_build = dev_build;
	/* Device upload happens at the first trace or frame after this. */
	s->static_dirty = 1;
	if (s->tlas_built)
		s->dynamic_dirty = 1;
}

void rend_prepdynamic(struct rdata *rd)
{
	if (G.tlas_gpu) {
		/* built on the device at the next upload; the host arena's TLAS
		 * region is left as it was */
		struct rl_scene *s = scene_add(rd);
		if (!s) {
			rl_fail(REND_ENOMEM, "scene registry full");
			return;
		}
		s->tlas_dev = 1;
		s->tlas_built = 1;
		s->dynamic_dirty = 1;
		return;
	}
	unsigned n = rd->instcnt;
	unsigned *imap = malloc(((size_t)n + 1) * sizeof(*imap));
	if (!imap) {
		rl_fail(REND_ENOMEM, "TLAS build buffers");
		return;
	}
	struct vec3 rmin = {FLT_MAX, FLT_MAX, FLT_MAX};
	struct vec3 rmax = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (unsigned i = 0; i < n; i++) {
		rmin = rv_min(rmin, rd->aabbs[i].min);
		rmax = rv_max(rmax, rd->aabbs[i].max);
		imap[i] = i;
	}
	struct bvh8_stats st;
	bvh8_build(&rd->bnodes[rd->tlasofs], rd->aabbs, imap, NULL, n, rmin,
	  rmax, &st);
	free(imap);
	struct rl_scene *s = scene_add(rd);
	if (!s) {
		rl_fail(REND_ENOMEM, "scene registry full");
		return;
	}
	s->hazards |= st.hazards;
	s->tlas_dev = 0;
	s->tlas_built = 1;
	s->dynamic_dirty = 1;
}

/* Uploads whatever rend_prepstatic/rend_prepdynamic rebuilt since the last
 * upload. */
int rl_sync_scene(struct rl_scene *s)
{
	int r = 0;
	if (s->static_dirty) {
		if ((r = upload_static(s->rd, s)))
			return r;
		s->static_dirty = 0;
		if (s->tlas_built)
			s->dynamic_dirty = 1;
	}
	if (s->dynamic_dirty) {
		if ((r = upload_dynamic(s->rd, s)))
			return r;
		s->dynamic_dirty = 0;
	}
	return 0;
}

int rend_hip_upload(struct rdata *rd)
{
	struct rl_scene *s = scene_add(rd);
	if (!s)
		return rl_fail(REND_ENOMEM, "scene registry full");
	s->static_dirty = 1;
	s->dynamic_dirty = s->tlas_built;
	return rl_sync_scene(s);
}

unsigned int rend_hip_hazards(struct rdata *rd)
{
	struct rl_scene *s = rl_scene(rd);
	return s ? s->hazards : 0;
}

long long rend_hip_arena_bytes(struct rdata *rd)
{
	struct rl_scene *s = rl_scene(rd);
	return s && s->dynamic_ok ? (long long)s->arena_used : -1;
}

/* ---- accumulation ------------------------------------------------------- */

void rend_resaccum(struct rdata *rd, unsigned int w, unsigned int h)
{
	free(rd->acc);
	rd->acc = zalloc64((size_t)w * h * sizeof(*rd->acc));
	rd->width = w;
	rd->height = h;
	rend_clraccum(rd);
}

void rend_clraccum(struct rdata *rd)
{
	if (rd->acc)
		memset(rd->acc, 0, (size_t)rd->width * rd->height * sizeof(*rd->acc));
	rd->samples = 0;
	struct rl_scene *s = rl_scene(rd);
	if (s)
		s->acc_dirty = 1;
}

/* ---- ray-buffer entry points ---------------------------------------------- */

int rl_trace_async(struct rl_scene *s, const void *d_rays, void *d_hits,
                   void *d_occ, unsigned n, int closest, int count,
                   hipStream_t st)
{
	unsigned grid;
	int r = rl_trace_grid(closest, count, G.rcp_ieee, &grid);
	if (r)
		return r;
	if ((r = rl_stream_order(st)))
		return r;
	unsigned slot = G.ctl_next++ % RL_CTL_SLOTS;
	int e = rk_trace_launch(s->arena, s->tlas, s->insts, d_rays, d_hits,
	  d_occ, n, G.ctl + RK_CTL_WORDS * slot, G.spill, count ? G.stats : NULL,
	  closest, count, G.rcp_ieee, G.variant, grid, G.refill, G.max_iter, NULL, G.diag,
	  rl_trace_flags(0), st);
	if (e)
		return rl_hipfail(e, "trace launch");
	G.last_slot = slot;
	return rl_stream_mark(st);
}

static int check_ready(struct rdata *rd, struct rl_scene **out)
{
	int r = rl_device();
	if (r)
		return r;
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->tlas_built)
		return rl_fail(REND_ESTATE, "scene not prepared (call rend_prepstatic "
		  "and rend_prepdynamic first)");
	if ((r = rl_sync_scene(s)))
		return r;
	*out = s;
	return 0;
}

int rend_hip_tlas_download(struct rdata *rd, void *dst, unsigned long long cap,
                           unsigned int *nodes)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	if (nodes)
		*nodes = (unsigned)(s->tlas_bytes / sizeof(struct bnode8));
	if (!dst)
		return 0;
	if (cap < s->tlas_bytes)
		return rl_fail(REND_EINVAL, "TLAS is %zu bytes, buffer %llu", s->tlas_bytes, cap);
	hipError_t e = hipMemcpyAsync(dst, s->arena + s->tlas, s->tlas_bytes,
	  hipMemcpyDeviceToHost, G.stream);
	if (!e)
		e = hipStreamSynchronize(G.stream);
	return e ? rl_hipfail(e, "TLAS download") : 0;
}

int rend_hip_blas_download(struct rdata *rd, unsigned int triofs, void *dst,
                           unsigned long long cap, unsigned long long *bytes)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	int b = blas_find(s, triofs);
	if (b < 0)
		return rl_fail(REND_EINVAL, "no BLAS at triangle offset %u", triofs);
	if (bytes)
		*bytes = s->blas_len[b];
	if (!dst)
		return 0;
	if (cap < s->blas_len[b])
		return rl_fail(REND_EINVAL, "BLAS is %llu bytes, buffer %llu", s->blas_len[b], cap);
	hipError_t e = hipMemcpyAsync(dst, s->arena + s->blas_dev[b], s->blas_len[b],
	  hipMemcpyDeviceToHost, G.stream);
	if (!e)
		e = hipStreamSynchronize(G.stream);
	return e ? rl_hipfail(e, "BLAS download") : 0;
}

int rend_hip_sync_scene(struct rdata *rd)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	hipError_t e = hipStreamSynchronize(G.stream);
	return e ? rl_hipfail(e, "scene upload") : 0;
}

/* Reads and clears the error word of the last launch (synchronises). */
int rl_check_errors(hipStream_t st)
{
	hipError_t e = hipStreamSynchronize(st);
	if (e != hipSuccess)
		return rl_hipfail(e, "trace");
	unsigned w[RK_DIAG_WORDS];
	e = hipMemcpy(w, G.diag, sizeof(w), hipMemcpyDeviceToHost);
	if (e != hipSuccess)
		return rl_hipfail(e, "trace status");
	if (!w[0])
		return 0;
	/* report and clear: the first failing ray of all launches since the
	 * last check */
	hipMemset(G.diag, 0, sizeof(w));
	float r[8];
	memcpy(r, &w[4], sizeof(r));
	return rl_fail(REND_EOVERFLOW, "traversal %s (first: ray %u o=(%.9g %.9g %.9g) "
	  "d=(%.9g %.9g %.9g) tmax=%.9g, %u steps)",
	  (w[14] & RK_ERR_STACK) ? "stack overflow (64 entries)" : "iteration limit",
	  w[12], (double)r[0], (double)r[1], (double)r[2], (double)r[4], (double)r[5],
	  (double)r[6], (double)r[3], w[13]);
}

int rend_hip_set_iter_limit(unsigned int steps)
{
	if (steps < 64)
		return rl_fail(REND_EINVAL, "iteration limit %u", steps);
	G.max_iter = steps;
	return 0;
}

int rend_hip_trace(struct rdata *rd, const struct rray *d_rays,
                   struct rhit *d_hits, unsigned int n, void *stream,
                   struct rtrace_stats *stats)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	if (stats) {
		if ((r = rl_stream_order(st)))
			return r;
		hipError_t e = hipMemsetAsync(G.stats, 0,
		  8 * sizeof(unsigned long long), st);
		if (e != hipSuccess)
			return rl_hipfail(e, "stats clear");
	}
	r = rl_trace_async(s, d_rays, d_hits, NULL, n, 1, stats != NULL, st);
	if (r || !stats)
		return r;
	unsigned long long c[8];
	hipError_t e = hipMemcpyAsync(c, G.stats, sizeof(c), hipMemcpyDeviceToHost,
	  st);
	if (e == hipSuccess)
		e = hipStreamSynchronize(st);
	if (e != hipSuccess)
		return rl_hipfail(e, "stats read");
	stats->nodes += c[0];
	stats->leaves += c[1];
	stats->insts += c[2];
	stats->rays += c[3];
	memcpy(G.lanes, &c[4], sizeof(G.lanes));
	return rl_check_errors(st);
}

/* Closest hit without any allocation: the scene must be on the device
 * already (rend_hip_sync_scene), the control ring and error block are the
 * ones rl_device allocated at set-up, and the traversal stack spills into
 * the static array of kern/trace.hip.  Asynchronous; errors are reported by
 * the next synchronous call, as for rend_hip_trace without stats. */
int rend_hip_trace_noalloc(struct rdata *rd, const struct rray *d_rays,
                           struct rhit *d_hits, unsigned int n, void *stream)
{
	if (!G.ready)
		return rl_fail(REND_ESTATE, "trace_noalloc: device not set up (upload the "
		  "scene with rend_hip_sync_scene first)");
	if (G.scratch_err)
		return rl_hipfail(G.scratch_err, "trace_noalloc: static spill area");
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->tlas_built || !s->dynamic_ok || s->static_dirty || s->dynamic_dirty)
		return rl_fail(REND_ESTATE, "trace_noalloc: scene not on the device (call "
		  "rend_hip_sync_scene after rend_prepstatic/rend_prepdynamic)");
	if (n > RK_NOALLOC_MAX_RAYS)
		return rl_fail(REND_EINVAL, "trace_noalloc: %u rays, at most %u per call", n,
		  RK_NOALLOC_MAX_RAYS);
	if (n && (!d_rays || !d_hits))
		return rl_fail(REND_EINVAL, "trace_noalloc: no ray or hit buffer");
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	int r = rl_stream_order(st);
	if (r)
		return r;
	unsigned slot = G.ctl_next++ % RL_CTL_SLOTS;
	int e = rk_trace_launch_noalloc(s->arena, s->tlas, s->insts, d_rays, d_hits, n,
	  G.ctl + RK_CTL_WORDS * slot, G.rcp_ieee, (unsigned)G.cus, G.refill, G.max_iter,
	  G.diag, rl_trace_flags(0), st);
	if (e)
		return rl_hipfail(e, "trace_noalloc launch");
	G.last_slot = slot;
	return rl_stream_mark(st);
}

int rend_hip_occluded(struct rdata *rd, const struct rray *d_rays,
                      unsigned char *d_occ, unsigned int n, void *stream)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	return rl_trace_async(s, d_rays, NULL, d_occ, n, 0, 0, st);
}

static int host_roundtrip(struct rdata *rd, const struct rray *rays,
                          void *out, size_t outsz, unsigned n, int closest,
                          struct rtrace_stats *stats)
{
	struct rl_scene *s;
	int r = check_ready(rd, &s);
	if (r)
		return r;
	if (n == 0)
		return 0;
	void *dr = NULL, *dout = NULL;
	hipError_t e = hipMalloc(&dr, (size_t)n * sizeof(*rays));
	if (e == hipSuccess)
		e = hipMalloc(&dout, (size_t)n * outsz);
	if (e != hipSuccess) {
		dfree(&dr);
		return rl_hipfail(e, "ray buffers");
	}
	e = hipMemcpy(dr, rays, (size_t)n * sizeof(*rays), hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		r = rl_hipfail(e, "ray upload");
		goto out;
	}
	if (closest)
		r = rend_hip_trace(rd, dr, dout, n, G.stream, stats);
	else
		r = rend_hip_occluded(rd, dr, dout, n, G.stream);
	if (!r)
		r = rl_check_errors(G.stream);
	if (!r) {
		e = hipMemcpy(out, dout, (size_t)n * outsz, hipMemcpyDeviceToHost);
		if (e != hipSuccess)
			r = rl_hipfail(e, "result download");
	}
out:
	dfree(&dr);
	dfree(&dout);
	return r;
}

int rend_hip_trace_host(struct rdata *rd, const struct rray *rays,
                        struct rhit *hits, unsigned int n,
                        struct rtrace_stats *stats)
{
	return host_roundtrip(rd, rays, hits, sizeof(*hits), n, 1, stats);
}

int rend_hip_occluded_host(struct rdata *rd, const struct rray *rays,
                           unsigned char *occ, unsigned int n)
{
	return host_roundtrip(rd, rays, occ, 1, n, 0, NULL);
}

/* ---- device memory and timing helpers --------------------------------------- */

void *rend_hip_malloc(unsigned long long bytes)
{
	if (rl_device())
		return NULL;
	void *p = NULL;
	hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
	if (e != hipSuccess) {
		rl_hipfail(e, "rend_hip_malloc");
		return NULL;
	}
	return p;
}

int rend_hip_free(void *d)
{
	hipError_t e = hipFree(d);
	return e == hipSuccess ? 0 : rl_hipfail(e, "rend_hip_free");
}

int rend_hip_memcpy_htod(void *d, const void *h, unsigned long long bytes)
{
	hipError_t e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
	return e == hipSuccess ? 0 : rl_hipfail(e, "htod");
}

int rend_hip_memcpy_dtoh(void *h, const void *d, unsigned long long bytes)
{
	hipError_t e = hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost);
	return e == hipSuccess ? 0 : rl_hipfail(e, "dtoh");
}

int rend_hip_sync(void)
{
	if (!G.ready)
		return 0;
	hipError_t e = hipStreamSynchronize(G.stream);
	if (e == hipSuccess)
		e = hipDeviceSynchronize();
	return e == hipSuccess ? 0 : rl_hipfail(e, "sync");
}

void *rend_hip_stream(void)
{
	return rl_device() ? NULL : (void *)G.stream;
}

void *rend_hip_event_create(void)
{
	if (rl_device())
		return NULL;
	hipEvent_t ev;
	return hipEventCreate(&ev) == hipSuccess ? (void *)ev : NULL;
}

int rend_hip_event_record(void *ev, void *stream)
{
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	hipError_t e = hipEventRecord((hipEvent_t)ev, st);
	return e == hipSuccess ? 0 : rl_hipfail(e, "event record");
}

float rend_hip_event_elapsed(void *start, void *stop)
{
	float ms = -1.0f;
	if (hipEventSynchronize((hipEvent_t)stop) != hipSuccess)
		return -1.0f;
	if (hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop)
	    != hipSuccess)
		return -1.0f;
	return ms;
}

int rend_hip_event_destroy(void *ev)
{
	return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : REND_ENODEV;
}
