/*
 * rend_render on the GPU (rend.c:3603-3708 protocol, main.c:386-427).
 *
 * The reference runs rend_render on every host core; each thread claims
 * 64x32 blocks with an atomic counter until blkcnt = floor(w/bx)*floor(h/by)
 * blocks are taken, so rows and columns past the last full block are never
 * rendered.  Here the first caller of a frame claims all blocks at once
 * (fetch_add(&blknum, blkcnt) returns 0 only for it), renders the block grid
 * on the GPU and copies the ARGB buffer to rd->buf; every other caller sees
 * blknum >= blkcnt and returns 0.  rays is incremented by the number of rays
 * traced, as the reference does.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_hip.h"
#include "librend.h"
#include "../dlayout.h"
#include "../rsqrt_table.h"

int rk_camrays_ref_setup(const unsigned *tab, void *stream);
int rk_camrays_ref(const struct CamArgs *c, unsigned bszx, unsigned bszy, unsigned blkcntx,
                   unsigned nblk, unsigned samples, const void *lens, void *out, void *stream);
int rk_camrays(const struct CamArgs *c, unsigned rw, unsigned rh,
               unsigned spp, unsigned seed, unsigned frame, void *out,
               void *stream);
int rk_shade_normals(const void *hits, const void *insts, const void *nrms,
                     const void *mtls, const float bg[3], unsigned rw,
                     unsigned rh, unsigned width, unsigned spp,
                     float invsamples, float *acc, unsigned *buf,
                     void *stream);
int rk_shade_normals_ref(const void *hits, const void *insts, const void *nrms,
                         const void *mtls, const float bg[3], unsigned bszx, unsigned bszy,
                         unsigned blkcntx, unsigned nblk, unsigned width, unsigned height,
                         float invsamples, float *acc, unsigned *buf, void *stream);
int rk_fill_f32(float *p, unsigned long long n, float v, void *stream);
int rk_diffuse_rays(const void *prim, const void *hits, const void *insts,
                    const void *tris, unsigned n, unsigned seed, void *out,
                    void *stream);

struct rl_frame {
	void    *rays, *hits;
	size_t   cap;        /* rays */
	unsigned frame;      /* frames rendered since the last clear */
};

void rl_frame_release(struct rl_scene *s)
{
	if (!s->frame)
		return;
	if (s->frame->rays)
		hipFree(s->frame->rays);
	if (s->frame->hits)
		hipFree(s->frame->hits);
	free(s->frame);
	s->frame = NULL;
}

void rl_camargs(const struct rcam *c, struct CamArgs *a)
{
	const float *f;
#define L0(v) (f = (const float *)&(v), f[0])
	a->eye[0] = L0(c->eyex8);
	a->eye[1] = L0(c->eyey8);
	a->eye[2] = L0(c->eyez8);
	a->ri[0] = L0(c->rix8);
	a->ri[1] = L0(c->riy8);
	a->ri[2] = L0(c->riz8);
	a->up[0] = L0(c->upx8);
	a->up[1] = L0(c->upy8);
	a->up[2] = L0(c->upz8);
	a->fwd[0] = L0(c->fwdx8);
	a->fwd[1] = L0(c->fwdy8);
	a->fwd[2] = L0(c->fwdz8);
	a->rw = L0(c->rw8);
	a->rh = L0(c->rh8);
	a->aspect = L0(c->aspect8);
	a->fdist = L0(c->fdist8);
	a->fovfdist = L0(c->fovfdist8);
	a->focrad = L0(c->focrad8);
#undef L0
	a->dof = c->tanfangle > 0.0f;
}

/* Device accumulator and colour buffer sized for the current resolution. */
int rl_acc_ready(struct rl_scene *s, hipStream_t st)
{
	struct rdata *rd = s->rd;
	size_t npx = (size_t)rd->width * rd->height;
	hipError_t e;
	if (npx > s->acc_cap) {
		if (s->acc)
			hipFree(s->acc);
		if (s->buf)
			hipFree(s->buf);
		s->acc = NULL;
		s->buf = NULL;
		s->acc_cap = 0;
		if ((e = hipMalloc((void **)&s->acc, npx * 3 * sizeof(float)))
		      != hipSuccess ||
		    (e = hipMalloc((void **)&s->buf, npx * sizeof(unsigned)))
		      != hipSuccess)
			return rl_hipfail(e, "accumulator");
		s->acc_cap = npx;
		s->acc_dirty = 1;
	}
	if (s->acc_dirty) {
		int r = rk_fill_f32(s->acc, npx * 3, 0.0f, st);
		if (r)
			return rl_hipfail(r, "accumulator clear");
		s->acc_dirty = 0;
		if (s->frame)
			s->frame->frame = 0;
	}
	return 0;
}

static int frame_buffers(struct rl_scene *s, size_t nrays)
{
	if (!s->frame) {
		s->frame = calloc(1, sizeof(*s->frame));
		if (!s->frame)
			return rl_fail(REND_ENOMEM, "frame state");
	}
	struct rl_frame *f = s->frame;
	if (nrays <= f->cap)
		return 0;
	if (f->rays)
		hipFree(f->rays);
	if (f->hits)
		hipFree(f->hits);
	f->rays = f->hits = NULL;
	f->cap = 0;
	hipError_t e;
	if ((e = hipMalloc(&f->rays, nrays * sizeof(struct rray))) != hipSuccess ||
	    (e = hipMalloc(&f->hits, nrays * sizeof(struct rhit))) != hipSuccess)
		return rl_hipfail(e, "frame ray buffers");
	f->cap = nrays;
	return 0;
}

static int rsqrt_table_ready(hipStream_t st);
static int lens_table_ready(const struct rcam *c, const void **lens);

static int is_pow2(unsigned v) { return v && !(v & (v - 1)); }

/* Normal-tint frame: the reference's active rend_render (rend.c:3603-3708).
 * Like the reference it traces one primary ray per pixel of the block grid
 * per frame and divides the accumulator by spp + samples.  With a pinhole
 * camera and blocks made of 4x2 packets in a Morton square (or 2:1) the rays
 * are the reference's own (make_camray_pckt with its xorshift128+ streams,
 * RSQRTPS and, for a thin lens, libmvec's cos/sin), so the frame equals the
 * reference's single-ray shading of its primaries bit for bit; with other
 * block shapes the rays come from k_camrays' jittered, exactly normalised
 * camera. */
static int render_normals(struct rl_scene *s, unsigned rw, unsigned rh,
                          unsigned *rays_out)
{
	struct rdata *rd = s->rd;
	hipStream_t st = G.stream;
	int r = rl_acc_ready(s, st);
	if (r)
		return r;
	const unsigned bx = rl_blkszx(), by = rl_blkszy();
	const int refrays = bx % 4 == 0 && by % 2 == 0 &&
	  is_pow2(by / 2) && (bx / 4 == by / 2 || bx / 4 == by);
	size_t nrays = (size_t)rw * rh;
	if ((r = frame_buffers(s, nrays)))
		return r;
	struct CamArgs ca;
	rl_camargs(&rd->cam, &ca);
	struct rl_frame *f = s->frame;
	const unsigned blkcntx = rd->width / bx, nblk = blkcntx * (rd->height / by);
	if (refrays) {
		const void *lens = NULL;
		if ((r = rsqrt_table_ready(st)) || (r = lens_table_ready(&rd->cam, &lens)))
			return r;
		if ((r = rk_camrays_ref(&ca, bx, by, blkcntx, nblk, rd->samples, lens, f->rays, st)))
			return rl_hipfail(r, "reference camera rays");
	} else if ((r = rk_camrays(&ca, rw, rh, 1, 0x2545f491u, rd->samples, f->rays, st))) {
		return rl_hipfail(r, "camera rays");
	}
	if ((r = rl_trace_async(s, f->rays, f->hits, NULL, (unsigned)nrays, 1, 0,
	  st)))
		return r;
	float bg[3] = {rd->bgcol.x, rd->bgcol.y, rd->bgcol.z};
	float inv = 1.0f / (float)(rd->spp + rd->samples);
	if (refrays)
		r = rk_shade_normals_ref(f->hits, s->insts, s->nrms, s->mtls, bg, bx, by, blkcntx,
		  nblk, rd->width, rd->height, inv, s->acc, s->buf, st);
	else
		r = rk_shade_normals(f->hits, s->insts, s->nrms, s->mtls, bg, rw, rh,
		  rd->width, 1, inv, s->acc, s->buf, st);
	if (r)
		return rl_hipfail(r, "shade");
	f->frame++;
	*rays_out = (unsigned)nrays;
	return 0;
}

int rend_render(void *d)
{
	struct rdata *rd = d;
	unsigned bx = rl_blkszx(), by = rl_blkszy();
	unsigned bcx = rd->width / bx, bcy = rd->height / by;
	int blkcnt = (int)(bcx * bcy);
	if (blkcnt <= 0)
		return 0;
	int first = __atomic_fetch_add(&rd->blknum, blkcnt, __ATOMIC_SEQ_CST);
	if (first != 0)
		return 0;

	unsigned rays = 0;
	int r = rend_hip_render_frame(rd, &rays);
	if (r) {
		/* The reference's thread entry returns 0 whatever happens; the
		 * failure stays readable through rend_hip_lasterror() (sticky)
		 * and is reported on stderr once per process. */
		static int noted;
		if (!__atomic_exchange_n(&noted, 1, __ATOMIC_SEQ_CST))
			fprintf(stderr, "librend: rend_render failed (%d): %s\n", r,
			  rend_hip_errstr());
		return 0;
	}
	__atomic_fetch_add(&rd->rays, rays, __ATOMIC_SEQ_CST);
	return 0;
}

static int g_mode = REND_MODE_NORMALS;

int rend_hip_set_mode(int mode)
{
	if (mode != REND_MODE_NORMALS && mode != REND_MODE_PATHTRACE)
		return rl_fail(REND_EINVAL, "render mode %d", mode);
	g_mode = mode;
	return 0;
}

int rend_hip_render_frame(struct rdata *rd, unsigned *rays)
{
	struct rl_scene *s = rl_scene(rd);
	int r = rl_device();
	if (r)
		return r;
	if (!s || !s->tlas_built)
		return rl_fail(REND_ESTATE, "scene not prepared");
	if ((r = rl_sync_scene(s)))
		return r;
	unsigned bx = rl_blkszx(), by = rl_blkszy();
	unsigned rw = (rd->width / bx) * bx, rh = (rd->height / by) * by;
	unsigned n = 0;
	if (g_mode == REND_MODE_PATHTRACE)
		r = rl_render_pt(s, rw, rh, &n);
	else
		r = render_normals(s, rw, rh, &n);
	if (r)
		return r;
	hipError_t e = hipStreamSynchronize(G.stream);
	if (e != hipSuccess)
		return rl_hipfail(e, "frame");
	r = rl_check_errors(G.stream);
	if (r)
		return r;
	if (rd->buf) {
		/* rd->buf has pitch = width (rend.c:3571). */
		e = hipMemcpy(rd->buf, s->buf, (size_t)rd->width * rd->height *
		  sizeof(unsigned), hipMemcpyDeviceToHost);
		if (e != hipSuccess)
			return rl_hipfail(e, "colour buffer download");
	}
	if (rd->acc) {
		/* the reference accumulates into rd->acc itself (rend.c:3571) */
		e = hipMemcpy(rd->acc, s->acc, (size_t)rd->width * rd->height *
		  3 * sizeof(float), hipMemcpyDeviceToHost);
		if (e != hipSuccess)
			return rl_hipfail(e, "accumulator download");
	}
	if (rays)
		*rays = n;
	return 0;
}

int rend_hip_gen_camrays(struct rdata *rd, struct rray *d_rays, unsigned int w,
                         unsigned int h, unsigned int spp, unsigned int seed,
                         void *stream)
{
	int r = rl_device();
	if (r)
		return r;
	struct CamArgs ca;
	rl_camargs(&rd->cam, &ca);
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	r = rk_camrays(&ca, w, h, spp ? spp : 1, seed, 0, d_rays, st);
	return r ? rl_hipfail(r, "camera rays") : 0;
}

static const unsigned g_rsqrt_xeon[8192] = RSQRT_TABLE_INIT;
static int g_rsqrt_dev = -1;  /* device the table was uploaded to */
static int g_rsqrt_mode = -1; /* 0: this host's capture, 1: the build host's */

/* RSQRTPS table of the selected host in the camera kernel's constant memory:
 * this host's (captured at set-up) unless the build host's tables are
 * selected (REND_RCP_XEON_TABLE). */
static int rsqrt_table_ready(hipStream_t st)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e)
		return rl_hipfail(e, "device");
	const int mode = G.rcp_mode == REND_RCP_XEON_TABLE;
	if (g_rsqrt_dev != dev || g_rsqrt_mode != mode) {
		int r = rk_camrays_ref_setup(mode ? g_rsqrt_xeon : G.rsqrt_host, st);
		if (r)
			return rl_hipfail(r, "rsqrt table");
		g_rsqrt_dev = dev;
		g_rsqrt_mode = mode;
	}
	return 0;
}

/* glibc libmvec, AVX2 variants: the cos/sin the reference's thin-lens
 * branch calls (rend.c:3328-3329) */
__m256 _ZGVdN8v_cosf(__m256 x);
__m256 _ZGVdN8v_sinf(__m256 x);

#define RL_LENS_N (1u << 23)  /* values of the 23-bit lens-angle mantissa */

static void *g_lens;          /* device table of RL_LENS_N (cos, sin) pairs */
static int g_lens_dev = -1;   /* device it lives on */

/* The reference's lens angle is 2 pi u3 with u3 = m 2^-23, m the 23 random
 * mantissa bits of randf8, so cos and sin take only 2^23 values: this host's
 * libmvec results for every m, computed once (8 lanes per call, as the
 * reference calls them) and kept on the device (64 MiB) until the process
 * ends.  Only needed for a thin-lens camera; *lens stays NULL otherwise. */
static int lens_table_ready(const struct rcam *c, const void **lens)
{
	*lens = NULL;
	if (!(c->tanfangle > 0.0f))
		return 0;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e)
		return rl_hipfail(e, "device");
	if (g_lens_dev != dev) {
		float *h = malloc((size_t)RL_LENS_N * 2 * sizeof(float));
		if (!h)
			return rl_fail(REND_ENOMEM, "lens table (host)");
		const __m256 twopi = _mm256_set1_ps(6.28318531f);
		const __m256i one = _mm256_set1_epi32(127 << 23);
		for (uint32_t m = 0; m < RL_LENS_N; m += 8) {
			const __m256i k = _mm256_add_epi32(_mm256_set1_epi32((int)m),
			  _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7));
			/* randf8: (m | 1.0f's exponent) - 1.0f */
			const __m256 u = _mm256_sub_ps(_mm256_castsi256_ps(_mm256_or_si256(k, one)),
			  _mm256_castsi256_ps(one));
			const __m256 th = _mm256_mul_ps(twopi, u);
			float cv[8], sv[8];
			_mm256_storeu_ps(cv, _ZGVdN8v_cosf(th));
			_mm256_storeu_ps(sv, _ZGVdN8v_sinf(th));
			for (int l = 0; l < 8; l++) {
				h[2 * (m + l)] = cv[l];
				h[2 * (m + l) + 1] = sv[l];
			}
		}
		void *d = NULL;
		e = hipMalloc(&d, (size_t)RL_LENS_N * 2 * sizeof(float));
		if (!e)
			e = hipMemcpy(d, h, (size_t)RL_LENS_N * 2 * sizeof(float), hipMemcpyHostToDevice);
		free(h);
		if (e) {
			if (d)
				hipFree(d);
			return rl_hipfail(e, "lens table");
		}
		if (g_lens)
			hipFree(g_lens);  /* a table of a device no longer current */
		g_lens = d;
		g_lens_dev = dev;
	}
	*lens = g_lens;
	return 0;
}

unsigned long long rend_hip_camrays_ref_count(struct rdata *rd)
{
	const unsigned bx = rl_blkszx(), by = rl_blkszy();
	if (!rd || bx % 4 || by % 2 || !bx || !by)
		return 0;
	return (unsigned long long)(rd->width / bx) * (rd->height / by) * (bx * by / 8) * 8;
}

int rend_hip_gen_camrays_ref(struct rdata *rd, struct rray *d_rays, unsigned int samples,
                             void *stream)
{
	int r = rl_device();
	if (r)
		return r;
	const unsigned bx = rl_blkszx(), by = rl_blkszy();
	if (bx % 4 || by % 2)
		return rl_fail(REND_EINVAL, "block %ux%u is not made of 4x2 packets", bx, by);
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	const void *lens = NULL;
	if ((r = rsqrt_table_ready(st)) || (r = lens_table_ready(&rd->cam, &lens)))
		return r;
	struct CamArgs ca;
	rl_camargs(&rd->cam, &ca);
	const unsigned blkcntx = rd->width / bx, nblk = blkcntx * (rd->height / by);
	r = rk_camrays_ref(&ca, bx, by, blkcntx, nblk, samples, lens, d_rays, st);
	return r ? rl_hipfail(r, "reference camera rays") : 0;
}

int rend_hip_gen_diffuse(struct rdata *rd, const struct rray *d_prim,
                         const struct rhit *d_hits, struct rray *d_out,
                         unsigned int n, unsigned int seed, void *stream)
{
	struct rl_scene *s = rl_scene(rd);
	int r = rl_device();
	if (r)
		return r;
	if (!s || !s->tlas_built)
		return rl_fail(REND_ESTATE, "scene not prepared");
	if ((r = rl_sync_scene(s)))
		return r;
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	r = rk_diffuse_rays(d_prim, d_hits, s->insts, s->tris, n, seed, d_out, st);
	return r ? rl_hipfail(r, "diffuse rays") : 0;
}

int rk_copy(const void *src, void *dst, unsigned long long bytes, void *stream);

int rend_hip_copy(const void *d_src, void *d_dst, unsigned long long bytes,
                  void *stream)
{
	int r = rl_device();
	if (r)
		return r;
	hipStream_t st = stream ? (hipStream_t)stream : G.stream;
	r = rk_copy(d_src, d_dst, bytes, st);
	return r ? rl_hipfail(r, "copy") : 0;
}

int rend_hip_download_acc(struct rdata *rd)
{
	struct rl_scene *s = rl_scene(rd);
	if (!s || !s->acc || !rd->acc)
		return rl_fail(REND_ESTATE, "no accumulator");
	hipError_t e = hipMemcpy(rd->acc, s->acc, (size_t)rd->width * rd->height *
	  3 * sizeof(float), hipMemcpyDeviceToHost);
	return e == hipSuccess ? 0 : rl_hipfail(e, "accumulator download");
}
