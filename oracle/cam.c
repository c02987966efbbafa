/*
 * ORACLE (test infrastructure only).  C restatement of the reference's
 * primary-ray generation for pinhole and thin-lens cameras, used to check
 * the GPU's reference-exact camera rays.  Follows:
 *   rend_render's block walk and per-block seed   rend.c:3603-3640
 *   4x2-pixel packets in Morton order              rend.c:131-162, util.c:133-145
 *   rand8_init / xorshift128+ / jump / randf8      util.c:29-75, 99-107
 *   make_camray_pckt, pinhole branch               rend.c:3293-3317, 3377-3401
 *   make_camray_pckt, thin-lens branch             rend.c:3320-3376: lens
 *                                                  angle 2pi * u3 through
 *                                                  glibc libmvec's 8-wide
 *                                                  cosf/sinf, the functions
 *                                                  the reference calls
 *   _mm256_rsqrt_ps                                this host's RSQRTPS (mode 0)
 *                                                  or the build host's table
 *                                                  (mode 2, gen_rsqrt_table.c rule)
 * Fused operations of the reference (fmsub/fmadd) are fmaf here; everything
 * else is separately rounded (-ffp-contract=off).
 * Pinned by: the reference's own make_camray_pckt compiled into
 * oracle/_ref (bit-exact rays, tests/test_cpu_camera.py), the RSQRTPS rule
 * against the instruction and its error bound, and unit-length directions.
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "rend.h"
#include "rend_rays.h"
#include "../nodebleed_amd/csrc/rsqrt_table.h"

/* glibc libmvec, AVX2 variants (what rend.c:3328-3329 calls) */
__m256 _ZGVdN8v_cosf(__m256 x);
__m256 _ZGVdN8v_sinf(__m256 x);

static const uint32_t rsq_tab[8192] = RSQRT_TABLE_INIT;
static int g_rsq_mode;  /* 0 host instruction, 2 build host's table */

void or_camrays_mode(int mode)
{
	g_rsq_mode = mode;
}

static float f32(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u32(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float or_rsqrt(float x)
{
	const uint32_t u = u32(x), s = u & 0x80000000u, e = (u >> 23) & 0xffu;
	const uint32_t m = u & 0x7fffffu;
	if (e == 255u)
		return f32(m ? (u | 0x00400000u) : (s ? 0xffc00000u : 0u));
	if (e == 0u)
		return f32(s | 0x7f800000u);
	if (s)
		return f32(0xffc00000u);
	const uint32_t p = e & 1u;
	const int sh = ((int)e - (p ? 127 : 128)) / 2;
	return f32((uint32_t)((int32_t)rsq_tab[(p << 12) | (m >> 11)] - sh * (1 << 23)));
}

static float cam_rsqrt(float x)
{
	if (g_rsq_mode == 2)
		return or_rsqrt(x);
	return _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
}

struct xs { uint64_t s0, s1; };

static uint64_t xs_next(struct xs *x)
{
	uint64_t a = x->s0;
	const uint64_t b = x->s1, res = a + b;
	x->s0 = b;
	a ^= a << 23;
	x->s1 = a ^ b ^ (a >> 18) ^ (b >> 5);
	return res;
}

static struct xs xs_jump(struct xs in)
{
	static const uint64_t J[2] = {0x8a5cd789635d2dffull, 0x121fd2155c472f96ull};
	struct xs out = {0, 0};
	for (int i = 0; i < 2; i++)
		for (int b = 0; b < 64; b++) {
			if (J[i] & (1ull << b)) {
				out.s0 ^= in.s0;
				out.s1 ^= in.s1;
			}
			xs_next(&in);
		}
	return out;
}

static unsigned mort(unsigned v)
{
	v &= 0x55555555u;
	v = (v ^ (v >> 1)) & 0x33333333u;
	v = (v ^ (v >> 2)) & 0x0f0f0f0fu;
	v = (v ^ (v >> 4)) & 0x00ff00ffu;
	return (v ^ (v >> 8)) & 0x0000ffffu;
}

static float lane0(__m256 v) { return _mm256_cvtss_f32(v); }

/* Rays of block blk (bszx * bszy / 8 packets x 8 lanes) of rd's camera for
 * a frame after `samples` samples. */
void or_camrays(const struct rdata *rd, unsigned bszx, unsigned bszy, unsigned samples,
                unsigned blk, struct rray *out)
{
	const struct rcam *c = &rd->cam;
	const float eye[3] = {lane0(c->eyex8), lane0(c->eyey8), lane0(c->eyez8)};
	const float ri[3] = {lane0(c->rix8), lane0(c->riy8), lane0(c->riz8)};
	const float up[3] = {lane0(c->upx8), lane0(c->upy8), lane0(c->upz8)};
	const float fw[3] = {lane0(c->fwdx8), lane0(c->fwdy8), lane0(c->fwdz8)};
	const float rw = lane0(c->rw8), rh = lane0(c->rh8), asp = lane0(c->aspect8);
	const float fd = lane0(c->fdist8), ff = lane0(c->fovfdist8);
	const float fr = lane0(c->focrad8);
	const int lens = c->tanfangle > 0.0f;
	const unsigned blkcntx = rd->width / bszx, npk = bszx * bszy / 8;
	struct xs g[4];
	g[0].s0 = (uint64_t)(int64_t)(((int)blk + 13) * 1571);
	g[0].s1 = (uint64_t)(samples * 23u);
	for (int l = 1; l < 4; l++)
		g[l] = xs_jump(g[l - 1]);
	for (unsigned k = 0; k < npk; k++) {
		/* u0, u1: pixel jitter; u2, u3: lens sample (thin lens only) */
		float u[4][8];
		for (int d = 0; d < (lens ? 4 : 2); d++)
			for (int l = 0; l < 4; l++) {
				const uint64_t v = xs_next(&g[l]);
				u[d][2 * l] = f32(((uint32_t)v >> 9) | 0x3f800000u) - 1.0f;
				u[d][2 * l + 1] = f32(((uint32_t)(v >> 32) >> 9) | 0x3f800000u) - 1.0f;
			}
		float lc[8], ls[8];
		if (lens) {
			const __m256 th = _mm256_mul_ps(_mm256_set1_ps(6.28318531f),
			  _mm256_loadu_ps(u[3]));
			_mm256_storeu_ps(lc, _ZGVdN8v_cosf(th));
			_mm256_storeu_ps(ls, _ZGVdN8v_sinf(th));
		}
		const unsigned x = (blk % blkcntx) * bszx + mort(k) * 4;
		const unsigned y = (blk / blkcntx) * bszy + mort(k >> 1) * 2;
		for (int l = 0; l < 8; l++) {
			float sx = (float)x + (float)(l & 3), sy = (float)y + (float)(l >> 2);
			sx = sx + (u[0][l] - 0.5f);
			sy = sy + (u[1][l] - 0.5f);
			sx = fmaf(sx, rw, -0.5f);
			sy = fmaf(sy, rh, -0.5f);
			sx = sx * ff;
			sy = sy * ff;
			sx = sx * asp;
			float t[3], o[3] = {eye[0], eye[1], eye[2]};
			if (lens) {
				const float rad = sqrtf(u[2][l]);
				const float rx = lc[l] * rad, ry = ls[l] * rad;
				for (int a = 0; a < 3; a++) {
					o[a] = fmaf(fmaf(ri[a], rx, up[a] * ry), fr, eye[a]);
					t[a] = (fmaf(fw[a], fd, eye[a]) + fmaf(ri[a], sx, -(up[a] * sy))) - o[a];
				}
			} else {
				for (int a = 0; a < 3; a++)
					t[a] = fmaf(fw[a], fd, fmaf(ri[a], sx, -(up[a] * sy)));
			}
			const float rl = cam_rsqrt(fmaf(t[0], t[0], fmaf(t[1], t[1], t[2] * t[2])));
			struct rray *r = &out[k * 8 + (unsigned)l];
			r->ox = o[0];
			r->oy = o[1];
			r->oz = o[2];
			r->tmax = 3.402823466e38f;
			r->dx = t[0] * rl;
			r->dy = t[1] * rl;
			r->dz = t[2] * rl;
			r->pad = 0.0f;
		}
	}
}
