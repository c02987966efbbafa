/*
 * ORACLE (test infrastructure only; never linked into librend).
 *
 * Scalar C restatement of the reference's single-ray traversal over the
 * reference's own host data model (struct rdata, bnode8 with perm8, leaf4,
 * rinst), used to check the GPU kernels bit for bit:
 *   intersect_blas        rend.c:714-968   -> blas_closest
 *   intersect_tlas        rend.c:1976-2155 -> tlas_closest
 *   intersect_any_blas    rend.c:1797-1973 -> blas_any
 *   intersect_any_tlas    rend.c:2883-2998 -> tlas_any
 *   mulpos / muldir       rend.c:171-187   -> xform
 * It keeps the reference's eager stack compaction and perm8-ordered pushes
 * (the GPU kernel uses lazy culling and an inverse permutation instead), the
 * x86 MAXPS/MINPS operand rules, the FMA shapes of the intrinsics and the
 * reference's reciprocal: this host's RCPSS instruction itself (mode 0, the
 * reference on this host), the correctly rounded 1/det (mode 1), or the
 * build host's RCPSS restated from its table (mode 2; the table is verified
 * against the instruction for all inputs by gen_rcp_table.c on that host).
 * Compile with -ffp-contract=off.
 *
 * Visit counters (interior nodes, leaf4 records, instance transforms) give
 * the algorithmic traffic of each ray.  Batches run on N threads claiming
 * 64-ray chunks from an atomic counter, as the reference claims blocks
 * (rend.c:3623).
 */
#include <float.h>
#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "rend.h"
#include "rend_rays.h"
#define RCP_TABLE_NO_HOST_ARRAY
#include "../nodebleed_amd/csrc/rcp_table.h"

static const unsigned int rcp_tab[4096] = RCP_TABLE_INIT;

#define LEAF 0x80000000u
#define F(v, i) (((const float *)&(v))[i])
#define U(v, i) (((const unsigned *)&(v))[i])

struct ocnt {
	unsigned long long nodes, leaves, insts;
};

static int g_rcp_mode;  /* 0 host instruction, 1 IEEE, 2 build host's table */

static float bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* _mm_rcp_ps on the build host (rule in gen_rcp_table.c). */
float or_rcp(float x)
{
	uint32_t u = f2bits(x), s = u & 0x80000000u, e = (u >> 23) & 0xffu,
	  m = u & 0x7fffffu;
	if (e == 0)
		return bits2f(s | 0x7f800000u);
	if (e == 255)
		return bits2f(m ? (u | 0x00400000u) : s);
	if (e >= 253)
		return bits2f(s);
	return bits2f(s | ((253u - e) << 23) | rcp_tab[m >> 11]);
}

/* The reciprocal of the triangle test in the current mode. */
static float det_rcp(float x)
{
	if (g_rcp_mode == 1)
		return 1.0f / x;
	if (g_rcp_mode == 2)
		return or_rcp(x);
	return _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
}

/* MAXPS/MINPS: second operand unless the first compares greater/less. */
static float xmax(float a, float b) { return a > b ? a : b; }
static float xmin(float a, float b) { return a < b ? a : b; }

struct ray {
	float ox, oy, oz, dx, dy, dz, idx, idy, idz, rx, ry, rz;
	int bx, by, bz;
};

static void ray_init(struct ray *r, float ox, float oy, float oz, float dx,
                     float dy, float dz)
{
	r->ox = ox; r->oy = oy; r->oz = oz;
	r->dx = dx; r->dy = dy; r->dz = dz;
	r->idx = 1.0f / dx;
	r->idy = 1.0f / dy;
	r->idz = 1.0f / dz;
	r->rx = ox * r->idx;
	r->ry = oy * r->idy;
	r->rz = oz * r->idz;
	r->bx = dx >= 0.0f;
	r->by = dy >= 0.0f;
	r->bz = dz >= 0.0f;
}

/* 8-wide slab test of rend.c:757-784; returns the hit mask. */
static unsigned slab8(const struct bnode8 *n, const struct ray *r, float t,
                      float tmin[8])
{
	unsigned hm = 0;
	for (unsigned i = 0; i < 8; i++) {
		float t0x = fmaf(r->bx ? F(n->minx8, i) : F(n->maxx8, i), r->idx, -r->rx);
		float t0y = fmaf(r->by ? F(n->miny8, i) : F(n->maxy8, i), r->idy, -r->ry);
		float t0z = fmaf(r->bz ? F(n->minz8, i) : F(n->maxz8, i), r->idz, -r->rz);
		float t1x = fmaf(r->bx ? F(n->maxx8, i) : F(n->minx8, i), r->idx, -r->rx);
		float t1y = fmaf(r->by ? F(n->maxy8, i) : F(n->miny8, i), r->idy, -r->ry);
		float t1z = fmaf(r->bz ? F(n->maxz8, i) : F(n->minz8, i), r->idz, -r->rz);
		float lo = xmax(xmax(xmax(t0x, t0y), t0z), 0.0f);
		float hi = xmin(xmin(xmin(t1x, t1y), t1z), t);
		tmin[i] = lo;
		if (lo <= hi)
			hm |= 1u << i;
	}
	return hm;
}

/* Moeller-Trumbore on one leaf4 (rend.c:846-905); writes per-lane t,u,v and
 * returns the accept mask. */
static unsigned mt4(const struct leaf4 *l, const struct ray *r, float tcur,
                    float t4[4], float u4[4], float v4[4])
{
	unsigned mask = 0;
	for (unsigned k = 0; k < 4; k++) {
		float e0x = F(l->e0x4, k), e0y = F(l->e0y4, k), e0z = F(l->e0z4, k);
		float e1x = F(l->e1x4, k), e1y = F(l->e1y4, k), e1z = F(l->e1z4, k);
		float pvx = fmaf(r->dy, e1z, -(r->dz * e1y));
		float pvy = fmaf(r->dz, e1x, -(r->dx * e1z));
		float pvz = fmaf(r->dx, e1y, -(r->dy * e1x));
		float tvx = r->ox - F(l->v0x4, k);
		float tvy = r->oy - F(l->v0y4, k);
		float tvz = r->oz - F(l->v0z4, k);
		float det = fmaf(e0x, pvx, fmaf(e0y, pvy, e0z * pvz));
		float qvx = fmaf(tvy, e0z, -(tvz * e0y));
		float qvy = fmaf(tvz, e0x, -(tvx * e0z));
		float qvz = fmaf(tvx, e0y, -(tvy * e0x));
		float idet = det_rcp(det);
		float u = idet * fmaf(tvx, pvx, fmaf(tvy, pvy, tvz * pvz));
		float v = idet * fmaf(r->dx, qvx, fmaf(r->dy, qvy, r->dz * qvz));
		float t = idet * fmaf(e1x, qvx, fmaf(e1y, qvy, e1z * qvz));
		t4[k] = t;
		u4[k] = u;
		v4[k] = v;
		if (u >= 0.0f && v >= 0.0f && (u + v) <= 1.0f && t > 0.0f && t < tcur)
			mask |= 1u << k;
	}
	return mask;
}

/* Keep stack entries whose distance is <= t, in order (rend.c:928-959). */
static unsigned compact(unsigned *stack, float *dstack, unsigned sp, float t)
{
	unsigned k = 0;
	for (unsigned i = 0; i < sp; i++)
		if (dstack[i] <= t) {
			stack[k] = stack[i];
			dstack[k] = dstack[i];
			k++;
		}
	return k;
}

static int blas_closest(float *t, float *u, float *v, unsigned *id,
                        const struct ray *r, const unsigned char *base,
                        unsigned instid, struct ocnt *c)
{
	unsigned stack[64];
	float dstack[64];
	unsigned sp = 0, ofs = 0;
	unsigned shft = (unsigned)((r->bz << 2) | (r->by << 1) | r->bx) * 3;
	for (;;) {
		while (!(ofs & LEAF)) {
			const struct bnode8 *n = (const struct bnode8 *)(base + ofs);
			float tmin[8];
			c->nodes++;
			unsigned hm = slab8(n, r, *t, tmin);
			unsigned cnt = (unsigned)__builtin_popcount(hm);
			if (cnt == 0) {
				if (!sp)
					return 0;
				ofs = stack[--sp];
			} else if (cnt == 1) {
				ofs = U(n->children8, 31 - __builtin_clz(hm));
			} else {
				if (sp + cnt > 64)
					return -1;
				unsigned k = 0;
				for (unsigned i = 0; i < 8; i++) {
					unsigned ch = (U(n->perm8, i) >> shft) & 7u;
					if ((hm >> ch) & 1u) {
						stack[sp + k] = U(n->children8, ch);
						dstack[sp + k] = tmin[ch];
						k++;
					}
				}
				sp += cnt - 1;
				ofs = stack[sp];
			}
		}
		const struct leaf4 *l = (const struct leaf4 *)(base + (ofs & ~LEAF));
		float t4[4], u4[4], v4[4];
		c->leaves++;
		unsigned mask = mt4(l, r, *t, t4, u4, v4);
		if (mask) {
			float tb[4];
			for (unsigned k = 0; k < 4; k++)
				tb[k] = ((mask >> k) & 1u) ? t4[k] : FLT_MAX;
			float m = xmin(xmin(tb[0], tb[1]), xmin(tb[2], tb[3]));
			unsigned sel = 0;
			for (unsigned k = 0; k < 4; k++)
				if (tb[k] == m)
					sel = k;
			*t = tb[sel];
			*u = u4[sel];
			*v = v4[sel];
			*id = (U(l->id4, sel) << 12) | instid;
			sp = compact(stack, dstack, sp, *t);
		}
		if (!sp)
			return 0;
		ofs = stack[--sp];
	}
}

static void xform(const float m[12], float x, float y, float z, float *ox,
                  float *oy, float *oz, int pos)
{
	if (pos) {
		*ox = m[0] * x + m[1] * y + m[2] * z + m[3];
		*oy = m[4] * x + m[5] * y + m[6] * z + m[7];
		*oz = m[8] * x + m[9] * y + m[10] * z + m[11];
	} else {
		*ox = m[0] * x + m[1] * y + m[2] * z;
		*oy = m[4] * x + m[5] * y + m[6] * z;
		*oz = m[8] * x + m[9] * y + m[10] * z;
	}
}

static int tlas_closest(const struct rdata *rd, const struct rray *in,
                        struct rhit *out, struct ocnt *c)
{
	unsigned stack[64];
	float dstack[64];
	unsigned sp = 0, ofs = 0;
	const unsigned char *tl = (const unsigned char *)&rd->bnodes[rd->tlasofs];
	struct ray r;
	ray_init(&r, in->ox, in->oy, in->oz, in->dx, in->dy, in->dz);
	unsigned shft = (unsigned)((r.bz << 2) | (r.by << 1) | r.bx) * 3;
	float t = in->tmax, u = 0.0f, v = 0.0f;
	unsigned id = 0xffffffffu;
	int err = 0;
	for (;;) {
		while (!(ofs & LEAF)) {
			const struct bnode8 *n = (const struct bnode8 *)(tl + ofs);
			float tmin[8];
			c->nodes++;
			unsigned hm = slab8(n, &r, t, tmin);
			unsigned cnt = (unsigned)__builtin_popcount(hm);
			if (cnt == 0) {
				if (!sp)
					goto done;
				ofs = stack[--sp];
			} else if (cnt == 1) {
				ofs = U(n->children8, 31 - __builtin_clz(hm));
			} else {
				if (sp + cnt > 64) {
					err = -1;
					goto done;
				}
				unsigned k = 0;
				for (unsigned i = 0; i < 8; i++) {
					unsigned ch = (U(n->perm8, i) >> shft) & 7u;
					if ((hm >> ch) & 1u) {
						stack[sp + k] = U(n->children8, ch);
						dstack[sp + k] = tmin[ch];
						k++;
					}
				}
				sp += cnt - 1;
				ofs = stack[sp];
			}
		}
		unsigned inst = ofs & ~LEAF;
		const struct rinst *ri = &rd->insts[inst];
		struct ray br;
		float ox, oy, oz, dx, dy, dz;
		xform(ri->globinv, r.ox, r.oy, r.oz, &ox, &oy, &oz, 1);
		xform(ri->globinv, r.dx, r.dy, r.dz, &dx, &dy, &dz, 0);
		ray_init(&br, ox, oy, oz, dx, dy, dz);
		c->insts++;
		float tt = t;
		if (blas_closest(&t, &u, &v, &id, &br,
		  (const unsigned char *)&rd->bnodes[(size_t)ri->triofs << 1], inst,
		  c) < 0) {
			err = -1;
			goto done;
		}
		if (t < tt)
			sp = compact(stack, dstack, sp, t);
		if (!sp)
			goto done;
		ofs = stack[--sp];
	}
done:
	out->t = t;
	out->u = u;
	out->v = v;
	out->id = id;
	return err;
}

static int blas_any(float tfar, const struct ray *r, const unsigned char *base,
                    struct ocnt *c)
{
	unsigned stack[64];
	unsigned sp = 0, ofs = 0;
	for (;;) {
		while (!(ofs & LEAF)) {
			const struct bnode8 *n = (const struct bnode8 *)(base + ofs);
			float tmin[8];
			c->nodes++;
			unsigned hm = slab8(n, r, tfar, tmin);
			unsigned cnt = (unsigned)__builtin_popcount(hm);
			if (cnt == 0) {
				if (!sp)
					return 0;
				ofs = stack[--sp];
			} else if (cnt == 1) {
				ofs = U(n->children8, 31 - __builtin_clz(hm));
			} else {
				if (sp + cnt > 64)
					return -1;
				unsigned k = 0;
				for (unsigned i = 0; i < 8; i++)
					if ((hm >> i) & 1u)
						stack[sp + k++] = U(n->children8, i);
				sp += cnt - 1;
				ofs = stack[sp];
			}
		}
		float t4[4], u4[4], v4[4];
		c->leaves++;
		if (mt4((const struct leaf4 *)(base + (ofs & ~LEAF)), r, tfar, t4, u4,
		  v4))
			return 1;
		if (!sp)
			return 0;
		ofs = stack[--sp];
	}
}

static int tlas_any(const struct rdata *rd, const struct rray *in,
                    struct ocnt *c)
{
	unsigned stack[64];
	unsigned sp = 0, ofs = 0;
	const unsigned char *tl = (const unsigned char *)&rd->bnodes[rd->tlasofs];
	struct ray r;
	ray_init(&r, in->ox, in->oy, in->oz, in->dx, in->dy, in->dz);
	for (;;) {
		while (!(ofs & LEAF)) {
			const struct bnode8 *n = (const struct bnode8 *)(tl + ofs);
			float tmin[8];
			c->nodes++;
			unsigned hm = slab8(n, &r, in->tmax, tmin);
			unsigned cnt = (unsigned)__builtin_popcount(hm);
			if (cnt == 0) {
				if (!sp)
					return 0;
				ofs = stack[--sp];
			} else if (cnt == 1) {
				ofs = U(n->children8, 31 - __builtin_clz(hm));
			} else {
				if (sp + cnt > 64)
					return -1;
				unsigned k = 0;
				for (unsigned i = 0; i < 8; i++)
					if ((hm >> i) & 1u)
						stack[sp + k++] = U(n->children8, i);
				sp += cnt - 1;
				ofs = stack[sp];
			}
		}
		const struct rinst *ri = &rd->insts[ofs & ~LEAF];
		struct ray br;
		float ox, oy, oz, dx, dy, dz;
		xform(ri->globinv, r.ox, r.oy, r.oz, &ox, &oy, &oz, 1);
		xform(ri->globinv, r.dx, r.dy, r.dz, &dx, &dy, &dz, 0);
		ray_init(&br, ox, oy, oz, dx, dy, dz);
		c->insts++;
		int h = blas_any(in->tmax, &br,
		  (const unsigned char *)&rd->bnodes[(size_t)ri->triofs << 1], c);
		if (h)
			return h;
		if (!sp)
			return 0;
		ofs = stack[--sp];
	}
}

/* ---- batches ----------------------------------------------------------- */

struct job {
	const struct rdata *rd;
	const struct rray *rays;
	struct rhit *hits;
	unsigned char *occ;
	unsigned n;
	unsigned next;          /* atomic chunk counter */
	int errors;             /* atomic */
	unsigned long long cnt[3];
	pthread_mutex_t lock;
	unsigned long long *per_ray;  /* optional 3 counters per ray */
};

static void *worker(void *arg)
{
	struct job *j = arg;
	struct ocnt c = {0, 0, 0};
	for (;;) {
		unsigned b = __atomic_fetch_add(&j->next, 64u, __ATOMIC_SEQ_CST);
		if (b >= j->n)
			break;
		unsigned e = b + 64 < j->n ? b + 64 : j->n;
		for (unsigned i = b; i < e; i++) {
			struct ocnt before = c;
			int r;
			if (j->hits) {
				r = tlas_closest(j->rd, &j->rays[i], &j->hits[i], &c);
			} else {
				r = tlas_any(j->rd, &j->rays[i], &c);
				j->occ[i] = r > 0;
			}
			if (r < 0)
				__atomic_fetch_add(&j->errors, 1, __ATOMIC_SEQ_CST);
			if (j->per_ray) {
				j->per_ray[3 * (size_t)i + 0] = c.nodes - before.nodes;
				j->per_ray[3 * (size_t)i + 1] = c.leaves - before.leaves;
				j->per_ray[3 * (size_t)i + 2] = c.insts - before.insts;
			}
		}
	}
	pthread_mutex_lock(&j->lock);
	j->cnt[0] += c.nodes;
	j->cnt[1] += c.leaves;
	j->cnt[2] += c.insts;
	pthread_mutex_unlock(&j->lock);
	return NULL;
}

static int run(struct job *j, int threads, unsigned long long cnt[3])
{
	if (threads < 1)
		threads = 1;
	if (threads > 256)
		threads = 256;
	pthread_t th[256];
	pthread_mutex_init(&j->lock, NULL);
	int started = 0;
	for (int i = 1; i < threads; i++)
		if (!pthread_create(&th[i], NULL, worker, j))
			started = i;
		else
			break;
	worker(j);
	for (int i = 1; i <= started; i++)
		pthread_join(th[i], NULL);
	pthread_mutex_destroy(&j->lock);
	if (cnt) {
		cnt[0] = j->cnt[0];
		cnt[1] = j->cnt[1];
		cnt[2] = j->cnt[2];
	}
	return j->errors;
}

/* Closest hit of n rays; returns the number of rays whose stack overflowed.
 * cnt (optional) receives total node/leaf/instance visits; per_ray (optional)
 * receives them per ray. */
int or_trace(const struct rdata *rd, const struct rray *rays,
             struct rhit *hits, unsigned n, int rcp_mode, int threads,
             unsigned long long cnt[3], unsigned long long *per_ray)
{
	struct job j;
	memset(&j, 0, sizeof(j));
	j.rd = rd;
	j.rays = rays;
	j.hits = hits;
	j.n = n;
	j.per_ray = per_ray;
	g_rcp_mode = rcp_mode;
	return run(&j, threads, cnt);
}

int or_occluded(const struct rdata *rd, const struct rray *rays,
                unsigned char *occ, unsigned n, int rcp_mode, int threads)
{
	struct job j;
	memset(&j, 0, sizeof(j));
	j.rd = rd;
	j.rays = rays;
	j.occ = occ;
	j.n = n;
	g_rcp_mode = rcp_mode;
	return run(&j, threads, NULL);
}

/* Brute force over every triangle of every instance: the closest t under the
 * same Moeller-Trumbore arithmetic (no BVH), for small scenes. */
void or_bruteforce(const struct rdata *rd, const struct rray *rays,
                   struct rhit *hits, unsigned n, int rcp_mode)
{
	g_rcp_mode = rcp_mode;
	for (unsigned i = 0; i < n; i++) {
		const struct rray *in = &rays[i];
		float best = in->tmax, bu = 0, bv = 0;
		unsigned bid = 0xffffffffu;
		for (unsigned k = 0; k < rd->instcnt; k++) {
			const struct rinst *ri = &rd->insts[k];
			float ox, oy, oz, dx, dy, dz;
			xform(ri->globinv, in->ox, in->oy, in->oz, &ox, &oy, &oz, 1);
			xform(ri->globinv, in->dx, in->dy, in->dz, &dx, &dy, &dz, 0);
			struct ray r;
			ray_init(&r, ox, oy, oz, dx, dy, dz);
			for (unsigned t = 0; t < ri->tricnt; t++) {
				const struct rtri *tr = &rd->tris[ri->triofs + t];
				struct leaf4 l;
				memset(&l, 0, sizeof(l));
				for (int q = 0; q < 4; q++) {
					((float *)&l.v0x4)[q] = tr->v0.x;
					((float *)&l.v0y4)[q] = tr->v0.y;
					((float *)&l.v0z4)[q] = tr->v0.z;
					((float *)&l.e0x4)[q] = tr->v1.x - tr->v0.x;
					((float *)&l.e0y4)[q] = tr->v1.y - tr->v0.y;
					((float *)&l.e0z4)[q] = tr->v1.z - tr->v0.z;
					((float *)&l.e1x4)[q] = tr->v2.x - tr->v0.x;
					((float *)&l.e1y4)[q] = tr->v2.y - tr->v0.y;
					((float *)&l.e1z4)[q] = tr->v2.z - tr->v0.z;
				}
				float t4[4], u4[4], v4[4];
				if (mt4(&l, &r, best, t4, u4, v4) & 1u) {
					best = t4[0];
					bu = u4[0];
					bv = v4[0];
					bid = (t << 12) | k;
				}
			}
		}
		hits[i].t = best;
		hits[i].u = bu;
		hits[i].v = bv;
		hits[i].id = bid;
	}
}

/* The 9 cases of the slab test's operand order for tests. */
float or_xmax(float a, float b) { return xmax(a, b); }
float or_xmin(float a, float b) { return xmin(a, b); }
