/*
 * ORACLE (test infrastructure only): FP64 path-tracing reference images.
 *
 * The same estimator as the GPU wavefront path tracer, evaluated in double
 * precision with an independent traversal and random number stream, so a
 * converged GPU image can be compared with it by relMSE:
 *   - camera: jittered pinhole rays of make_camray_pckt (rend.c:3293-3318)
 *     with exact normalisation;
 *   - Lambertian BRDF albedo/pi (trace3, rend.c:3240-3287), shading normal
 *     interpolated as calc_nrm (rend.c:3109-3116), both normals turned
 *     towards the incoming ray;
 *   - next-event estimation to emissive triangles chosen by area x luminance,
 *     uniform on the triangle, shadow ray from p + 1e-3 l (EPS2, util.h:8)
 *     ending 2e-3 short of the light;
 *   - emission counted for camera rays only, background radiance on a miss
 *     (rend.c:3220), cosine-weighted continuation (rand_hemicos,
 *     rend.c:3135-3140), Russian roulette with survival min(max(thr), 0.95)
 *     from rr_depth on, at most maxdepth segments.
 * Geometry is read from the reference data model (struct rdata, bnode8,
 * rtri, rnrm, rinst).  Traversal uses conservative double slab tests and a
 * double Moeller-Trumbore test.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rend.h"

#define LEAF 0x80000000u
#define EMISSIVE 0x1000u
#define EPS 1e-3
#define PI64 3.14159265358979323846

struct dray {
	double o[3], d[3], inv[3];
};

struct pt64_cam {
	double eye[3], ri[3], up[3], fwd[3];
	double rw, rh, aspect, fdist, fovfdist;
};

struct dlight {
	double v0[3], e0[3], e1[3], n[3], E[3], area, pick, cdf;
};

struct pt64 {
	const struct rdata *rd;
	struct pt64_cam cam;
	unsigned w, h, spp, maxdepth, rr_depth;
	unsigned long long seed;
	double bg[3];
	struct dlight *lights;
	unsigned nlights;
	double (*fwd)[12];   /* world matrix (inverse of globinv) per instance */
	int use_tlas;        /* prune instances with the host TLAS */
	double *img;
	unsigned next_row;
};

/* ---- random numbers: splitmix64 ----------------------------------------- */

static uint64_t sm64(uint64_t *s)
{
	uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static double rnd(uint64_t *s)
{
	return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0);
}

/* ---- geometry ------------------------------------------------------------- */

static double dot3(const double *a, const double *b)
{
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void norm3(double *a)
{
	double l = sqrt(dot3(a, a));
	a[0] /= l;
	a[1] /= l;
	a[2] /= l;
}

static void mk_ray(struct dray *r, const double o[3], const double d[3])
{
	for (int k = 0; k < 3; k++) {
		r->o[k] = o[k];
		r->d[k] = d[k];
		r->inv[k] = 1.0 / d[k];
	}
}

/* Conservative slab test: NaN (0 * inf) terms are ignored. */
static int box_hit(const struct dray *r, double mn[3], double mx[3], double tmax,
                   double *tnear)
{
	double lo = 0.0, hi = tmax;
	for (int k = 0; k < 3; k++) {
		double a = (mn[k] - r->o[k]) * r->inv[k];
		double b = (mx[k] - r->o[k]) * r->inv[k];
		double t0 = fmin(a, b), t1 = fmax(a, b);
		if (t0 == t0 && t0 > lo)
			lo = t0;
		if (t1 == t1 && t1 < hi)
			hi = t1;
	}
	*tnear = lo;
	return lo <= hi * (1.0 + 1e-12) + 1e-12;
}

static int tri_hit(const struct dray *r, const struct rtri *t, double tmax,
                   double *tt, double *uu, double *vv)
{
	double v0[3] = {t->v0.x, t->v0.y, t->v0.z};
	double e0[3] = {t->v1.x - v0[0], t->v1.y - v0[1], t->v1.z - v0[2]};
	double e1[3] = {t->v2.x - v0[0], t->v2.y - v0[1], t->v2.z - v0[2]};
	double pv[3] = {r->d[1] * e1[2] - r->d[2] * e1[1],
	  r->d[2] * e1[0] - r->d[0] * e1[2], r->d[0] * e1[1] - r->d[1] * e1[0]};
	double det = dot3(e0, pv);
	if (fabs(det) < 1e-300)
		return 0;
	double idet = 1.0 / det;
	double tv[3] = {r->o[0] - v0[0], r->o[1] - v0[1], r->o[2] - v0[2]};
	double u = dot3(tv, pv) * idet;
	if (u < 0.0 || u > 1.0)
		return 0;
	double qv[3] = {tv[1] * e0[2] - tv[2] * e0[1], tv[2] * e0[0] - tv[0] * e0[2],
	  tv[0] * e0[1] - tv[1] * e0[0]};
	double v = dot3(r->d, qv) * idet;
	if (v < 0.0 || u + v > 1.0)
		return 0;
	double th = dot3(e1, qv) * idet;
	if (!(th > 0.0 && th < tmax))
		return 0;
	*tt = th;
	*uu = u;
	*vv = v;
	return 1;
}

#define F(v, i) ((double)((const float *)&(v))[i])
#define U(v, i) (((const unsigned *)&(v))[i])

/* Closest (any = 0) or any (any = 1) hit in one BLAS. */
static int blas64(const struct rdata *rd, unsigned triofs, const struct dray *r,
                  double *t, double *u, double *v, unsigned *tri, int any)
{
	const unsigned char *base = (const unsigned char *)&rd->bnodes[(size_t)triofs << 1];
	unsigned stack[256];
	unsigned sp = 0;
	int hit = 0;
	stack[sp++] = 0;
	while (sp) {
		unsigned ofs = stack[--sp];
		if (ofs & LEAF) {
			const struct leaf4 *l = (const struct leaf4 *)(base + (ofs & ~LEAF));
			for (unsigned k = 0; k < l->tricnt; k++) {
				unsigned id = U(l->id4, k);
				double tt, uu, vv;
				if (tri_hit(r, &rd->tris[triofs + id], *t, &tt, &uu, &vv)) {
					*t = tt;
					*u = uu;
					*v = vv;
					*tri = id;
					hit = 1;
					if (any)
						return 1;
				}
			}
			continue;
		}
		const struct bnode8 *n = (const struct bnode8 *)(base + ofs);
		for (unsigned i = 0; i < 8; i++) {
			unsigned c = U(n->children8, i);
			if (c == 0)
				continue;
			double mn[3] = {F(n->minx8, i), F(n->miny8, i), F(n->minz8, i)};
			double mx[3] = {F(n->maxx8, i), F(n->maxy8, i), F(n->maxz8, i)};
			double tn;
			if (box_hit(r, mn, mx, *t, &tn) && sp < 256)
				stack[sp++] = c;
		}
	}
	return hit;
}

/* One instance: the world ray through globinv into the BLAS. */
static int inst64(const struct rdata *rd, unsigned i, const struct dray *wr,
                  double *t, double *u, double *v, unsigned *tri, int any)
{
	const struct aabb *b = &rd->aabbs[i];
	double mn[3] = {b->min.x, b->min.y, b->min.z};
	double mx[3] = {b->max.x, b->max.y, b->max.z};
	double tn;
	if (!box_hit(wr, mn, mx, *t, &tn))
		return 0;
	const float *g = rd->insts[i].globinv;
	double o[3], d[3];
	for (int k = 0; k < 3; k++) {
		o[k] = g[4 * k] * wr->o[0] + g[4 * k + 1] * wr->o[1] +
		  g[4 * k + 2] * wr->o[2] + g[4 * k + 3];
		d[k] = g[4 * k] * wr->d[0] + g[4 * k + 1] * wr->d[1] +
		  g[4 * k + 2] * wr->d[2];
	}
	struct dray br;
	mk_ray(&br, o, d);
	return blas64(rd, rd->insts[i].triofs, &br, t, u, v, tri, any);
}

/* Closest or any hit over all instances.  The TLAS (rd->bnodes at tlasofs,
 * children8: byte offsets, leaves NODE_LEAF | instance id) only prunes:
 * its boxes enclose the instance boxes, each tested conservatively, so the
 * result is that of testing every instance. */
static int scene64(const struct pt64 *P, const struct dray *wr, double tmax,
                   double *t, double *u, double *v, unsigned *inst,
                   unsigned *tri, int any)
{
	const struct rdata *rd = P->rd;
	*t = tmax;
	int hit = 0;
	if (P->use_tlas) {
		const unsigned char *base = (const unsigned char *)&rd->bnodes[rd->tlasofs];
		unsigned stack[512];
		unsigned sp = 0;
		stack[sp++] = 0;
		while (sp) {
			unsigned ofs = stack[--sp];
			if (ofs & LEAF) {
				unsigned i = ofs & ~LEAF;
				if (i < rd->instcnt && inst64(rd, i, wr, t, u, v, tri, any)) {
					*inst = i;
					hit = 1;
					if (any)
						return 1;
				}
				continue;
			}
			const struct bnode8 *n = (const struct bnode8 *)(base + ofs);
			for (unsigned c8 = 0; c8 < 8; c8++) {
				unsigned c = U(n->children8, c8);
				if (c == 0)
					continue;
				double mn[3] = {F(n->minx8, c8), F(n->miny8, c8), F(n->minz8, c8)};
				double mx[3] = {F(n->maxx8, c8), F(n->maxy8, c8), F(n->maxz8, c8)};
				double tn;
				if (box_hit(wr, mn, mx, *t, &tn) && sp < 512)
					stack[sp++] = c;
			}
		}
		return hit;
	}
	for (unsigned i = 0; i < rd->instcnt; i++) {
		const struct aabb *b = &rd->aabbs[i];
		double mn[3] = {b->min.x, b->min.y, b->min.z};
		double mx[3] = {b->max.x, b->max.y, b->max.z};
		double tn;
		if (!box_hit(wr, mn, mx, *t, &tn))
			continue;
		const float *g = rd->insts[i].globinv;
		double o[3], d[3];
		for (int k = 0; k < 3; k++) {
			o[k] = g[4 * k] * wr->o[0] + g[4 * k + 1] * wr->o[1] +
			  g[4 * k + 2] * wr->o[2] + g[4 * k + 3];
			d[k] = g[4 * k] * wr->d[0] + g[4 * k + 1] * wr->d[1] +
			  g[4 * k + 2] * wr->d[2];
		}
		struct dray br;
		mk_ray(&br, o, d);
		if (blas64(rd, rd->insts[i].triofs, &br, t, u, v, tri, any)) {
			*inst = i;
			hit = 1;
			if (any)
				return 1;
		}
	}
	return hit;
}

/* normal of object-space vector through the inverse transpose of globinv */
static void nrm_world(const float *g, const double in[3], double out[3])
{
	for (int k = 0; k < 3; k++)
		out[k] = g[k] * in[0] + g[4 + k] * in[1] + g[8 + k] * in[2];
	norm3(out);
}

static void trace_path(const struct pt64 *P, const double o0[3], const double d0[3],
                       uint64_t *rs, double L[3])
{
	const struct rdata *rd = P->rd;
	double thr[3] = {1, 1, 1};
	double o[3] = {o0[0], o0[1], o0[2]}, d[3] = {d0[0], d0[1], d0[2]};
	L[0] = L[1] = L[2] = 0.0;
	for (unsigned depth = 0; depth < P->maxdepth; depth++) {
		struct dray r;
		mk_ray(&r, o, d);
		double t, u, v;
		unsigned inst = 0, tri = 0;
		if (!scene64(P, &r, DBL_MAX, &t, &u, &v, &inst, &tri, 0)) {
			for (int k = 0; k < 3; k++)
				L[k] += thr[k] * P->bg[k];
			return;
		}
		const struct rinst *ri = &rd->insts[inst];
		const struct rtri *T = &rd->tris[ri->triofs + tri];
		const struct rnrm *rn = &rd->nrms[ri->triofs + tri];
		const struct rmtl *mt = &rd->mtls[rn->mtlid];
		double e0[3] = {T->v1.x - T->v0.x, T->v1.y - T->v0.y, T->v1.z - T->v0.z};
		double e1[3] = {T->v2.x - T->v0.x, T->v2.y - T->v0.y, T->v2.z - T->v0.z};
		double g0[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
		  e0[0] * e1[1] - e0[1] * e1[0]};
		double g[3];
		nrm_world(ri->globinv, g0, g);
		int front = dot3(g, d) < 0.0;
		if (mt->flags & EMISSIVE) {
			if (depth == 0 && front) {
				L[0] += thr[0] * mt->col.x;
				L[1] += thr[1] * mt->col.y;
				L[2] += thr[2] * mt->col.z;
			}
			return;
		}
		double w = 1.0 - u - v;
		double s0[3] = {rn->n1.x * u + rn->n2.x * v + rn->n0.x * w,
		  rn->n1.y * u + rn->n2.y * v + rn->n0.y * w,
		  rn->n1.z * u + rn->n2.z * v + rn->n0.z * w};
		double n[3];
		nrm_world(ri->globinv, s0, n);
		if (!front)
			for (int k = 0; k < 3; k++)
				g[k] = -g[k];
		if (dot3(n, g) < 0.0)
			for (int k = 0; k < 3; k++)
				n[k] = -n[k];
		double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
		double alb[3] = {mt->col.x, mt->col.y, mt->col.z};
		if (P->nlights) {
			double u0 = rnd(rs), u1 = rnd(rs), u2 = rnd(rs);
			unsigned k = 0;
			while (k + 1 < P->nlights && P->lights[k].cdf <= u0)
				k++;
			const struct dlight *Lt = &P->lights[k];
			double su = sqrt(u1), b0 = su * (1.0 - u2), b1 = su * u2;
			double q[3], l[3];
			for (int j = 0; j < 3; j++) {
				q[j] = Lt->v0[j] + Lt->e0[j] * b0 + Lt->e1[j] * b1;
				l[j] = q[j] - p[j];
			}
			double d2 = dot3(l, l), dist = sqrt(d2);
			for (int j = 0; j < 3; j++)
				l[j] /= dist;
			double cs = dot3(n, l), cl = -dot3(Lt->n, l);
			if (cs > 0.0 && cl > 0.0 && dot3(g, l) > 0.0 && dist > 4.0 * EPS) {
				double so[3] = {p[0] + l[0] * EPS, p[1] + l[1] * EPS, p[2] + l[2] * EPS};
				struct dray sr;
				mk_ray(&sr, so, l);
				double st, su2, sv;
				unsigned si, stt;
				if (!scene64(P, &sr, dist * (1.0 - 1e-4) - 2.0 * EPS, &st, &su2, &sv,
				  &si, &stt, 1)) {
					double f = (1.0 / PI64) * cs * cl / (d2 * (Lt->pick / Lt->area));
					for (int j = 0; j < 3; j++)
						L[j] += thr[j] * alb[j] * Lt->E[j] * f;
				}
			}
		}
		if (depth + 1 >= P->maxdepth)
			return;
		double sgn = n[2] >= 0.0 ? 1.0 : -1.0;
		double a = -1.0 / (sgn + n[2]), b = n[0] * n[1] * a;
		double b1[3] = {1.0 + sgn * n[0] * n[0] * a, sgn * b, -sgn * n[0]};
		double b2[3] = {b, sgn + n[1] * n[1] * a, -n[1]};
		double phi = 2.0 * PI64 * rnd(rs), r1 = rnd(rs);
		double sr1 = sqrt(r1), lx = cos(phi) * sr1, ly = sin(phi) * sr1, lz = sqrt(1.0 - r1);
		double nd[3];
		for (int j = 0; j < 3; j++)
			nd[j] = b1[j] * lx + b2[j] * ly + n[j] * lz;
		if (dot3(g, nd) <= 0.0)
			return;
		for (int j = 0; j < 3; j++)
			thr[j] *= alb[j];
		if (depth + 1 >= P->rr_depth) {
			double mx = fmax(thr[0], fmax(thr[1], thr[2]));
			double qs = fmin(mx, 0.95);
			if (rnd(rs) >= qs)
				return;
			for (int j = 0; j < 3; j++)
				thr[j] /= qs;
		}
		for (int j = 0; j < 3; j++) {
			o[j] = p[j] + nd[j] * EPS;
			d[j] = nd[j];
		}
	}
}

static void *row_worker(void *arg)
{
	struct pt64 *P = arg;
	for (;;) {
		unsigned y = __atomic_fetch_add(&P->next_row, 1u, __ATOMIC_SEQ_CST);
		if (y >= P->h)
			break;
		for (unsigned x = 0; x < P->w; x++) {
			uint64_t rs = P->seed ^ ((uint64_t)y << 32 | x) * 0x9e3779b97f4a7c15ull;
			double acc[3] = {0, 0, 0};
			for (unsigned s = 0; s < P->spp; s++) {
				const struct pt64_cam *c = &P->cam;
				double u0 = rnd(&rs), u1 = rnd(&rs);
				double sx = (((double)x + u0 - 0.5) * c->rw - 0.5) * c->fovfdist * c->aspect;
				double sy = (((double)y + u1 - 0.5) * c->rh - 0.5) * c->fovfdist;
				double d[3];
				for (int k = 0; k < 3; k++)
					d[k] = c->fwd[k] * c->fdist + c->ri[k] * sx - c->up[k] * sy;
				norm3(d);
				double L[3];
				trace_path(P, c->eye, d, &rs, L);
				for (int k = 0; k < 3; k++)
					acc[k] += L[k];
			}
			double *o = &P->img[3 * ((size_t)y * P->w + x)];
			for (int k = 0; k < 3; k++)
				o[k] = acc[k] / P->spp;
		}
	}
	return NULL;
}

static int build_lights(struct pt64 *P)
{
	const struct rdata *rd = P->rd;
	unsigned cap = 64, n = 0;
	struct dlight *L = malloc(cap * sizeof(*L));
	if (!L)
		return -1;
	for (unsigned i = 0; i < rd->instcnt; i++) {
		const float *g = rd->insts[i].globinv;
		double a[9] = {g[0], g[1], g[2], g[4], g[5], g[6], g[8], g[9], g[10]};
		double det = a[0] * (a[4] * a[8] - a[5] * a[7]) -
		  a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
		double m[9] = {
		  (a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
		  (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
		  (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
		  (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
		  (a[0] * a[4] - a[1] * a[3]) / det};
		double tr[3] = {g[3], g[7], g[11]};
		for (unsigned k = 0; k < rd->insts[i].tricnt; k++) {
			const struct rnrm *rn = &rd->nrms[rd->insts[i].triofs + k];
			const struct rmtl *mt = &rd->mtls[rn->mtlid];
			if (!(mt->flags & EMISSIVE))
				continue;
			const struct rtri *T = &rd->tris[rd->insts[i].triofs + k];
			const struct vec3 *vv[3] = {&T->v0, &T->v1, &T->v2};
			double w[3][3];
			for (int q = 0; q < 3; q++) {
				double po[3] = {vv[q]->x, vv[q]->y, vv[q]->z};
				for (int r = 0; r < 3; r++)
					w[q][r] = m[3 * r] * (po[0] - tr[0]) + m[3 * r + 1] * (po[1] - tr[1]) +
					  m[3 * r + 2] * (po[2] - tr[2]);
			}
			if (n == cap) {
				cap *= 2;
				struct dlight *nl = realloc(L, cap * sizeof(*L));
				if (!nl) {
					free(L);
					return -1;
				}
				L = nl;
			}
			struct dlight *l = &L[n++];
			for (int r = 0; r < 3; r++) {
				l->v0[r] = w[0][r];
				l->e0[r] = w[1][r] - w[0][r];
				l->e1[r] = w[2][r] - w[0][r];
			}
			l->n[0] = l->e0[1] * l->e1[2] - l->e0[2] * l->e1[1];
			l->n[1] = l->e0[2] * l->e1[0] - l->e0[0] * l->e1[2];
			l->n[2] = l->e0[0] * l->e1[1] - l->e0[1] * l->e1[0];
			double len = sqrt(dot3(l->n, l->n));
			for (int r = 0; r < 3; r++)
				l->n[r] /= len;
			l->area = 0.5 * len;
			l->E[0] = mt->col.x;
			l->E[1] = mt->col.y;
			l->E[2] = mt->col.z;
			l->pick = l->area * (0.2126 * l->E[0] + 0.7152 * l->E[1] + 0.0722 * l->E[2]);
		}
	}
	double tot = 0.0, acc = 0.0;
	for (unsigned i = 0; i < n; i++)
		tot += L[i].pick;
	for (unsigned i = 0; i < n; i++) {
		acc += L[i].pick;
		L[i].pick /= tot;
		L[i].cdf = acc / tot;
	}
	P->lights = L;
	P->nlights = n;
	return 0;
}

/* The host TLAS may prune only if it is sound: every instance reachable
 * exactly once, and every child box containing the boxes below it (the
 * builder's degenerate splits can leave stale boxes).  Returns 1 if so. */
static int tlas_check(const struct rdata *rd, unsigned ofs, const double *bmin,
                      const double *bmax, unsigned char *seen, int depth)
{
	if (depth > 64)
		return 0;
	if (ofs & LEAF) {
		unsigned i = ofs & ~LEAF;
		if (i >= rd->instcnt || seen[i])
			return 0;
		seen[i] = 1;
		const struct aabb *b = &rd->aabbs[i];
		double mn[3] = {b->min.x, b->min.y, b->min.z}, mx[3] = {b->max.x, b->max.y, b->max.z};
		for (int k = 0; k < 3; k++)
			if (bmin && (mn[k] < bmin[k] || mx[k] > bmax[k]))
				return 0;
		return 1;
	}
	if (ofs % 32 || ofs >= 2u * (rd->instcnt + 1) * sizeof(struct bnode8))
		return 0;
	const struct bnode8 *n = (const struct bnode8 *)((const unsigned char *)
	  &rd->bnodes[rd->tlasofs] + ofs);
	for (unsigned c8 = 0; c8 < 8; c8++) {
		unsigned c = U(n->children8, c8);
		if (c == 0)
			continue;
		double mn[3] = {F(n->minx8, c8), F(n->miny8, c8), F(n->minz8, c8)};
		double mx[3] = {F(n->maxx8, c8), F(n->maxy8, c8), F(n->maxz8, c8)};
		for (int k = 0; k < 3; k++)
			if (bmin && (mn[k] < bmin[k] || mx[k] > bmax[k]))
				return 0;
		if (!tlas_check(rd, c, mn, mx, seen, depth + 1))
			return 0;
	}
	return 1;
}

static int tlas_sound(const struct rdata *rd)
{
	if (rd->instcnt < 2)
		return 0;
	unsigned char *seen = calloc(rd->instcnt, 1);
	if (!seen)
		return 0;
	int ok = tlas_check(rd, 0, NULL, NULL, seen, 0);
	for (unsigned i = 0; ok && i < rd->instcnt; i++)
		ok = seen[i];
	free(seen);
	return ok;
}

static int g_tlas_mode = 1;

/* 0: test every instance; 1: prune with the host TLAS where it is sound. */
void or_pt64_tlas(int mode)
{
	g_tlas_mode = mode;
}

/* Renders rd's camera (cam: eye, ri, up, fwd, rw, rh, aspect, fdist,
 * fovfdist as 17 doubles) at w x h with spp samples per pixel into img
 * (3 doubles per pixel, averaged). */
int or_pt64(const struct rdata *rd, const double *cam, unsigned w, unsigned h,
            unsigned spp, unsigned maxdepth, unsigned rr_depth,
            unsigned long long seed, int threads, double *img)
{
	struct pt64 P;
	memset(&P, 0, sizeof(P));
	P.rd = rd;
	memcpy(P.cam.eye, cam, 3 * sizeof(double));
	memcpy(P.cam.ri, cam + 3, 3 * sizeof(double));
	memcpy(P.cam.up, cam + 6, 3 * sizeof(double));
	memcpy(P.cam.fwd, cam + 9, 3 * sizeof(double));
	P.cam.rw = cam[12];
	P.cam.rh = cam[13];
	P.cam.aspect = cam[14];
	P.cam.fdist = cam[15];
	P.cam.fovfdist = cam[16];
	P.w = w;
	P.h = h;
	P.spp = spp;
	P.maxdepth = maxdepth;
	P.rr_depth = rr_depth ? rr_depth : 1000;
	P.seed = seed;
	P.bg[0] = rd->bgcol.x;
	P.bg[1] = rd->bgcol.y;
	P.bg[2] = rd->bgcol.z;
	P.img = img;
	P.use_tlas = g_tlas_mode && rd->instcnt > 16 && tlas_sound(rd);
	if (build_lights(&P))
		return -1;
	if (threads < 1)
		threads = 1;
	if (threads > 256)
		threads = 256;
	pthread_t th[256];
	int started = 0;
	for (int i = 1; i < threads; i++)
		if (!pthread_create(&th[i], NULL, row_worker, &P))
			started = i;
		else
			break;
	row_worker(&P);
	for (int i = 1; i <= started; i++)
		pthread_join(th[i], NULL);
	free(P.lights);
	return 0;
}
