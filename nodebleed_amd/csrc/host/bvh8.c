/*
 * SAH-binned 8-wide BVH builder (host side) and the repack into the device
 * node layout.
 *
 * The tree must be byte-identical to the reference's build_bvh8
 * (rend.c:458-709): the GPU traversal visits children in the order stored
 * here, and equal-t hits are resolved by visit order, so any difference in
 * split choice, child order or arena offsets changes hit ids.  The steps of
 * the algorithm therefore follow the reference (binning rend.c:274-387,
 * partition rend.c:389-414, leaf embedding rend.c:416-452, horizontal then
 * vertical splitting rend.c:491-656, octant ordering rend.c:658-687); the
 * float expressions keep its operation order and are compiled with
 * -ffp-contract=off.  Where the reference's behaviour is undefined (no split
 * with a finite cost leaves its split record uninitialised) this builder
 * falls back to a median split with exact bounds and reports a hazard bit.
 */
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"

#define BINS  16  /* INTERVAL_CNT, rend.c:27 */

struct split {
	float        cost;
	unsigned     axis;
	unsigned     pos;
	struct vec3  lmin, lmax, rmin, rmax;
	struct vec3  invd;
};

#define F32(v, i)  (((float *)&(v))[i])
#define U32(v, i)  (((unsigned int *)&(v))[i])

/* x86 CVTTSS2SI: truncation, with NaN and out-of-range mapped to INT_MIN. */
static int cvtt(float f)
{
	if (f >= -2147483648.0f && f < 2147483648.0f)
		return (int)f;
	return INT_MIN;
}

static unsigned bin_of(float c)
{
	int b = cvtt(c);
	b = b > 0 ? b : 0;
	b = b < BINS - 1 ? b : BINS - 1;
	return (unsigned)b;
}

static float half_area(struct vec3 mi, struct vec3 ma)
{
	struct vec3 d = rv_sub(ma, mi);
	return d.x * d.y + d.y * d.z + d.z * d.x;
}

static void best_split(struct split *best, unsigned start, unsigned cnt,
                       struct vec3 nmin, struct vec3 nmax, struct vec3 minext,
                       const struct aabb *aabbs, const unsigned *imap)
{
	struct vec3 bmin[3][BINS], bmax[3][BINS];
	unsigned bcnt[3][BINS];
	const struct vec3 big = {FLT_MAX, FLT_MAX, FLT_MAX};
	const struct vec3 nbig = {-FLT_MAX, -FLT_MAX, -FLT_MAX};

	for (unsigned a = 0; a < 3; a++)
		for (unsigned i = 0; i < BINS; i++) {
			bmin[a][i] = big;
			bmax[a][i] = nbig;
			bcnt[a][i] = 0;
		}

	best->invd.x = (float)BINS / (nmax.x - nmin.x);
	best->invd.y = (float)BINS / (nmax.y - nmin.y);
	best->invd.z = (float)BINS / (nmax.z - nmin.z);

	for (unsigned i = 0; i < cnt; i++) {
		const struct aabb *b = &aabbs[imap[start + i]];
		struct vec3 c = rv_mul(rv_sub(rv_scale(rv_add(b->min, b->max),
		  0.5f), nmin), best->invd);
		unsigned k[3] = {bin_of(c.x), bin_of(c.y), bin_of(c.z)};
		for (unsigned a = 0; a < 3; a++) {
			bcnt[a][k[a]]++;
			bmin[a][k[a]] = rv_min(bmin[a][k[a]], b->min);
			bmax[a][k[a]] = rv_max(bmax[a][k[a]], b->max);
		}
	}

	best->cost = FLT_MAX;

	for (unsigned a = 0; a < 3; a++) {
		if (rv_c(nmax, a) - rv_c(nmin, a) < rv_c(minext, a))
			continue;

		struct vec3 lmin[BINS - 1], lmax[BINS - 1];
		struct vec3 rmin[BINS - 1], rmax[BINS - 1];
		float lcost[BINS - 1], rcost[BINS - 1];
		struct vec3 lmi = big, lma = nbig, rmi = big, rma = nbig;
		unsigned lc = 0, rc = 0;

		for (unsigned i = 0; i < BINS - 1; i++) {
			unsigned r = BINS - 1 - i;
			lmi = rv_min(lmi, bmin[a][i]);
			lma = rv_max(lma, bmax[a][i]);
			lmin[i] = lmi;
			lmax[i] = lma;
			lc += bcnt[a][i];
			lcost[i] = lc > 0 ?
			  (float)lc * half_area(lmi, lma) / BRANCH_MAX : FLT_MAX;

			rmi = rv_min(rmi, bmin[a][r]);
			rma = rv_max(rma, bmax[a][r]);
			rmin[r - 1] = rmi;
			rmax[r - 1] = rma;
			rc += bcnt[a][r];
			rcost[r - 1] = rc > 0 ?
			  (float)rc * half_area(rmi, rma) / BRANCH_MAX : FLT_MAX;
		}

		for (unsigned i = 0; i < BINS - 1; i++) {
			float c = lcost[i] + rcost[i];
			if (c < best->cost) {
				best->cost = c;
				best->axis = a;
				best->pos = i;
				best->lmin = lmin[i];
				best->lmax = lmax[i];
				best->rmin = rmin[i];
				best->rmax = rmax[i];
			}
		}
	}
}

/* In-place partition of imap[start, start+cnt) by bin <= pos; returns the
 * left count.  Right-side elements are swapped in from the end. */
static unsigned partition(unsigned start, unsigned cnt, struct vec3 nmin,
                          const struct aabb *aabbs, unsigned *imap,
                          const struct split *s)
{
	unsigned l = start, r = start + cnt;
	float invd = rv_c(s->invd, s->axis);
	float lo = rv_c(nmin, s->axis);
	for (unsigned i = 0; i < cnt; i++) {
		unsigned id = imap[l];
		const struct aabb *b = &aabbs[id];
		float c = ((rv_c(b->min, s->axis) + rv_c(b->max, s->axis)) * 0.5f
		  - lo) * invd;
		if (bin_of(c) <= s->pos) {
			l++;
		} else {
			imap[l] = imap[--r];
			imap[r] = id;
		}
	}
	return l - start;
}

/* Median split with exact bounds, used only where the reference is
 * undefined (no finite SAH cost). */
static unsigned median_split(struct split *s, unsigned start, unsigned cnt,
                             const struct aabb *aabbs, const unsigned *imap)
{
	unsigned lcnt = cnt / 2;
	struct vec3 big = {FLT_MAX, FLT_MAX, FLT_MAX};
	struct vec3 nbig = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	s->lmin = s->rmin = big;
	s->lmax = s->rmax = nbig;
	for (unsigned i = 0; i < cnt; i++) {
		const struct aabb *b = &aabbs[imap[start + i]];
		if (i < lcnt) {
			s->lmin = rv_min(s->lmin, b->min);
			s->lmax = rv_max(s->lmax, b->max);
		} else {
			s->rmin = rv_min(s->rmin, b->min);
			s->rmax = rv_max(s->rmax, b->max);
		}
	}
	return lcnt;
}

static unsigned put_leaf4(unsigned char *base, unsigned ofs, unsigned start,
                          unsigned cnt, const unsigned *imap,
                          const struct rtri *tris)
{
	struct leaf4 *l = (struct leaf4 *)(base + ofs);
	unsigned k = start;
	for (unsigned i = 0; i < 4; i++) {
		const struct rtri *t = &tris[imap[k]];
		struct vec3 e0 = rv_sub(t->v1, t->v0);
		struct vec3 e1 = rv_sub(t->v2, t->v0);
		F32(l->v0x4, i) = t->v0.x;
		F32(l->v0y4, i) = t->v0.y;
		F32(l->v0z4, i) = t->v0.z;
		F32(l->e0x4, i) = e0.x;
		F32(l->e0y4, i) = e0.y;
		F32(l->e0z4, i) = e0.z;
		F32(l->e1x4, i) = e1.x;
		F32(l->e1y4, i) = e1.y;
		F32(l->e1z4, i) = e1.z;
		U32(l->id4, i) = imap[k];
		if (i + 1 < cnt)  /* lanes past cnt repeat the last triangle */
			k++;
	}
	l->tricnt = cnt;
	return (unsigned)sizeof(*l);
}

struct did {
	float          dist;
	unsigned char  id;
};

/* Top-down merge sort, descending by dist, taking the left run's element
 * whenever !(left < right).  This is the order glibc 2.35's qsort (merge
 * sort for small arrays) gives with the reference's comparator
 * (rend.c:97-101), which never returns 0. */
static void msort_desc(struct did *v, unsigned n, struct did *tmp)
{
	if (n <= 1)
		return;
	unsigned n1 = n / 2, n2 = n - n1;
	struct did *a = v, *b = v + n1;
	msort_desc(a, n1, tmp);
	msort_desc(b, n2, tmp);
	unsigned k = 0;
	while (n1 > 0 && n2 > 0) {
		if (!(a->dist < b->dist)) {
			tmp[k++] = *a++;
			n1--;
		} else {
			tmp[k++] = *b++;
			n2--;
		}
	}
	while (n1 > 0) {
		tmp[k++] = *a++;
		n1--;
	}
	memcpy(v, tmp, k * sizeof(*v));
}

static void set_leaf(struct bnode8 *n, unsigned j, unsigned char *base,
                     unsigned *ofs, unsigned start, unsigned cnt,
                     const unsigned *imap, const struct rtri *tris,
                     struct bvh8_stats *st)
{
	if (tris) {
		U32(n->children8, j) = NODE_LEAF | *ofs;
		*ofs += put_leaf4(base, *ofs, start, cnt, imap, tris);
	} else {
		U32(n->children8, j) = NODE_LEAF | imap[start];
	}
	U32(n->perm8, j) = 0;
	st->leaves++;
}

static void set_child_box(struct bnode8 *n, unsigned j, struct vec3 mi,
                          struct vec3 ma)
{
	F32(n->minx8, j) = mi.x;
	F32(n->miny8, j) = mi.y;
	F32(n->minz8, j) = mi.z;
	F32(n->maxx8, j) = ma.x;
	F32(n->maxy8, j) = ma.y;
	F32(n->maxz8, j) = ma.z;
}

static struct vec3 child_min(const struct bnode8 *n, unsigned j)
{
	struct vec3 v = {F32(n->minx8, j), F32(n->miny8, j), F32(n->minz8, j)};
	return v;
}

static struct vec3 child_max(const struct bnode8 *n, unsigned j)
{
	struct vec3 v = {F32(n->maxx8, j), F32(n->maxy8, j), F32(n->maxz8, j)};
	return v;
}

unsigned int bvh8_build(struct bnode8 *nodes, const struct aabb *aabbs,
                        unsigned int *imap, const struct rtri *tris,
                        unsigned int pcnt, struct vec3 rootmin,
                        struct vec3 rootmax, struct bvh8_stats *st)
{
	unsigned char *base = (unsigned char *)nodes;
	unsigned stack[64];
	unsigned sp = 0;
	unsigned leafmax = tris ? BLAS_LEAF_MAX : TLAS_LEAF_MAX;
	struct bvh8_stats local;
	if (!st)
		st = &local;
	memset(st, 0, sizeof(*st));

	struct bnode8 *n = nodes;
	unsigned ofs = (unsigned)sizeof(*n);
	unsigned childcnt = 1;
	unsigned ncnt = 1;

	memset(n, 0, sizeof(*n));
	set_child_box(n, 0, rootmin, rootmax);
	U32(n->children8, 0) = pcnt;  /* count; perm8 holds the start index */

	struct vec3 minext = rv_scale(rv_sub(rootmax, rootmin), 0.00000001f);

	for (;;) {
		/* Horizontal phase: split the widest splittable child until the
		 * node has 8 children. */
		while (childcnt < BRANCH_MAX) {
			int j = -1;
			float bcost = 0.0f;
			struct vec3 bmin = {0, 0, 0}, bmax = {0, 0, 0};
			for (unsigned i = 0; i < childcnt; i++) {
				if (U32(n->children8, i) & NODE_LEAF)
					continue;
				struct vec3 mi = child_min(n, i), ma = child_max(n, i);
				float c = half_area(mi, ma);
				if (c > bcost) {
					bcost = c;
					j = (int)i;
					bmin = mi;
					bmax = ma;
				}
			}
			if (j < 0)
				break;

			unsigned cnt = U32(n->children8, j);
			unsigned start = U32(n->perm8, j);
			if (cnt <= leafmax) {
				set_leaf(n, (unsigned)j, base, &ofs, start, cnt,
				  imap, tris, st);
				continue;
			}

			struct split s;
			unsigned lcnt;
			best_split(&s, start, cnt, bmin, bmax, minext, aabbs, imap);
			if (s.cost == FLT_MAX) {
				st->hazards |= BVH_HAZ_NOSPLIT;
				lcnt = median_split(&s, start, cnt, aabbs, imap);
			} else {
				lcnt = partition(start, cnt, bmin, aabbs, imap, &s);
			}
			if (lcnt == 0 || lcnt == cnt) {
				/* rend.c:547-551: force a 1 : n-1 split, bounds stay. */
				st->hazards |= BVH_HAZ_HDEGEN;
				lcnt = lcnt == 0 ? 1 : cnt - 1;
			}

			U32(n->children8, childcnt) = cnt - lcnt;
			U32(n->perm8, childcnt) = start + lcnt;
			set_child_box(n, childcnt, s.rmin, s.rmax);
			childcnt++;

			U32(n->children8, j) = lcnt;
			set_child_box(n, (unsigned)j, s.lmin, s.lmax);
		}

		/* Vertical phase: every child still holding a range becomes a
		 * leaf or a new two-child node scheduled for splitting. */
		for (unsigned j = 0; j < childcnt; j++) {
			unsigned cnt = U32(n->children8, j);
			if (cnt & NODE_LEAF)
				continue;
			unsigned start = U32(n->perm8, j);
			if (cnt <= leafmax) {
				set_leaf(n, j, base, &ofs, start, cnt, imap, tris, st);
				continue;
			}

			struct vec3 mi = child_min(n, j), ma = child_max(n, j);
			struct split s;
			unsigned lcnt;
			best_split(&s, start, cnt, mi, ma, minext, aabbs, imap);
			if (s.cost == FLT_MAX) {
				st->hazards |= BVH_HAZ_NOSPLIT;
				lcnt = median_split(&s, start, cnt, aabbs, imap);
			} else {
				lcnt = partition(start, cnt, mi, aabbs, imap, &s);
			}
			if (lcnt == 0 || lcnt == cnt) {
				/* The reference leaves the count in children8 here
				 * (rend.c:615-620), which traversal then reads as a
				 * node offset.  Split at the median instead. */
				st->hazards |= BVH_HAZ_VDEGEN;
				lcnt = median_split(&s, start, cnt, aabbs, imap);
			}

			struct bnode8 *c = (struct bnode8 *)(base + ofs);
			memset(c, 0, sizeof(*c));
			U32(c->children8, 0) = lcnt;
			U32(c->perm8, 0) = start;
			set_child_box(c, 0, s.lmin, s.lmax);
			U32(c->children8, 1) = cnt - lcnt;
			U32(c->perm8, 1) = start + lcnt;
			set_child_box(c, 1, s.rmin, s.rmax);

			U32(n->children8, j) = ofs;
			U32(n->perm8, j) = 0;

			if (sp >= 64) {
				st->hazards |= BVH_HAZ_STACK;
				st->used = ofs;
				st->nodes = ncnt;
				return ncnt;
			}
			stack[sp++] = ofs;
			ofs += (unsigned)sizeof(*c);
			ncnt++;
		}

		/* Per-octant traversal order: children sorted by descending
		 * distance of their near corner along the octant direction. */
		for (unsigned o = 0; o < BRANCH_MAX; o++) {
			struct vec3 dir = {(o & 1) ? 1.0f : -1.0f,
			  (o & 2) ? 1.0f : -1.0f, (o & 4) ? 1.0f : -1.0f};
			struct did d[BRANCH_MAX], tmp[BRANCH_MAX];
			for (unsigned i = 0; i < BRANCH_MAX; i++) {
				d[i].id = (unsigned char)i;
				if (i < childcnt) {
					struct vec3 corner = {
					  (o & 1) ? F32(n->minx8, i) : F32(n->maxx8, i),
					  (o & 2) ? F32(n->miny8, i) : F32(n->maxy8, i),
					  (o & 4) ? F32(n->minz8, i) : F32(n->maxz8, i)};
					d[i].dist = rv_dot(dir, corner);
				} else {
					d[i].dist = FLT_MAX;
				}
			}
			msort_desc(d, BRANCH_MAX, tmp);
			for (unsigned i = 0; i < BRANCH_MAX; i++)
				U32(n->perm8, i) |= (unsigned)d[i].id << (o * 3);
		}

		for (; childcnt < BRANCH_MAX; childcnt++) {
			struct vec3 big = {FLT_MAX, FLT_MAX, FLT_MAX};
			struct vec3 nbig = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
			set_child_box(n, childcnt, big, nbig);
		}

		if (sp == 0)
			break;
		n = (struct bnode8 *)(base + stack[--sp]);
		childcnt = 2;
	}

	st->nodes = ncnt;
	st->used = ofs;
	return ncnt;
}

long long bvh8_extent(const struct bnode8 *root, int is_blas, size_t limit)
{
	const unsigned char *base = (const unsigned char *)root;
	unsigned stack[1024];
	unsigned sp = 0;
	long long end = (long long)sizeof(struct bnode8);
	if (limit < sizeof(struct bnode8))
		return -1;
	stack[sp++] = 0;
	while (sp) {
		unsigned ofs = stack[--sp];
		const struct bnode8 *n = (const struct bnode8 *)(base + ofs);
		for (unsigned i = 0; i < BRANCH_MAX; i++) {
			unsigned c = U32(n->children8, i);
			if (c == 0)
				continue;  /* empty slot */
			if (c & NODE_LEAF) {
				if (!is_blas)
					continue;
				unsigned lo = c & ~NODE_LEAF;
				if ((lo & 31u) || lo + sizeof(struct leaf4) > limit)
					return -1;
				if ((long long)(lo + sizeof(struct leaf4)) > end)
					end = (long long)(lo + sizeof(struct leaf4));
			} else {
				if (c <= ofs || (c & 31u) ||
				    c + sizeof(struct bnode8) > limit || sp >= 1024)
					return -1;
				if ((long long)(c + sizeof(struct bnode8)) > end)
					end = (long long)(c + sizeof(struct bnode8));
				stack[sp++] = c;
			}
		}
	}
	return end;
}

/* Device repack.  Nodes are renumbered in breadth-first order at 256-byte
 * strides (one node = two 128-byte lines), leaves follow at 192-byte strides
 * in the order they are reached.  Child words keep their meaning with the
 * new offsets; empty slots keep the reference's zero word.  perm8 becomes
 * the per-octant inverse permutation (dlayout.h). */
long long bvh8_repack_size(const struct bnode8 *root, int is_blas,
                           size_t limit, unsigned *nnodes, unsigned *nleaves)
{
	long long ext = bvh8_extent(root, is_blas, limit);
	if (ext < 0)
		return -1;
	const unsigned char *base = (const unsigned char *)root;
	unsigned nn = 0, nl = 0;
	/* Tree walk without a queue: visit nodes by scanning the extent. */
	unsigned *q = malloc(((size_t)ext / sizeof(struct bnode8) + 1) * sizeof(*q));
	if (!q)
		return -1;
	size_t qh = 0, qt = 0;
	q[qt++] = 0;
	while (qh < qt) {
		const struct bnode8 *n = (const struct bnode8 *)(base + q[qh++]);
		nn++;
		for (unsigned i = 0; i < BRANCH_MAX; i++) {
			unsigned c = U32(n->children8, i);
			if (c == 0)
				continue;
			if (c & NODE_LEAF)
				nl += is_blas ? 1u : 0u;
			else
				q[qt++] = c;
		}
	}
	free(q);
	if (nnodes)
		*nnodes = nn;
	if (nleaves)
		*nleaves = nl;
	return (long long)nn * (long long)sizeof(struct bnode8) +
	  (long long)nl * (long long)sizeof(struct leaf4);
}

long long bvh8_repack(void *dst, size_t cap, const struct bnode8 *root,
                      int is_blas, size_t limit)
{
	unsigned nn, nl;
	long long need = bvh8_repack_size(root, is_blas, limit, &nn, &nl);
	if (need < 0 || (size_t)need > cap)
		return -1;
	const unsigned char *src = (const unsigned char *)root;
	unsigned char *out = (unsigned char *)dst;
	unsigned *q = malloc(((size_t)nn + 1) * sizeof(*q));
	if (!q)
		return -1;
	size_t qh = 0, qt = 0;
	unsigned nextnode = 1, nextleaf = 0;
	const size_t leafbase = (size_t)nn * sizeof(struct bnode8);
	q[qt++] = 0;
	while (qh < qt) {
		size_t k = qh;
		const struct bnode8 *n = (const struct bnode8 *)(src + q[qh++]);
		struct bnode8 *d = (struct bnode8 *)(out + k * sizeof(struct bnode8));
		memcpy(d, n, 6 * sizeof(__m256));
		for (unsigned i = 0; i < BRANCH_MAX; i++) {
			unsigned c = U32(n->children8, i), nc = c;
			if (c == 0) {
				nc = 0;
			} else if (c & NODE_LEAF) {
				if (is_blas) {
					size_t lo = leafbase + (size_t)nextleaf++ *
					  sizeof(struct leaf4);
					memcpy(out + lo, src + (c & ~NODE_LEAF),
					  sizeof(struct leaf4));
					nc = NODE_LEAF | (unsigned)lo;
				}
			} else {
				nc = (unsigned)((size_t)nextnode++ * sizeof(struct bnode8));
				q[qt++] = c;
			}
			U32(d->children8, i) = nc;
		}
		for (unsigned o = 0; o < BRANCH_MAX; o++) {
			unsigned inv = 0;
			for (unsigned r = 0; r < BRANCH_MAX; r++) {
				unsigned c = (U32(n->perm8, r) >> (3 * o)) & 7u;
				inv |= r << (3 * c);
			}
			U32(d->perm8, o) = inv;
		}
	}
	free(q);
	return need;
}
