/*
 * Procedural scenes and the scene -> rdata flattening (rend_scene.h).
 *
 * Flattening keeps the reference's bridge semantics: triangles and vertex
 * normals of every mesh are concatenated with a per-triangle material id
 * (cpy_rdata, main.c:53-112); instances store the first 12 floats of the
 * inverse world matrix and the world AABB of the 8 transformed corners of
 * the BLAS root bounds (upd_rinsts, main.c:117-157); the world matrix is
 * T * R(q) * S (combine_tfcomp, scene.c:176-185); the camera constants are
 * those of upd_rcam (main.c:159-195).  The matrix inverse is the cofactor
 * expansion with a double-precision reciprocal of the determinant, term for
 * term as mat4_inv (mat4.c:299-362), so globinv bits match the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_scene.h"
#include "rend_hip.h"
#include "librend.h"

#define PI_D  3.14159265  /* util.h:9, a double constant */

/* ---- value noise ---------------------------------------------------------- */

static uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

static float lattice(int ix, int iy, int iz, uint32_t seed)
{
	uint32_t h = mix32(seed ^ mix32((uint32_t)ix ^ mix32((uint32_t)iy ^
	  mix32((uint32_t)iz + 0x51ed270bu))));
	return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

static float fade(float t)
{
	return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

static float vnoise(float x, float y, float z, uint32_t seed)
{
	float fx = floorf(x), fy = floorf(y), fz = floorf(z);
	int ix = (int)fx, iy = (int)fy, iz = (int)fz;
	float tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
	float c[2][2][2];
	for (int k = 0; k < 2; k++)
		for (int j = 0; j < 2; j++)
			for (int i = 0; i < 2; i++)
				c[k][j][i] = lattice(ix + i, iy + j, iz + k, seed);
	float a[2][2];
	for (int k = 0; k < 2; k++)
		for (int j = 0; j < 2; j++)
			a[k][j] = c[k][j][0] + (c[k][j][1] - c[k][j][0]) * tx;
	float b[2];
	for (int k = 0; k < 2; k++)
		b[k] = a[k][0] + (a[k][1] - a[k][0]) * ty;
	return b[0] + (b[1] - b[0]) * tz;
}

/* 4 octaves, lacunarity 2, gain 0.5, normalised to [-1, 1]. */
float rs_noise(float x, float y, float z, unsigned int seed)
{
	float sum = 0.0f, amp = 1.0f, norm = 0.0f;
	for (unsigned o = 0; o < 4; o++) {
		sum += amp * vnoise(x, y, z, seed + o * 0x9e3779b9u);
		norm += amp;
		amp *= 0.5f;
		x *= 2.0f;
		y *= 2.0f;
		z *= 2.0f;
	}
	return sum / norm;
}

/* ---- scene builder ------------------------------------------------------ */

struct rs_mesh {
	unsigned   nv, ntri;
	float     *v;     /* 3 per vertex */
	float     *n;     /* 3 per vertex */
	unsigned  *idx;   /* 3 per triangle */
	unsigned  *mtl;   /* 1 per triangle */
};

struct rs_inst {
	unsigned  mesh, flags;
	float     glob[16];
	int       node;   /* scene-graph node it follows, or -1 */
};

/* Scene graph of imported scenes (scene.h:47-130): nodes with transform
 * components, local and world matrices, keyframe tracks and samplers over
 * one block of animation data. */
struct rs_node {
	int       parent;
	int       mesh, cam;
	unsigned  flags;
	float     trans[3], rot[4], scale[3];
	float     loc[16], glob[16];
};

struct rs_track {
	unsigned  sampler, node, key;
	int       tgt;    /* RS_TGT_* */
};

struct rs_sampler {
	unsigned  kcnt, kofs, dofs;  /* key count, key and value offsets (floats) */
	int       interp;            /* RS_INTERP_* */
};

struct rs_cam {
	float     vfov, focdist, focangle;  /* degrees */
	int       node;
};

struct rs_scene {
	struct rmtl    *mtls;
	unsigned        nmtl, capmtl;
	struct rs_mesh *meshes;
	unsigned        nmesh, capmesh;
	struct rs_inst *insts;
	unsigned        ninst, capinst;
	struct rs_mesh  cur;
	unsigned        capv, capt;
	int             open;
	float           eye[3], target[3], vfov, focdist, focangle;
	float           bg[3];
	/* scene graph (imported scenes) */
	struct rs_node   *nodes;
	unsigned          nnode, capnode;
	struct rs_track  *tracks;
	unsigned          ntrack, captrack;
	struct rs_sampler *samplers;
	unsigned          nsampler, capsampler;
	float            *animdata;
	unsigned          nanim, capanim;   /* floats */
	struct rs_cam    *cams;
	unsigned          ncam, capcam;
	unsigned          dirty;            /* RS_DIRTY_* */
};

#define GROW(p, n, cap, extra) do { \
	if ((n) + (extra) > (cap)) { \
		size_t nc_ = (cap) ? (size_t)(cap) * 2 : 64; \
		while (nc_ < (size_t)(n) + (extra)) nc_ *= 2; \
		void *np_ = realloc((p), nc_ * sizeof(*(p))); \
		if (!np_) return -1; \
		(p) = np_; (cap) = (unsigned)nc_; \
	} } while (0)

struct rs_scene *rs_new(void)
{
	struct rs_scene *s = calloc(1, sizeof(*s));
	if (!s)
		return NULL;
	s->eye[2] = 1.0f;
	s->vfov = 45.0f;
	s->focdist = 1.0f;
	return s;
}

static void mesh_free(struct rs_mesh *m)
{
	free(m->v);
	free(m->n);
	free(m->idx);
	free(m->mtl);
	memset(m, 0, sizeof(*m));
}

void rs_free(struct rs_scene *s)
{
	if (!s)
		return;
	for (unsigned i = 0; i < s->nmesh; i++)
		mesh_free(&s->meshes[i]);
	mesh_free(&s->cur);
	free(s->meshes);
	free(s->mtls);
	free(s->insts);
	free(s->nodes);
	free(s->tracks);
	free(s->samplers);
	free(s->animdata);
	free(s->cams);
	free(s);
}

int rs_material(struct rs_scene *s, float r, float g, float b,
                unsigned int flags)
{
	GROW(s->mtls, s->nmtl, s->capmtl, 1);
	struct rmtl *m = &s->mtls[s->nmtl];
	memset(m, 0, sizeof(*m));
	m->col.x = r;
	m->col.y = g;
	m->col.z = b;
	m->roughness = 1.0f;
	m->ior = 1.5f;
	m->flags = flags;
	return (int)s->nmtl++;
}

int rs_mesh_begin(struct rs_scene *s)
{
	if (s->open)
		return -1;
	memset(&s->cur, 0, sizeof(s->cur));
	s->capv = s->capt = 0;
	s->open = 1;
	return 0;
}

static int add_vert(struct rs_scene *s, float x, float y, float z)
{
	struct rs_mesh *m = &s->cur;
	if (m->nv + 1 > s->capv) {
		unsigned nc = s->capv ? s->capv * 2 : 1024;
		float *nv = realloc(m->v, (size_t)nc * 3 * sizeof(float));
		if (!nv)
			return -1;
		m->v = nv;
		float *nn = realloc(m->n, (size_t)nc * 3 * sizeof(float));
		if (!nn)
			return -1;
		m->n = nn;
		s->capv = nc;
	}
	m->v[3 * m->nv + 0] = x;
	m->v[3 * m->nv + 1] = y;
	m->v[3 * m->nv + 2] = z;
	m->n[3 * m->nv + 0] = m->n[3 * m->nv + 1] = m->n[3 * m->nv + 2] = 0.0f;
	return (int)m->nv++;
}

static int add_tri(struct rs_scene *s, unsigned a, unsigned b, unsigned c,
                   unsigned mtl)
{
	struct rs_mesh *m = &s->cur;
	if (m->ntri + 1 > s->capt) {
		unsigned nc = s->capt ? s->capt * 2 : 1024;
		unsigned *ni = realloc(m->idx, (size_t)nc * 3 * sizeof(unsigned));
		if (!ni)
			return -1;
		m->idx = ni;
		unsigned *nm = realloc(m->mtl, (size_t)nc * sizeof(unsigned));
		if (!nm)
			return -1;
		m->mtl = nm;
		s->capt = nc;
	}
	m->idx[3 * m->ntri + 0] = a;
	m->idx[3 * m->ntri + 1] = b;
	m->idx[3 * m->ntri + 2] = c;
	m->mtl[m->ntri] = mtl;
	m->ntri++;
	return 0;
}

/* Area-weighted vertex normals over the triangles added since vertex v0. */
static void smooth_normals(struct rs_scene *s, unsigned v0, unsigned t0)
{
	struct rs_mesh *m = &s->cur;
	for (unsigned t = t0; t < m->ntri; t++) {
		unsigned *ix = &m->idx[3 * t];
		const float *a = &m->v[3 * ix[0]], *b = &m->v[3 * ix[1]],
		  *c = &m->v[3 * ix[2]];
		float e0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
		float e1[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
		float n[3] = {e0[1] * e1[2] - e0[2] * e1[1],
		  e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
		for (int k = 0; k < 3; k++)
			for (int j = 0; j < 3; j++)
				m->n[3 * ix[k] + j] += n[j];
	}
	for (unsigned v = v0; v < m->nv; v++) {
		float *n = &m->n[3 * v];
		float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
		if (l > 0.0f) {
			n[0] /= l;
			n[1] /= l;
			n[2] /= l;
		} else {
			n[1] = 1.0f;
		}
	}
}

int rs_mesh_data(struct rs_scene *s, const float *v, unsigned int nv, const float *n,
                 const unsigned int *idx, unsigned int ntri, const unsigned int *mtls,
                 unsigned int mtl)
{
	if (!s->open || !v || !idx)
		return -1;
	for (unsigned t = 0; t < 3 * ntri; t++)
		if (idx[t] >= nv)
			return -1;
	for (unsigned t = 0; mtls && t < ntri; t++)
		if (mtls[t] >= s->nmtl)
			return -1;
	if (!mtls && ntri && mtl >= s->nmtl)
		return -1;
	const unsigned v0 = s->cur.nv, t0 = s->cur.ntri;
	for (unsigned i = 0; i < nv; i++) {
		if (add_vert(s, v[3 * i], v[3 * i + 1], v[3 * i + 2]) < 0)
			return -1;
		if (n)
			memcpy(&s->cur.n[3 * (v0 + i)], &n[3 * i], 3 * sizeof(float));
	}
	for (unsigned t = 0; t < ntri; t++)
		if (add_tri(s, v0 + idx[3 * t], v0 + idx[3 * t + 1], v0 + idx[3 * t + 2],
		  mtls ? mtls[t] : mtl) < 0)
			return -1;
	if (!n)
		smooth_normals(s, v0, t0);
	return 0;
}

unsigned int rs_mtlcount(const struct rs_scene *s)
{
	return s->nmtl;
}

int rs_mesh_end(struct rs_scene *s)
{
	if (!s->open)
		return -1;
	GROW(s->meshes, s->nmesh, s->capmesh, 1);
	s->meshes[s->nmesh] = s->cur;
	memset(&s->cur, 0, sizeof(s->cur));
	s->open = 0;
	return (int)s->nmesh++;
}

/* Welding map for subdivision vertices keyed by exact position bits. */
struct wmap {
	uint32_t  *keys;  /* 3 per slot */
	int       *val;
	size_t     cap;
};

static int wmap_init(struct wmap *w, size_t n)
{
	w->cap = 1;
	while (w->cap < n * 2)
		w->cap <<= 1;
	w->keys = malloc(w->cap * 3 * sizeof(uint32_t));
	w->val = malloc(w->cap * sizeof(int));
	if (!w->keys || !w->val)
		return -1;
	for (size_t i = 0; i < w->cap; i++)
		w->val[i] = -1;
	return 0;
}

static void wmap_free(struct wmap *w)
{
	free(w->keys);
	free(w->val);
}

static int wmap_get(struct rs_scene *s, struct wmap *w, const float p[3])
{
	uint32_t k[3];
	memcpy(k, p, sizeof(k));
	size_t h = mix32(k[0] ^ mix32(k[1] ^ mix32(k[2]))) & (w->cap - 1);
	for (;;) {
		if (w->val[h] < 0) {
			int id = add_vert(s, p[0], p[1], p[2]);
			if (id < 0)
				return -1;
			memcpy(&w->keys[3 * h], k, sizeof(k));
			w->val[h] = id;
			return id;
		}
		if (!memcmp(&w->keys[3 * h], k, sizeof(k)))
			return w->val[h];
		h = (h + 1) & (w->cap - 1);
	}
}

static void unit3(float p[3])
{
	float l = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
	p[0] /= l;
	p[1] /= l;
	p[2] /= l;
}

static int ico_tri(struct rs_scene *s, struct wmap *w, const float *a,
                   const float *b, const float *c, unsigned depth,
                   unsigned mtl)
{
	if (depth == 0) {
		int ia = wmap_get(s, w, a), ib = wmap_get(s, w, b),
		  ic = wmap_get(s, w, c);
		if (ia < 0 || ib < 0 || ic < 0)
			return -1;
		return add_tri(s, (unsigned)ia, (unsigned)ib, (unsigned)ic, mtl);
	}
	float ab[3], bc[3], ca[3];
	for (int k = 0; k < 3; k++) {
		ab[k] = (a[k] + b[k]) * 0.5f;
		bc[k] = (b[k] + c[k]) * 0.5f;
		ca[k] = (c[k] + a[k]) * 0.5f;
	}
	unit3(ab);
	unit3(bc);
	unit3(ca);
	if (ico_tri(s, w, a, ab, ca, depth - 1, mtl) ||
	    ico_tri(s, w, ab, b, bc, depth - 1, mtl) ||
	    ico_tri(s, w, ca, bc, c, depth - 1, mtl) ||
	    ico_tri(s, w, ab, bc, ca, depth - 1, mtl))
		return -1;
	return 0;
}

int rs_icosphere(struct rs_scene *s, unsigned int subdiv, float cx, float cy,
                 float cz, float r, float amp, float freq, unsigned int seed,
                 unsigned int mtl)
{
	if (!s->open || subdiv > 9)
		return -1;
	const float t = 1.6180339887f;
	float v[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
	  {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
	  {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
	static const unsigned f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7},
	  {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4}, {11, 10, 2},
	  {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
	  {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
	for (int i = 0; i < 12; i++)
		unit3(v[i]);
	unsigned v0 = s->cur.nv, t0 = s->cur.ntri;
	size_t nv = 10u * ((size_t)1 << (2 * subdiv)) + 2;
	struct wmap w;
	if (wmap_init(&w, nv + 16))
		return -1;
	for (int i = 0; i < 20; i++)
		if (ico_tri(s, &w, v[f[i][0]], v[f[i][1]], v[f[i][2]], subdiv,
		  mtl)) {
			wmap_free(&w);
			return -1;
		}
	wmap_free(&w);
	for (unsigned i = v0; i < s->cur.nv; i++) {
		float *p = &s->cur.v[3 * i];
		float d = r * (1.0f + amp * rs_noise(freq * p[0], freq * p[1],
		  freq * p[2], seed));
		p[0] = cx + p[0] * d;
		p[1] = cy + p[1] * d;
		p[2] = cz + p[2] * d;
	}
	smooth_normals(s, v0, t0);
	return 0;
}

int rs_torus(struct rs_scene *s, unsigned int nu, unsigned int nv, float R,
             float r, float amp, float freq, unsigned int seed,
             unsigned int mtl)
{
	if (!s->open || nu < 3 || nv < 3)
		return -1;
	unsigned v0 = s->cur.nv, t0 = s->cur.ntri;
	for (unsigned i = 0; i < nu; i++) {
		double a = 2.0 * PI_D * i / nu;
		for (unsigned j = 0; j < nv; j++) {
			double b = 2.0 * PI_D * j / nv;
			float nx = (float)(cos(a) * cos(b)), ny = (float)sin(b),
			  nz = (float)(sin(a) * cos(b));
			float cxp = (float)(R * cos(a)), czp = (float)(R * sin(a));
			float px = cxp + r * nx, py = r * ny, pz = czp + r * nz;
			float d = r * amp * rs_noise(freq * px, freq * py, freq * pz,
			  seed);
			if (add_vert(s, px + d * nx, py + d * ny, pz + d * nz) < 0)
				return -1;
		}
	}
	for (unsigned i = 0; i < nu; i++)
		for (unsigned j = 0; j < nv; j++) {
			unsigned a = v0 + i * nv + j, b = v0 + ((i + 1) % nu) * nv + j;
			unsigned c = v0 + ((i + 1) % nu) * nv + (j + 1) % nv;
			unsigned d = v0 + i * nv + (j + 1) % nv;
			if (add_tri(s, a, d, c, mtl) || add_tri(s, a, c, b, mtl))
				return -1;
		}
	smooth_normals(s, v0, t0);
	return 0;
}

int rs_terrain(struct rs_scene *s, unsigned int nx, unsigned int nz,
               float x0, float x1, float z0, float z1, float amp, float freq,
               unsigned int seed, unsigned int mtl)
{
	if (!s->open || nx < 1 || nz < 1)
		return -1;
	unsigned v0 = s->cur.nv, t0 = s->cur.ntri;
	for (unsigned j = 0; j <= nz; j++) {
		float z = z0 + (z1 - z0) * (float)j / (float)nz;
		for (unsigned i = 0; i <= nx; i++) {
			float x = x0 + (x1 - x0) * (float)i / (float)nx;
			float y = amp * rs_noise(freq * x, 0.37f, freq * z, seed);
			if (add_vert(s, x, y, z) < 0)
				return -1;
		}
	}
	for (unsigned j = 0; j < nz; j++)
		for (unsigned i = 0; i < nx; i++) {
			unsigned a = v0 + j * (nx + 1) + i, b = a + 1;
			unsigned c = a + (nx + 1), d = c + 1;
			/* counter-clockwise seen from +y */
			if (add_tri(s, a, c, b, mtl) || add_tri(s, b, c, d, mtl))
				return -1;
		}
	smooth_normals(s, v0, t0);
	return 0;
}

int rs_quad(struct rs_scene *s, const float p0[3], const float p1[3],
            const float p2[3], const float p3[3], unsigned int mtl)
{
	if (!s->open)
		return -1;
	unsigned v0 = s->cur.nv, t0 = s->cur.ntri;
	int a = add_vert(s, p0[0], p0[1], p0[2]);
	int b = add_vert(s, p1[0], p1[1], p1[2]);
	int c = add_vert(s, p2[0], p2[1], p2[2]);
	int d = add_vert(s, p3[0], p3[1], p3[2]);
	if (a < 0 || b < 0 || c < 0 || d < 0)
		return -1;
	if (add_tri(s, (unsigned)a, (unsigned)b, (unsigned)c, mtl) ||
	    add_tri(s, (unsigned)a, (unsigned)c, (unsigned)d, mtl))
		return -1;
	smooth_normals(s, v0, t0);
	return 0;
}

/* ---- matrices (mat4.c semantics) ------------------------------------------ */

static void m4_mul(float d[16], const float a[16], const float b[16])
{
	float r[16];
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			r[4 * i + j] = a[4 * i + 0] * b[j] + a[4 * i + 1] * b[4 + j] +
			  a[4 * i + 2] * b[8 + j] + a[4 * i + 3] * b[12 + j];
	memcpy(d, r, sizeof(r));
}

static void m4_fromquat(float d[16], float x, float y, float z, float w)
{
	float xx = x * x, xy = x * y, xz = x * z, xw = x * w;
	float yy = y * y, yz = y * z, yw = y * w, zz = z * z, zw = z * w;
	d[0] = 1.0f - 2.0f * (yy + zz);
	d[1] = 2.0f * (xy - zw);
	d[2] = 2.0f * (xz + yw);
	d[4] = 2.0f * (xy + zw);
	d[5] = 1.0f - 2.0f * (xx + zz);
	d[6] = 2.0f * (yz - xw);
	d[8] = 2.0f * (xz - yw);
	d[9] = 2.0f * (yz + xw);
	d[10] = 1.0f - 2.0f * (xx + yy);
	d[3] = d[7] = d[11] = d[12] = d[13] = d[14] = 0.0f;
	d[15] = 1.0f;
}

/* Cofactor inverse (the public-domain Mesa/GLU gluInvertMatrix expansion,
 * which the reference's mat4_inv, mat4.c:297-362, credits and uses), with
 * the determinant's reciprocal in double as there, so instance inverses
 * match the reference's bit for bit; returns 1 when singular. */
static int m4_inv(float dst[16], const float m[16])
{
	float v[16];
	v[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
	  m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
	v[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
	  m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
	v[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
	  m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
	v[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
	  m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
	v[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
	  m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
	v[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
	  m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
	v[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
	  m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
	v[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
	  m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
	v[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
	  m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
	v[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
	  m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
	v[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
	  m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
	v[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
	  m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
	v[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
	  m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
	v[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
	  m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
	v[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
	  m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
	v[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
	  m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
	float det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
	if (det == 0)
		return 1;
	det = (float)(1.0 / det);
	for (int i = 0; i < 16; i++)
		dst[i] = v[i] * det;
	return 0;
}

int rs_instance(struct rs_scene *s, unsigned int mesh, const float t[3],
                const float q[4], const float scale[3], unsigned int flags)
{
	if (mesh >= s->nmesh || s->ninst >= 4096)
		return -1;
	GROW(s->insts, s->ninst, s->capinst, 1);
	struct rs_inst *in = &s->insts[s->ninst];
	float ms[16] = {0}, mr[16], mt[16] = {0};
	ms[0] = scale[0];
	ms[5] = scale[1];
	ms[10] = scale[2];
	ms[15] = 1.0f;
	m4_fromquat(mr, q[0], q[1], q[2], q[3]);
	mt[0] = mt[5] = mt[10] = mt[15] = 1.0f;
	mt[3] = t[0];
	mt[7] = t[1];
	mt[11] = t[2];
	m4_mul(in->glob, mr, ms);
	m4_mul(in->glob, mt, in->glob);
	in->mesh = mesh;
	in->flags = flags | 0x0001u;  /* MESH */
	in->node = -1;
	return (int)s->ninst++;
}

/* ---- scene graph and keyframe animation (scene.c:176-399) ----------------- */

/* Local matrix T * R(q) * S of transform components (combine_tfcomp,
 * scene.c:176-185). */
static void combine_trs(float d[16], const float t[3], const float q[4], const float sc[3])
{
	float ms[16] = {0}, mr[16], mt[16] = {0};
	ms[0] = sc[0];
	ms[5] = sc[1];
	ms[10] = sc[2];
	ms[15] = 1.0f;
	m4_fromquat(mr, q[0], q[1], q[2], q[3]);
	mt[0] = mt[5] = mt[10] = mt[15] = 1.0f;
	mt[3] = t[0];
	mt[7] = t[1];
	mt[11] = t[2];
	m4_mul(d, mr, ms);
	m4_mul(d, mt, d);
}

int rsi_node(struct rs_scene *s, int parent, int mesh, int cam, unsigned flags,
             const float t[3], const float q[4], const float sc[3])
{
	if (parent >= (int)s->nnode)
		return -1;
	GROW(s->nodes, s->nnode, s->capnode, 1);
	struct rs_node *n = &s->nodes[s->nnode];
	memset(n, 0, sizeof(*n));
	n->parent = parent;
	n->mesh = mesh;
	n->cam = cam;
	n->flags = flags;
	memcpy(n->trans, t, sizeof(n->trans));
	memcpy(n->rot, q, sizeof(n->rot));
	memcpy(n->scale, sc, sizeof(n->scale));
	combine_trs(n->loc, t, q, sc);
	memcpy(n->glob, n->loc, sizeof(n->glob));
	return (int)s->nnode++;
}

unsigned rsi_node_flags(struct rs_scene *s, unsigned node, unsigned set)
{
	if (node >= s->nnode)
		return 0;
	s->nodes[node].flags |= set;
	return s->nodes[node].flags;
}

int rsi_cam(struct rs_scene *s, float vfov, float focdist, float focangle, int node)
{
	GROW(s->cams, s->ncam, s->capcam, 1);
	struct rs_cam *c = &s->cams[s->ncam];
	c->vfov = vfov;
	c->focdist = focdist;
	c->focangle = focangle;
	c->node = node;
	return (int)s->ncam++;
}

int rsi_cam_node(struct rs_scene *s, unsigned cam, int node)
{
	if (cam >= s->ncam)
		return -1;
	s->cams[cam].node = node;
	return 0;
}

int rsi_mtl_set(struct rs_scene *s, unsigned id, float metallic, float roughness,
                unsigned flags, const float *col)
{
	if (id >= s->nmtl)
		return -1;
	struct rmtl *m = &s->mtls[id];
	m->metallic = metallic;
	m->roughness = roughness;
	m->flags = flags;
	if (col) {
		m->col.x = col[0];
		m->col.y = col[1];
		m->col.z = col[2];
	}
	return 0;
}

unsigned rsi_mtl_flags(const struct rs_scene *s, unsigned id)
{
	return id < s->nmtl ? s->mtls[id].flags : 0u;
}

/* Nothing added yet: no materials, meshes, instances, nodes or cameras. */
int rsi_empty(const struct rs_scene *s)
{
	return !s->nmtl && !s->nmesh && !s->ninst && !s->nnode && !s->ncam && !s->open;
}

/* Appends floats to the animation data; returns their offset or -1. */
long rsi_animdata(struct rs_scene *s, const float *v, unsigned n)
{
	GROW(s->animdata, s->nanim, s->capanim, n ? n : 1);
	memcpy(&s->animdata[s->nanim], v, (size_t)n * sizeof(float));
	long at = (long)s->nanim;
	s->nanim += n;
	return at;
}

int rsi_sampler(struct rs_scene *s, unsigned kcnt, unsigned kofs, unsigned dofs, int interp)
{
	GROW(s->samplers, s->nsampler, s->capsampler, 1);
	s->samplers[s->nsampler] = (struct rs_sampler){kcnt, kofs, dofs, interp};
	return (int)s->nsampler++;
}

/* A track marks its node as animated (scene_inittrack, scene.c:233-246). */
int rsi_track(struct rs_scene *s, unsigned sampler, unsigned node, int tgt)
{
	if (node >= s->nnode)
		return -1;
	GROW(s->tracks, s->ntrack, s->captrack, 1);
	s->tracks[s->ntrack] = (struct rs_track){sampler, node, 0, tgt};
	s->nodes[node].flags |= RS_DYNAMIC;
	return (int)s->ntrack++;
}

/* Instance of every mesh node, in node order (cpy_rdata, main.c:98-108). */
int rsi_instances_from_nodes(struct rs_scene *s)
{
	for (unsigned i = 0; i < s->nnode; i++) {
		struct rs_node *n = &s->nodes[i];
		if (!(n->flags & RS_MESH))
			continue;
		if (n->mesh < 0 || (unsigned)n->mesh >= s->nmesh || s->ninst >= 4096)
			return -1;
		GROW(s->insts, s->ninst, s->capinst, 1);
		struct rs_inst *in = &s->insts[s->ninst++];
		in->mesh = (unsigned)n->mesh;
		in->flags = n->flags;
		in->node = (int)i;
		memcpy(in->glob, n->glob, sizeof(in->glob));
	}
	return 0;
}

/* Key interval for time t, searched from the track's last key
 * (find_key, scene.c:248-254: a later time never moves it back). */
static unsigned find_key(const float *keys, unsigned last, unsigned cnt, float t)
{
	for (int i = (int)last; i < (int)cnt - 1; i++)
		if (t <= keys[i])
			return (unsigned)(i - 1 > 0 ? i - 1 : 0);
	return cnt - 2;
}

static void interp_linear(float *d, const float *a, const float *b, float t)
{
	for (int k = 0; k < 3; k++)
		d[k] = (1.0f - t) * a[k] + t * b[k];
}

/* Spherical interpolation of unit quaternions along the shorter arc
 * (spher_lerp, scene.c:263-278). */
static void interp_slerp(float *d, const float *a, const float *b, float t)
{
	float c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
	float ang = acosf(c);
	float isin = 1.0f / sinf(ang);
	float e = sinf(ang * (1.0f - t)) * isin;
	float f = (c / fabsf(c)) * sinf(ang * t) * isin;
	for (int k = 0; k < 4; k++)
		d[k] = e * a[k] + f * b[k];
}

/* Cubic Hermite spline over (in-tangent, value, out-tangent) keyframes,
 * tangents scaled by the interval (glTF CUBICSPLINE; cubic,
 * scene.c:280-300). */
static void interp_cubic(float *d, const float *a, const float *b, float t, float dur,
                         unsigned n)
{
	float t2 = t * t, t3 = t2 * t;
	float c0 = 2 * t3 - 3 * t2 + 1;
	float c1 = dur * (t3 - 2 * t2 + t);
	float c2 = -2.0f * t3 + 3 * t2;
	float c3 = dur * (t3 - t2);
	for (unsigned k = 0; k < n; k++)
		d[k] = c0 * a[n + k] + c1 * a[2 * n + k] + c2 * b[n + k] + c3 * b[k];
}

/* Tracks at time t into the nodes' transform components and local
 * matrices (scene_updanims, scene.c:302-384); sets RS_DIRTY_TRANSFORM for
 * mesh nodes, RS_DIRTY_CAM for camera nodes. */
static void upd_anims(struct rs_scene *s, float time)
{
	for (unsigned j = 0; j < s->ntrack; j++) {
		struct rs_track *tr = &s->tracks[j];
		const struct rs_sampler *sa = &s->samplers[tr->sampler];
		const float *keys = &s->animdata[sa->kofs];
		unsigned k = find_key(keys, tr->key, sa->kcnt, time);
		tr->key = k;
		float t0 = keys[k], t1 = keys[k + 1];
		float tc = rf_max(rf_min(time, t1), t0);
		float dur = t1 - t0;
		float t = (tc - t0) / dur;
		unsigned comp = tr->tgt == RS_TGT_ROT ? 4 : 3;
		unsigned elem = sa->interp == RS_INTERP_CUBIC ? 3 : 1;
		const float *v0 = &s->animdata[sa->dofs + elem * comp * k];
		const float *v1 = &s->animdata[sa->dofs + elem * comp * (k + 1)];
		float v[4];
		if (sa->interp == RS_INTERP_STEP) {
			for (unsigned i = 0; i < comp; i++)
				v[i] = v0[i];
		} else if (sa->interp == RS_INTERP_LINEAR) {
			if (comp == 4)
				interp_slerp(v, v0, v1, t);
			else
				interp_linear(v, v0, v1, t);
		} else {
			interp_cubic(v, v0, v1, t, dur, comp);
			if (comp == 4) {
				float l = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] +
				  v[3] * v[3]);
				for (int i = 0; i < 4; i++)
					v[i] *= l;
			}
		}
		struct rs_node *n = &s->nodes[tr->node];
		float *dst = tr->tgt == RS_TGT_TRANS ? n->trans :
		  (tr->tgt == RS_TGT_ROT ? n->rot : n->scale);
		for (unsigned i = 0; i < comp; i++)
			dst[i] = v[i];
		/* the local matrix once per run of tracks on one node */
		if (j + 1 < s->ntrack && s->tracks[j + 1].node == tr->node)
			continue;
		combine_trs(n->loc, n->trans, n->rot, n->scale);
		if (n->flags & RS_MESH)
			s->dirty |= RS_DIRTY_TRANSFORM;
		else if (n->flags & RS_CAM)
			s->dirty |= RS_DIRTY_CAM;
	}
}

/* World matrices, parents before children (scene_updtransforms,
 * scene.c:386-399: node 0 is the root). */
static void upd_transforms(struct rs_scene *s)
{
	if (!s->nnode)
		return;
	memcpy(s->nodes[0].glob, s->nodes[0].loc, sizeof(s->nodes[0].glob));
	for (unsigned i = 1; i < s->nnode; i++)
		m4_mul(s->nodes[i].glob, s->nodes[s->nodes[i].parent].glob, s->nodes[i].loc);
}

int rs_camera(struct rs_scene *s, const float eye[3], const float target[3],
              float vfov, float focdist, float focangle)
{
	memcpy(s->eye, eye, sizeof(s->eye));
	memcpy(s->target, target, sizeof(s->target));
	s->vfov = vfov;
	s->focdist = focdist;
	s->focangle = focangle;
	return 0;
}

int rs_background(struct rs_scene *s, float r, float g, float b)
{
	s->bg[0] = r;
	s->bg[1] = g;
	s->bg[2] = b;
	return 0;
}

unsigned int rs_tricount(const struct rs_scene *s)
{
	unsigned n = 0;
	for (unsigned i = 0; i < s->nmesh; i++)
		n += s->meshes[i].ntri;
	return n;
}

static struct vec3 v3(float x, float y, float z)
{
	struct vec3 r = {x, y, z};
	return r;
}

static struct vec3 v3unit(struct vec3 v)
{
	float il = 1.0f / sqrtf(rv_dot(v, v));
	return rv_scale(v, il);
}

static struct vec3 v3cross(struct vec3 a, struct vec3 b)
{
	return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
	  a.x * b.y - a.y * b.x);
}

/* rcam of a camera at e looking along the unit vector fwd (calc_cam,
 * scene.c:401-409, then upd_rcam, main.c:159-195). */
static void set_rcam(struct rdata *rd, struct vec3 e, struct vec3 fwd, float vfov,
                     float focdist, float focangle, unsigned int w, unsigned int h)
{
	struct vec3 ri = v3unit(v3cross(fwd, v3(0.0f, 1.0f, 0.0f)));
	struct vec3 up = v3unit(v3cross(ri, fwd));
	float fw = (float)w, fh = (float)h;
	float aspect = fw / fh;
	float tanvfov = tanf((float)(0.5f * vfov * PI_D / 180.0f));
	float tanfangle = tanf((float)(0.5f * focangle * PI_D / 180.0f));
	struct rcam *c = &rd->cam;
	memset(c, 0, sizeof(*c));
	c->eye = e;
	c->tanvfov = tanvfov;
	c->ri = ri;
	c->focdist = focdist;
	c->up = up;
	c->tanfangle = tanfangle;
	c->fwd = fwd;
	c->aspect = aspect;
	c->eyex8 = _mm256_set1_ps(e.x);
	c->eyey8 = _mm256_set1_ps(e.y);
	c->eyez8 = _mm256_set1_ps(e.z);
	c->rix8 = _mm256_set1_ps(ri.x);
	c->riy8 = _mm256_set1_ps(ri.y);
	c->riz8 = _mm256_set1_ps(ri.z);
	c->upx8 = _mm256_set1_ps(up.x);
	c->upy8 = _mm256_set1_ps(up.y);
	c->upz8 = _mm256_set1_ps(up.z);
	c->fwdx8 = _mm256_set1_ps(fwd.x);
	c->fwdy8 = _mm256_set1_ps(fwd.y);
	c->fwdz8 = _mm256_set1_ps(fwd.z);
	c->rw8 = _mm256_set1_ps(1.0f / fw);
	c->rh8 = _mm256_set1_ps(1.0f / fh);
	c->aspect8 = _mm256_set1_ps(aspect);
	c->fdist8 = _mm256_set1_ps(focdist);
	c->fovfdist8 = _mm256_set1_ps(2.0f * tanvfov * focdist);
	c->focangle8 = _mm256_set1_ps(tanfangle);
	c->focrad8 = _mm256_set1_ps(tanfangle * focdist);
}

int rs_set_rcam(struct rdata *rd, const float eye[3], const float target[3],
                float vfov, float focdist, float focangle, unsigned int w,
                unsigned int h)
{
	struct vec3 e = v3(eye[0], eye[1], eye[2]);
	struct vec3 fwd = v3unit(rv_sub(v3(target[0], target[1], target[2]), e));
	set_rcam(rd, e, fwd, vfov, focdist, focangle, w, h);
	return 0;
}

/* The scene graph's current camera (camera 0) from its node's world matrix:
 * eye = translation, forward = -z column (calc_cam, scene.c:401-409). */
static void node_rcam(struct rs_scene *s, struct rdata *rd)
{
	const struct rs_cam *c = &s->cams[0];
	const float *m = s->nodes[c->node >= 0 && (unsigned)c->node < s->nnode ?
	  (unsigned)c->node : 0].glob;
	struct vec3 e = v3(m[3], m[7], m[11]);
	struct vec3 col2 = v3(m[2], m[6], m[10]);
	struct vec3 fwd = v3unit(v3(-col2.x, -col2.y, -col2.z));
	set_rcam(rd, e, fwd, c->vfov, c->focdist, c->focangle, rd->width, rd->height);
}

static int inst_transforms(struct rs_scene *s, struct rdata *rd);

int rs_animate(struct rs_scene *s, struct rdata *rd, float time)
{
	if (!s->nnode)
		return rl_fail(REND_EINVAL, "rs_animate: the scene has no scene graph");
	int changed = 0;
	upd_anims(s, time);
	if (s->dirty & RS_DIRTY_TRANSFORM) {
		upd_transforms(s);
		int r = inst_transforms(s, rd);
		if (r)
			return r;
		rend_prepdynamic(rd);
		changed |= RS_DIRTY_TRANSFORM;
	}
	if (s->dirty & RS_DIRTY_CAM) {
		/* as in the reference's update (main.c:267-272), world matrices
		 * are recomputed only when a mesh node moved: an animated camera
		 * alone keeps its previous world matrix */
		if (s->ncam)
			node_rcam(s, rd);
		changed |= RS_DIRTY_CAM;
	}
	s->dirty = 0;
	return changed;
}

/* World AABB of an instance from its BLAS root (upd_rinsts, main.c:117-157). */
static void inst_aabb(struct rdata *rd, unsigned i, const float glob[16])
{
	struct rinst *ri = &rd->insts[i];
	const struct bnode8 *n = &rd->bnodes[(size_t)ri->triofs << 1];
	const float *mnx = (const float *)&n->minx8, *mny = (const float *)&n->miny8;
	const float *mnz = (const float *)&n->minz8, *mxx = (const float *)&n->maxx8;
	const float *mxy = (const float *)&n->maxy8, *mxz = (const float *)&n->maxz8;
	float nmix = mnx[0], nmiy = mny[0], nmiz = mnz[0];
	float nmax = mxx[0], nmay = mxy[0], nmaz = mxz[0];
	for (int k = 1; k < 8; k++) {
		nmix = rf_min(nmix, mnx[k]);
		nmiy = rf_min(nmiy, mny[k]);
		nmiz = rf_min(nmiz, mnz[k]);
		nmax = rf_max(nmax, mxx[k]);
		nmay = rf_max(nmay, mxy[k]);
		nmaz = rf_max(nmaz, mxz[k]);
	}
	struct vec3 mi = {FLT_MAX, FLT_MAX, FLT_MAX};
	struct vec3 ma = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (unsigned c = 0; c < 8; c++) {
		struct vec3 v;
		mulpos(&v.x, &v.y, &v.z, (float *)glob, (c & 1) ? nmix : nmax,
		  (c & 2) ? nmiy : nmay, (c & 4) ? nmiz : nmaz);
		mi = rv_min(mi, v);
		ma = rv_max(ma, v);
	}
	memset(&rd->aabbs[i], 0, sizeof(rd->aabbs[i]));
	rd->aabbs[i].min = mi;
	rd->aabbs[i].max = ma;
}

int rs_commit(struct rs_scene *s, struct rdata *rd, unsigned int w,
              unsigned int h, unsigned int spp)
{
	if (s->open || !s->nmesh || !s->ninst)
		return rl_fail(REND_EINVAL, "rs_commit: no meshes/instances or "
		  "mesh still open");
	unsigned ntri = rs_tricount(s);
	for (unsigned i = 0; i < s->nmesh; i++)
		if (s->meshes[i].ntri >= (1u << 20))
			return rl_fail(REND_EINVAL, "mesh %u has %u triangles, the id "
			  "encoding allows 2^20-1 (rend.c:24)", i, s->meshes[i].ntri);
	rend_init(rd, s->nmtl ? s->nmtl : 1, ntri, s->ninst, spp);
	if (!rd->bnodes)
		return rl_fail(REND_ENOMEM, "rend_init failed");

	unsigned *triofs = malloc(s->nmesh * sizeof(unsigned));
	if (!triofs)
		return rl_fail(REND_ENOMEM, "rs_commit");
	unsigned ofs = 0;
	for (unsigned k = 0; k < s->nmesh; k++) {
		const struct rs_mesh *m = &s->meshes[k];
		for (unsigned t = 0; t < m->ntri; t++) {
			struct rtri *rt = &rd->tris[ofs + t];
			struct rnrm *rn = &rd->nrms[ofs + t];
			memset(rt, 0, sizeof(*rt));
			memset(rn, 0, sizeof(*rn));
			const unsigned *ix = &m->idx[3 * t];
			memcpy(&rt->v0, &m->v[3 * ix[0]], sizeof(rt->v0));
			memcpy(&rt->v1, &m->v[3 * ix[1]], sizeof(rt->v1));
			memcpy(&rt->v2, &m->v[3 * ix[2]], sizeof(rt->v2));
			memcpy(&rn->n0, &m->n[3 * ix[0]], sizeof(rn->n0));
			memcpy(&rn->n1, &m->n[3 * ix[1]], sizeof(rn->n1));
			memcpy(&rn->n2, &m->n[3 * ix[2]], sizeof(rn->n2));
			rn->mtlid = m->mtl[t];
		}
		triofs[k] = ofs;
		ofs += m->ntri;
	}
	if (s->nmtl)
		memcpy(rd->mtls, s->mtls, s->nmtl * sizeof(*rd->mtls));
	for (unsigned i = 0; i < s->ninst; i++) {
		struct rinst *ri = &rd->insts[i];
		memset(ri, 0, sizeof(*ri));
		ri->flags = s->insts[i].flags;
		ri->triofs = triofs[s->insts[i].mesh];
		ri->tricnt = s->meshes[s->insts[i].mesh].ntri;
	}
	rd->instcnt = s->ninst;
	rd->bgcol = v3(s->bg[0], s->bg[1], s->bg[2]);
	free(triofs);

	rend_prepstatic(rd);
	if (s->nnode) {
		/* the imported scene at time 0 (main.c:243-245) */
		upd_anims(s, 0.0f);
		upd_transforms(s);
	}
	int r = inst_transforms(s, rd);
	if (r)
		return r;
	rend_prepdynamic(rd);
	rend_resaccum(rd, w, h);
	if (s->nnode && s->ncam)
		node_rcam(s, rd);
	else
		rs_set_rcam(rd, s->eye, s->target, s->vfov, s->focdist, s->focangle, w, h);
	/* the first rs_animate refreshes instances and camera
	 * (scene_init's dirty flags, scene.c:78) */
	s->dirty = RS_DIRTY_TRANSFORM | RS_DIRTY_CAM;
	return 0;
}

/* Instance inverse matrices and world boxes from their transforms
 * (upd_rinsts, main.c:117-157); instances of scene-graph nodes take the
 * node's world matrix. */
static int inst_transforms(struct rs_scene *s, struct rdata *rd)
{
	for (unsigned i = 0; i < s->ninst; i++) {
		struct rs_inst *in = &s->insts[i];
		if (in->node >= 0)
			memcpy(in->glob, s->nodes[in->node].glob, sizeof(in->glob));
		float inv[16];
		if (m4_inv(inv, in->glob))
			return rl_fail(REND_EINVAL, "instance %u: singular transform", i);
		memcpy(rd->insts[i].globinv, inv, sizeof(rd->insts[i].globinv));
		inst_aabb(rd, i, in->glob);
	}
	return 0;
}

/* ---- benchmark configurations -------------------------------------------- */

static float urand(uint32_t *st)
{
	*st = mix32(*st + 0x9e3779b9u);
	return (float)(*st >> 8) * (1.0f / 16777216.0f);
}

static int config_box(struct rs_scene *s, unsigned seed, float detail)
{
	uint32_t rs = seed * 0x2c1b3c6du + 1;
	float a[3];
	for (int k = 0; k < 3; k++)
		a[k] = 0.2f + 0.6f * urand(&rs);
	int msph = rs_material(s, a[0], a[1], a[2], 0);
	for (int k = 0; k < 3; k++)
		a[k] = 0.2f + 0.6f * urand(&rs);
	int mbox = rs_material(s, a[0], a[1], a[2], 0);
	int mlit = rs_material(s, 12.0f, 12.0f, 12.0f, RFLAG_EMISSIVE);
	unsigned sub = detail >= 1.0f ? 6 : (detail >= 0.25f ? 5 : 3);
	if (rs_mesh_begin(s) ||
	    rs_icosphere(s, sub, 0.0f, 0.0f, 0.0f, 1.0f, 0.1f, 2.0f, seed,
	      (unsigned)msph))
		return -1;
	const float X0 = -2.0f, X1 = 2.0f, Y0 = -1.1f, Y1 = 2.9f, Z0 = -2.0f,
	  Z1 = 2.0f;
	float p[8][3] = {{X0, Y0, Z0}, {X1, Y0, Z0}, {X1, Y1, Z0}, {X0, Y1, Z0},
	  {X0, Y0, Z1}, {X1, Y0, Z1}, {X1, Y1, Z1}, {X0, Y1, Z1}};
	/* floor, ceiling, back, left, right; front (z = Z1) open */
	if (rs_quad(s, p[4], p[5], p[1], p[0], (unsigned)mbox) ||
	    rs_quad(s, p[3], p[2], p[6], p[7], (unsigned)mbox) ||
	    rs_quad(s, p[0], p[1], p[2], p[3], (unsigned)mbox) ||
	    rs_quad(s, p[4], p[0], p[3], p[7], (unsigned)mbox) ||
	    rs_quad(s, p[1], p[5], p[6], p[2], (unsigned)mbox))
		return -1;
	float l0[3] = {-0.5f, Y1 - 0.01f, -0.5f}, l1[3] = {0.5f, Y1 - 0.01f, -0.5f};
	float l2[3] = {0.5f, Y1 - 0.01f, 0.5f}, l3[3] = {-0.5f, Y1 - 0.01f, 0.5f};
	if (rs_quad(s, l0, l1, l2, l3, (unsigned)mlit))
		return -1;
	int mesh = rs_mesh_end(s);
	float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, sc[3] = {1, 1, 1};
	if (mesh < 0 || rs_instance(s, (unsigned)mesh, t, q, sc, 0) < 0)
		return -1;
	float eye[3] = {0.0f, 0.9f, 6.5f}, tg[3] = {0.0f, 0.9f, 0.0f};
	rs_camera(s, eye, tg, 45.0f, 1.0f, 0.0f);
	rs_background(s, 0.0f, 0.0f, 0.0f);
	return 0;
}

static int config_terrain(struct rs_scene *s, unsigned seed, float detail,
                          int lights)
{
	uint32_t rs = seed * 0x2c1b3c6du + 2;
	float a[3];
	for (int k = 0; k < 3; k++)
		a[k] = 0.2f + 0.6f * urand(&rs);
	int mter = rs_material(s, a[0], a[1], a[2], 0);
	unsigned nx = 1000, nz = 500;
	if (detail < 1.0f) {
		nx = (unsigned)(nx * detail);
		nz = (unsigned)(nz * detail);
		if (nx < 4)
			nx = 4;
		if (nz < 2)
			nz = 2;
	}
	if (rs_mesh_begin(s) ||
	    rs_terrain(s, nx, nz, -1.0f, 1.0f, -0.5f, 0.5f, 0.05f, 4.0f, seed,
	      (unsigned)mter))
		return -1;
	if (lights) {
		/* Occluded lights: 8 small emitters facing up, each capping a
		 * narrow chimney above the camera, so that no terrain point sees
		 * one; they light spots on a diffuse ceiling, and the terrain
		 * receives light only after a bounce off those spots.  The caps
		 * sit on the chimney tops rather than inside them: an emitter
		 * inside would light the chimney walls at vanishing distance,
		 * and paths reaching those walls through the opening would turn
		 * the 1/d^2 of next-event estimation into fireflies.  The caps
		 * sit 0.03 below the ceiling, so each lit spot is small (radius
		 * about 0.03) and bright: most of a terrain point's incident
		 * radiance comes from a few small, well separated directions. */
		int mwall = rs_material(s, 0.3f, 0.3f, 0.3f, 0);
		int mceil = rs_material(s, 0.8f, 0.8f, 0.8f, 0);
		int mlit = rs_material(s, 150.0f, 135.0f, 112.0f, RFLAG_EMISSIVE);
		/* a room: ceiling, back and side walls, open towards the camera */
		const float cy = 0.5f, yb = -0.06f;
		float c0[3] = {-1.1f, cy, -0.6f}, c1[3] = {1.1f, cy, -0.6f};
		float c2[3] = {1.1f, cy, 0.6f}, c3[3] = {-1.1f, cy, 0.6f};
		float b0[3] = {-1.1f, yb, -0.55f}, b1[3] = {1.1f, yb, -0.55f};
		float b2[3] = {1.1f, cy, -0.55f}, b3[3] = {-1.1f, cy, -0.55f};
		float l0[3] = {-1.05f, yb, 0.6f}, l1[3] = {-1.05f, yb, -0.6f};
		float l2[3] = {-1.05f, cy, -0.6f}, l3[3] = {-1.05f, cy, 0.6f};
		float r0[3] = {1.05f, yb, -0.6f}, r1[3] = {1.05f, yb, 0.6f};
		float r2[3] = {1.05f, cy, 0.6f}, r3[3] = {1.05f, cy, -0.6f};
		if (rs_quad(s, c0, c1, c2, c3, (unsigned)mceil) ||
		    rs_quad(s, b0, b1, b2, b3, (unsigned)mwall) ||
		    rs_quad(s, l0, l1, l2, l3, (unsigned)mwall) ||
		    rs_quad(s, r0, r1, r2, r3, (unsigned)mwall))
			return -1;
		for (int i = 0; i < 8; i++) {
			float cx = -0.6f + 0.4f * (float)(i & 3);
			float cz = (i & 4) ? 0.2f : -0.25f;
			const float hw = 0.015f, y0 = -0.06f, y1 = 0.47f;
			/* four chimney walls */
			const float ox[4] = {-hw, hw, hw, -hw}, oz[4] = {-hw, -hw, hw, hw};
			for (int k = 0; k < 4; k++) {
				int k2 = (k + 1) & 3;
				float w0[3] = {cx + ox[k], y0, cz + oz[k]};
				float w1[3] = {cx + ox[k2], y0, cz + oz[k2]};
				float w2[3] = {cx + ox[k2], y1, cz + oz[k2]};
				float w3[3] = {cx + ox[k], y1, cz + oz[k]};
				if (rs_quad(s, w0, w1, w2, w3, (unsigned)mwall))
					return -1;
			}
			/* emitter closing the opening, normal +y */
			const float lw = hw, ly = y1;
			float q0[3] = {cx - lw, ly, cz - lw}, q1[3] = {cx - lw, ly, cz + lw};
			float q2[3] = {cx + lw, ly, cz + lw}, q3[3] = {cx + lw, ly, cz - lw};
			if (rs_quad(s, q0, q1, q2, q3, (unsigned)mlit))
				return -1;
		}
	}
	int mesh = rs_mesh_end(s);
	float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, sc[3] = {1, 1, 1};
	if (mesh < 0 || rs_instance(s, (unsigned)mesh, t, q, sc, 0) < 0)
		return -1;
	if (lights) {
		/* below the chimney tops, looking down at the terrain: neither
		 * the emitters nor the lit ceiling spots are in view */
		float eye[3] = {0.0f, 0.3f, 0.55f}, tg[3] = {0.0f, 0.0f, -0.1f};
		rs_camera(s, eye, tg, 40.0f, 1.0f, 0.0f);
		rs_background(s, 0.0f, 0.0f, 0.0f);
		return 0;
	}
	/* Sees only terrain at 16:9 (every primary ray hits). */
	float eye[3] = {0.0f, 0.95f, 0.25f}, tg[3] = {0.0f, 0.0f, -0.05f};
	rs_camera(s, eye, tg, 34.0f, 1.0f, 0.0f);
	rs_background(s, 0.4f, 0.4f, 0.4f);
	return 0;
}

static int config_instances(struct rs_scene *s, unsigned seed, float detail)
{
	uint32_t rs = seed * 0x2c1b3c6du + 3;
	int meshes[16];
	for (int m = 0; m < 16; m++) {
		float a[3];
		for (int k = 0; k < 3; k++)
			a[k] = 0.2f + 0.6f * urand(&rs);
		int mt = rs_material(s, a[0], a[1], a[2], 0);
		/* log-uniform triangle budget in [50k, 500k] */
		float tri = 50000.0f * powf(10.0f, urand(&rs)) * detail;
		if (rs_mesh_begin(s))
			return -1;
		if (m < 8) {
			unsigned sub = tri >= 4.0f * 81920.0f / 2.0f ? 7 :
			  (tri >= 81920.0f / 2.0f ? 6 : (tri >= 5120 ? 4 : 3));
			if (rs_icosphere(s, sub, 0, 0, 0, 1.0f, 0.1f, 2.0f,
			  seed + (unsigned)m, (unsigned)mt))
				return -1;
		} else {
			float R = 1.0f, r = 0.35f;
			unsigned nu = (unsigned)sqrtf(tri / 2.0f * R / r);
			unsigned nv = (unsigned)(tri / 2.0f / (float)(nu ? nu : 1));
			if (nu < 3)
				nu = 3;
			if (nv < 3)
				nv = 3;
			if (rs_torus(s, nu, nv, R, r, 0.05f, 3.0f, seed + (unsigned)m,
			  (unsigned)mt))
				return -1;
		}
		if ((meshes[m] = rs_mesh_end(s)) < 0)
			return -1;
	}
	int mlit = rs_material(s, 3.0f, 3.0f, 3.0f, RFLAG_EMISSIVE);
	float l0[3] = {-20, 140, -20}, l1[3] = {120, 140, -20};
	float l2[3] = {120, 140, 120}, l3[3] = {-20, 140, 120};
	/* winding l0..l3 gives the normal -y: the light faces the instances */
	if (rs_mesh_begin(s) || rs_quad(s, l0, l1, l2, l3, (unsigned)mlit))
		return -1;
	int light = rs_mesh_end(s);
	unsigned ninst = 4095;
	for (unsigned i = 0; i < ninst; i++) {
		float t[3] = {100.0f * urand(&rs), 100.0f * urand(&rs),
		  100.0f * urand(&rs)};
		/* uniform unit quaternion (Shoemake) */
		float u1 = urand(&rs), u2 = urand(&rs), u3 = urand(&rs);
		float q[4] = {sqrtf(1 - u1) * sinf(6.2831853f * u2),
		  sqrtf(1 - u1) * cosf(6.2831853f * u2),
		  sqrtf(u1) * sinf(6.2831853f * u3), sqrtf(u1) * cosf(6.2831853f * u3)};
		float sc = 0.5f + 1.5f * urand(&rs);
		float scv[3] = {sc, sc, sc};
		unsigned m = (unsigned)(urand(&rs) * 16.0f) & 15u;
		if (rs_instance(s, (unsigned)meshes[m], t, q, scv, 0) < 0)
			return -1;
	}
	float t0[3] = {0, 0, 0}, q0[4] = {0, 0, 0, 1}, s1[3] = {1, 1, 1};
	if (rs_instance(s, (unsigned)light, t0, q0, s1, 0) < 0)
		return -1;
	float eye[3] = {-60.0f, 70.0f, -60.0f}, tg[3] = {50.0f, 45.0f, 50.0f};
	rs_camera(s, eye, tg, 50.0f, 1.0f, 0.0f);
	rs_background(s, 0.05f, 0.05f, 0.08f);
	return 0;
}

int rs_config(struct rs_scene *s, int config, unsigned int seed, float detail)
{
	if (detail <= 0.0f || detail > 1.0f)
		return rl_fail(REND_EINVAL, "detail %f", (double)detail);
	int r;
	switch (config) {
	case 1: r = config_box(s, seed, detail); break;
	case 2: r = config_terrain(s, seed, detail, 0); break;
	case 3: r = config_instances(s, seed, detail); break;
	case 5: r = config_terrain(s, seed, detail, 1); break;
	default: return rl_fail(REND_EINVAL, "config %d", config);
	}
	return r ? rl_fail(REND_ENOMEM, "config %d build failed", config) : 0;
}

int rs_instance_glob(const struct rs_scene *s, unsigned int i, float m[16])
{
	if (i >= s->ninst)
		return -1;
	memcpy(m, s->insts[i].glob, 16 * sizeof(float));
	return 0;
}

void rs_rdata_view(const struct rdata *rd, struct rs_view *v)
{
	v->mtls = rd->mtls;
	v->tris = rd->tris;
	v->nrms = rd->nrms;
	v->insts = rd->insts;
	v->aabbs = rd->aabbs;
	v->bnodes = rd->bnodes;
	v->acc = rd->acc;
	v->buf = rd->buf;
	v->instcnt = rd->instcnt;
	v->tlasofs = rd->tlasofs;
	v->width = rd->width;
	v->height = rd->height;
	v->spp = rd->spp;
	v->samples = rd->samples;
	v->rays = rd->rays;
	v->blknum = rd->blknum;
	v->bgcol[0] = rd->bgcol.x;
	v->bgcol[1] = rd->bgcol.y;
	v->bgcol[2] = rd->bgcol.z;
	v->cam = &rd->cam;
	v->camsize = (unsigned)sizeof(rd->cam);
}

void rs_rdata_camera(const struct rdata *rd, float out[18])
{
	struct CamArgs a;
	rl_camargs(&rd->cam, &a);
	memcpy(out, a.eye, 3 * sizeof(float));
	memcpy(out + 3, a.ri, 3 * sizeof(float));
	memcpy(out + 6, a.up, 3 * sizeof(float));
	memcpy(out + 9, a.fwd, 3 * sizeof(float));
	out[12] = a.rw;
	out[13] = a.rh;
	out[14] = a.aspect;
	out[15] = a.fdist;
	out[16] = a.fovfdist;
	out[17] = a.focrad;
}

void rs_rdata_setframe(struct rdata *rd, int blknum, unsigned int rays,
                       unsigned int samples, unsigned int *buf)
{
	rd->blknum = blknum;
	rd->rays = rays;
	rd->samples = samples;
	rd->buf = buf;
}

struct rdata *rs_rdata_new(void)
{
	struct rdata *rd = aligned_alloc(64, (sizeof(struct rdata) + 63) & ~(size_t)63);
	if (rd)
		memset(rd, 0, sizeof(*rd));
	return rd;
}

void rs_rdata_free(struct rdata *rd)
{
	free(rd);
}

unsigned int rs_sizeof_rdata(void)
{
	return (unsigned)sizeof(struct rdata);
}
