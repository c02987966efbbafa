/*
 * glTF 2.0 import into the scene builder (rs_gltf): the reference's
 * importer (gltf.c gltf_init, gltfimp.c imp_gltf, gltfimp.c:330-435) as a
 * small JSON reader of our own plus the same subset and rules, so that a
 * scene file gives the rdata the reference's cpy_rdata/upd_rinsts/upd_rcam
 * (main.c:53-195) give it:
 *   - the first scene's root nodes, each subtree depth first, under one
 *     root node (gltfimp.c:391-400, 205-233); a node with a camera is a
 *     camera node, else a node with a mesh is a mesh node;
 *   - materials: base colour, metallic, roughness; emission =
 *     emissiveFactor x KHR_materials_emissive_strength, and an emissive
 *     material's colour is its emission (gltf.c:179-243, gltfimp.c:97-125);
 *     none in the file gives one (0.5, 0, 0.5) material;
 *   - cameras: perspective yfov (0.5 rad if absent), focus distance 10, no
 *     lens; none gives one camera on the root node whose field of view is
 *     60 * 180 / pi degrees, as the reference computes it (gltfimp.c:366);
 *   - triangle primitives with VEC3 float POSITION and NORMAL and SCALAR
 *     u8/u16/u32 indices (others are skipped), vertex strides honoured, a
 *     primitive without material takes material 0 (gltfimp.c:127-203);
 *   - animation channels on translation, rotation and scale with STEP,
 *     LINEAR and CUBICSPLINE samplers; the key and value data of a sampler
 *     are read from the start of its accessor's buffer view, each
 *     referenced buffer view copied once per accessor in accessor order
 *     (gltfimp.c:235-328, 337-352, 402-411).
 * Numbers are converted as the reference's atof/atoi do: through double.
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rend_int.h"
#include "rend_scene.h"
#include "rend_hip.h"
#include "librend.h"

/* ---- JSON ------------------------------------------------------------------ */

enum { J_NULL, J_BOOL, J_NUM, J_STR, J_ARR, J_OBJ };

struct jtok {
	int type;
	int start, end;  /* text span (strings: inside the quotes) */
	int size;        /* array elements or object pairs */
	int next;        /* token after this one's subtree */
};

struct json {
	const char  *s;
	int          len;
	struct jtok *t;
	int          n, cap;
	int          depth;
};

static int jnew(struct json *j, int type, int start)
{
	if (j->n == j->cap) {
		int nc = j->cap ? 2 * j->cap : 256;
		struct jtok *nt = realloc(j->t, (size_t)nc * sizeof(*nt));
		if (!nt)
			return -1;
		j->t = nt;
		j->cap = nc;
	}
	j->t[j->n] = (struct jtok){type, start, start, 0, 0};
	return j->n++;
}

static void jspace(const struct json *j, int *p)
{
	while (*p < j->len && (j->s[*p] == ' ' || j->s[*p] == '\t' || j->s[*p] == '\n' ||
	  j->s[*p] == '\r'))
		(*p)++;
}

static int jvalue(struct json *j, int *p);

static int jstring(struct json *j, int *p)
{
	int k = jnew(j, J_STR, *p + 1);
	if (k < 0)
		return -1;
	for ((*p)++; *p < j->len; (*p)++) {
		char c = j->s[*p];
		if (c == '\\') {
			(*p)++;
			continue;
		}
		if (c == '"') {
			j->t[k].end = (*p)++;
			j->t[k].next = j->n;
			return k;
		}
	}
	return -1;
}

static int jvalue(struct json *j, int *p)
{
	jspace(j, p);
	if (*p >= j->len || ++j->depth > 256)
		return -1;
	const char c = j->s[*p];
	int k;
	if (c == '{' || c == '[') {
		const int obj = c == '{';
		k = jnew(j, obj ? J_OBJ : J_ARR, *p);
		if (k < 0)
			return -1;
		(*p)++;
		jspace(j, p);
		if (*p < j->len && j->s[*p] == (obj ? '}' : ']')) {
			(*p)++;
		} else {
			for (;;) {
				if (obj) {
					jspace(j, p);
					if (*p >= j->len || j->s[*p] != '"' || jstring(j, p) < 0)
						return -1;
					jspace(j, p);
					if (*p >= j->len || j->s[*p] != ':')
						return -1;
					(*p)++;
				}
				if (jvalue(j, p) < 0)
					return -1;
				j->t[k].size++;
				jspace(j, p);
				if (*p < j->len && j->s[*p] == ',') {
					(*p)++;
					continue;
				}
				if (*p < j->len && j->s[*p] == (obj ? '}' : ']')) {
					(*p)++;
					break;
				}
				return -1;
			}
		}
		j->t[k].end = *p;
	} else if (c == '"') {
		k = jstring(j, p);
		if (k < 0)
			return -1;
	} else {
		int type = J_NUM;
		if (!strncmp(j->s + *p, "true", 4) || !strncmp(j->s + *p, "false", 5))
			type = J_BOOL;
		else if (!strncmp(j->s + *p, "null", 4))
			type = J_NULL;
		k = jnew(j, type, *p);
		if (k < 0)
			return -1;
		while (*p < j->len && !strchr(",]} \t\r\n", j->s[*p]))
			(*p)++;
		if (j->t[k].start == *p)
			return -1;
		j->t[k].end = *p;
	}
	j->t[k].next = j->n;
	j->depth--;
	return k;
}

/* Value of key in object token o, or -1. */
static int jget(const struct json *j, int o, const char *key)
{
	if (o < 0 || j->t[o].type != J_OBJ)
		return -1;
	size_t kl = strlen(key);
	int k = o + 1;
	for (int i = 0; i < j->t[o].size; i++) {
		const struct jtok *kt = &j->t[k];
		int v = kt->next;
		if ((size_t)(kt->end - kt->start) == kl && !memcmp(j->s + kt->start, key, kl))
			return v;
		k = j->t[v].next;
	}
	return -1;
}

/* Element i of array token a, or -1. */
static int jat(const struct json *j, int a, int i)
{
	if (a < 0 || j->t[a].type != J_ARR || i < 0 || i >= j->t[a].size)
		return -1;
	int k = a + 1;
	while (i--)
		k = j->t[k].next;
	return k;
}

static int jlen(const struct json *j, int a)
{
	return a >= 0 && j->t[a].type == J_ARR ? j->t[a].size : 0;
}

/* atof/atoi of a token's text (the reference's conversions). */
static double jnum(const struct json *j, int k, double dflt)
{
	if (k < 0)
		return dflt;
	char buf[64];
	int n = j->t[k].end - j->t[k].start;
	if (n <= 0 || n >= (int)sizeof(buf))
		return dflt;
	memcpy(buf, j->s + j->t[k].start, (size_t)n);
	buf[n] = 0;
	return strtod(buf, NULL);
}

static long jint(const struct json *j, int k, long dflt)
{
	if (k < 0)
		return dflt;
	char buf[32];
	int n = j->t[k].end - j->t[k].start;
	if (n <= 0 || n >= (int)sizeof(buf))
		return dflt;
	memcpy(buf, j->s + j->t[k].start, (size_t)n);
	buf[n] = 0;
	return strtol(buf, NULL, 10);
}

/* Substring test on a string token (the reference matches with strstr). */
static int jhas(const struct json *j, int k, const char *sub)
{
	if (k < 0)
		return 0;
	char buf[64];
	int n = j->t[k].end - j->t[k].start;
	if (n < 0 || n >= (int)sizeof(buf))
		return 0;
	memcpy(buf, j->s + j->t[k].start, (size_t)n);
	buf[n] = 0;
	return strstr(buf, sub) != NULL;
}

/* n floats of array key `key` of object o into v when it has exactly n. */
static void jfloats(const struct json *j, int o, const char *key, float *v, int n)
{
	int a = jget(j, o, key);
	if (jlen(j, a) != n)
		return;
	for (int i = 0; i < n; i++)
		v[i] = (float)jnum(j, jat(j, a, i), 0.0);
}

/* ---- files ----------------------------------------------------------------- */

static void *slurp(const char *path, size_t *len)
{
	FILE *f = fopen(path, "rb");
	if (!f)
		return NULL;
	fseek(f, 0, SEEK_END);
	long n = ftell(f);
	fseek(f, 0, SEEK_SET);
	char *b = n >= 0 ? malloc((size_t)n + 1) : NULL;
	if (b && fread(b, 1, (size_t)n, f) != (size_t)n) {
		free(b);
		b = NULL;
	}
	fclose(f);
	if (b) {
		b[n] = 0;
		*len = (size_t)n;
	}
	return b;
}

/* ---- import ------------------------------------------------------------------ */

struct acc {
	int       view;
	unsigned  count, byteofs, comp;
	int       vec;   /* components: 1 SCALAR, 2, 3, 4, (9, 16 for matrices) */
};

struct imp {
	struct json         j;
	const unsigned char *bin;
	size_t               binlen;
	struct rs_scene     *s;
	struct acc          *accs;
	int                  nacc;
	unsigned            *mesh_flags;
	int                 *mesh_id;
	int                 *nodemap;
	int                  nnode;
};

static int view_field(const struct imp *m, int view, const char *key, long dflt)
{
	int bvs = jget(&m->j, 0, "bufferViews");
	return (int)jint(&m->j, jget(&m->j, jat(&m->j, bvs, view), key), dflt);
}

static int oom(void)
{
	return rl_fail(REND_ENOMEM, "gltf: out of memory");
}

static int read_accessors(struct imp *m)
{
	int a = jget(&m->j, 0, "accessors");
	m->nacc = jlen(&m->j, a);
	m->accs = calloc((size_t)(m->nacc ? m->nacc : 1), sizeof(*m->accs));
	if (!m->accs)
		return oom();
	for (int i = 0; i < m->nacc; i++) {
		int o = jat(&m->j, a, i);
		struct acc *c = &m->accs[i];
		c->view = (int)jint(&m->j, jget(&m->j, o, "bufferView"), -1);
		c->count = (unsigned)jint(&m->j, jget(&m->j, o, "count"), 0);
		c->byteofs = (unsigned)jint(&m->j, jget(&m->j, o, "byteOffset"), 0);
		c->comp = (unsigned)jint(&m->j, jget(&m->j, o, "componentType"), 0);
		int t = jget(&m->j, o, "type");
		c->vec = jhas(&m->j, t, "SCALAR") ? 1 : jhas(&m->j, t, "VEC2") ? 2 :
		  jhas(&m->j, t, "VEC3") ? 3 : jhas(&m->j, t, "VEC4") ? 4 :
		  jhas(&m->j, t, "MAT2") ? 4 : jhas(&m->j, t, "MAT3") ? 9 :
		  jhas(&m->j, t, "MAT4") ? 16 : 0;
	}
	return 0;
}

/* Bytes [ofs, ofs + n) of the binary buffer, or NULL when out of range. */
static const unsigned char *bin_at(const struct imp *m, size_t ofs, size_t n)
{
	return ofs <= m->binlen && n <= m->binlen - ofs ? m->bin + ofs : NULL;
}

static int import_materials(struct imp *m)
{
	int ms = jget(&m->j, 0, "materials");
	int n = jlen(&m->j, ms);
	for (int i = 0; i < n; i++) {
		int o = jat(&m->j, ms, i);
		float col[3] = {1.0f, 1.0f, 1.0f}, metallic = 0.0f, rough = 0.5f;
		float refractive = 0.0f, estrength = 0.0f, efac[3] = {1.0f, 1.0f, 1.0f};
		int pbr = jget(&m->j, o, "pbrMetallicRoughness");
		int bc = jget(&m->j, pbr, "baseColorFactor");
		if (jlen(&m->j, bc) == 4)
			for (int k = 0; k < 3; k++)
				col[k] = (float)jnum(&m->j, jat(&m->j, bc, k), 1.0);
		metallic = (float)jnum(&m->j, jget(&m->j, pbr, "metallicFactor"), 0.0);
		rough = (float)jnum(&m->j, jget(&m->j, pbr, "roughnessFactor"), 0.5);
		jfloats(&m->j, o, "emissiveFactor", efac, 3);
		int ext = jget(&m->j, o, "extensions");
		estrength = (float)jnum(&m->j, jget(&m->j,
		  jget(&m->j, ext, "KHR_materials_emissive_strength"), "emissiveStrength"), 0.0);
		refractive = (float)jnum(&m->j, jget(&m->j,
		  jget(&m->j, ext, "KHR_materials_transmission"), "transmissionFactor"), 0.0);
		float em[3] = {efac[0] * estrength, efac[1] * estrength, efac[2] * estrength};
		unsigned flags = (refractive != 0.0f ? RS_REFRACTIVE : 0u) |
		  (em[0] + em[1] + em[2] > 0.0f ? RS_EMISSIVE : 0u);
		int id = rs_material(m->s, col[0], col[1], col[2], 0);
		if (id < 0 || rsi_mtl_set(m->s, (unsigned)id, metallic, rough, flags,
		  (flags & RS_EMISSIVE) ? em : NULL))
			return oom();
	}
	if (!n) {
		int id = rs_material(m->s, 0.5f, 0.0f, 0.5f, 0);
		if (id < 0 || rsi_mtl_set(m->s, (unsigned)id, 0.0f, 0.5f, 0, NULL))
			return oom();
	}
	return 0;
}

static int import_cameras(struct imp *m)
{
	int cs = jget(&m->j, 0, "cameras");
	int n = jlen(&m->j, cs);
	for (int i = 0; i < n; i++) {
		int o = jat(&m->j, cs, i);
		float yfov = (float)jnum(&m->j, jget(&m->j, jget(&m->j, o, "perspective"), "yfov"),
		  0.5);
		if (rsi_cam(m->s, (float)(yfov * 180.0f / 3.14159265), 10.0f, 0.0f, -1) < 0)
			return oom();
	}
	if (!n && rsi_cam(m->s, (float)(60.0f * 180.0f / 3.14159265), 10.0f, 0.0f, 0) < 0)
		return oom();
	return 0;
}

static int import_meshes(struct imp *m)
{
	int ms = jget(&m->j, 0, "meshes");
	int n = jlen(&m->j, ms);
	m->mesh_flags = calloc((size_t)(n ? n : 1), sizeof(unsigned));
	m->mesh_id = calloc((size_t)(n ? n : 1), sizeof(int));
	if (!m->mesh_flags || !m->mesh_id)
		return oom();
	unsigned nmtl = rs_mtlcount(m->s);
	for (int i = 0; i < n; i++) {
		int prims = jget(&m->j, jat(&m->j, ms, i), "primitives");
		if (rs_mesh_begin(m->s))
			return oom();
		for (int k = 0; k < jlen(&m->j, prims); k++) {
			int p = jat(&m->j, prims, k);
			int at = jget(&m->j, p, "attributes");
			long ii = jint(&m->j, jget(&m->j, p, "indices"), -1);
			long pi = jint(&m->j, jget(&m->j, at, "POSITION"), -1);
			long ni = jint(&m->j, jget(&m->j, at, "NORMAL"), -1);
			long mi = jint(&m->j, jget(&m->j, p, "material"), -1);
			if (ii < 0 || pi < 0 || ni < 0 || ii >= m->nacc || pi >= m->nacc || ni >= m->nacc)
				continue;
			const struct acc *ia = &m->accs[ii], *va = &m->accs[pi], *na = &m->accs[ni];
			if (ia->vec != 1 || va->vec != 3 || va->comp != 5126 || na->vec != 3 ||
			    na->comp != 5126 || (ia->comp != 5121 && ia->comp != 5123 && ia->comp != 5125))
				continue;
			if (va->count != na->count || ia->view < 0 || va->view < 0 || na->view < 0)
				return rl_fail(REND_EINVAL, "gltf: mesh %d primitive %d: accessors", i, k);
			/* vertices and normals: stride beyond 12 bytes skipped */
			unsigned nv = va->count;
			float *v = malloc((size_t)(nv ? nv : 1) * 6 * sizeof(float));
			unsigned *idx = malloc((size_t)(ia->count ? ia->count : 1) * sizeof(unsigned));
			if (!v || !idx) {
				free(v);
				free(idx);
				return oom();
			}
			int bad = 0;
			const struct acc *src[2] = {va, na};
			for (int a = 0; a < 2 && !bad; a++) {
				int stride = view_field(m, src[a]->view, "byteStride", 0);
				size_t step = stride > 12 ? (size_t)stride : 12;
				size_t base = (size_t)view_field(m, src[a]->view, "byteOffset", 0) +
				  src[a]->byteofs;
				for (unsigned q = 0; q < nv; q++) {
					const unsigned char *b = bin_at(m, base + q * step, 12);
					if (!b) {
						bad = 1;
						break;
					}
					memcpy(&v[(size_t)a * 3 * nv + 3 * q], b, 12);
				}
			}
			unsigned isz = ia->comp == 5121 ? 1 : (ia->comp == 5123 ? 2 : 4);
			size_t ibase = (size_t)view_field(m, ia->view, "byteOffset", 0) + ia->byteofs;
			for (unsigned q = 0; q < ia->count && !bad; q++) {
				const unsigned char *b = bin_at(m, ibase + (size_t)q * isz, isz);
				if (!b) {
					bad = 1;
					break;
				}
				uint32_t x = 0;
				memcpy(&x, b, isz);
				idx[q] = x;
			}
			unsigned mtl = mi < 0 ? 0u : (unsigned)mi;
			int r = bad || mtl >= nmtl ? -1 :
			  rs_mesh_data(m->s, v, nv, v + 3 * nv, idx, ia->count / 3, NULL, mtl);
			free(v);
			free(idx);
			if (r)
				return rl_fail(REND_EINVAL, "gltf: mesh %d primitive %d: data out of "
				  "range", i, k);
			/* the mesh carries its materials' flags (gltfimp.c:196) */
			m->mesh_flags[i] |= rsi_mtl_flags(m->s, mtl);
		}
		m->mesh_id[i] = rs_mesh_end(m->s);
		if (m->mesh_id[i] < 0)
			return oom();
	}
	return 0;
}

static int import_node(struct imp *m, int id, int parent, int depth)
{
	int ns = jget(&m->j, 0, "nodes");
	if (id < 0 || id >= m->nnode || depth > 256)
		return rl_fail(REND_EINVAL, "gltf: node %d", id);
	int o = jat(&m->j, ns, id);
	long mesh = jint(&m->j, jget(&m->j, o, "mesh"), -1);
	long cam = jint(&m->j, jget(&m->j, o, "camera"), -1);
	float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, sc[3] = {1, 1, 1};
	jfloats(&m->j, o, "translation", t, 3);
	jfloats(&m->j, o, "rotation", q, 4);
	jfloats(&m->j, o, "scale", sc, 3);
	unsigned flags = 0;
	int nmesh = jlen(&m->j, jget(&m->j, 0, "meshes"));
	if (cam >= 0)
		flags = RS_CAM;
	else if (mesh >= 0) {
		if (mesh >= nmesh)
			return rl_fail(REND_EINVAL, "gltf: node %d: mesh %ld", id, mesh);
		flags = RS_MESH | m->mesh_flags[mesh];
	}
	int nid = rsi_node(m->s, parent, cam < 0 && mesh >= 0 ? m->mesh_id[mesh] : -1,
	  (int)cam, flags, t, q, sc);
	if (nid < 0)
		return oom();
	if (cam >= 0 && rsi_cam_node(m->s, (unsigned)cam, nid))
		return rl_fail(REND_EINVAL, "gltf: node %d: camera %ld", id, cam);
	m->nodemap[id] = nid;
	int ch = jget(&m->j, o, "children");
	for (int i = 0; i < jlen(&m->j, ch); i++)
		if (import_node(m, (int)jint(&m->j, jat(&m->j, ch, i), -1), nid, depth + 1))
			return oom();
	return 0;
}

static int import_animations(struct imp *m)
{
	int as = jget(&m->j, 0, "animations");
	int na = jlen(&m->j, as);
	if (!na)
		return 0;
	/* accessors of any sampler: their buffer views, copied in accessor
	 * order (gltfimp.c:337-352, 402-411) */
	long *ofs = malloc((size_t)(m->nacc ? m->nacc : 1) * sizeof(long));
	if (!ofs)
		return oom();
	for (int i = 0; i < m->nacc; i++)
		ofs[i] = -1;
	for (int a = 0; a < na; a++) {
		int ss = jget(&m->j, jat(&m->j, as, a), "samplers");
		for (int i = 0; i < jlen(&m->j, ss); i++) {
			int sa = jat(&m->j, ss, i);
			long in = jint(&m->j, jget(&m->j, sa, "input"), -1);
			long out = jint(&m->j, jget(&m->j, sa, "output"), -1);
			if (in < 0 || out < 0 || in >= m->nacc || out >= m->nacc) {
				free(ofs);
				return rl_fail(REND_EINVAL, "gltf: animation %d sampler %d", a, i);
			}
			ofs[in] = ofs[out] = 0;
		}
	}
	int r = 0;
	for (int i = 0; i < m->nacc && !r; i++) {
		if (ofs[i] < 0)
			continue;
		int view = m->accs[i].view;
		size_t vo = (size_t)view_field(m, view, "byteOffset", 0);
		size_t vl = (size_t)view_field(m, view, "byteLength", 0);
		const unsigned char *b = bin_at(m, vo, vl);
		float *tmp = malloc(vl + 4);
		if (!b || !tmp || vl % 4) {
			free(tmp);
			r = rl_fail(REND_EINVAL, "gltf: animation data of accessor %d", i);
			break;
		}
		memcpy(tmp, b, vl);
		ofs[i] = rsi_animdata(m->s, tmp, (unsigned)(vl / 4));
		free(tmp);
		if (ofs[i] < 0)
			r = oom();
	}
	unsigned sofs = 0;
	for (int a = 0; a < na && !r; a++) {
		int an = jat(&m->j, as, a);
		int cs = jget(&m->j, an, "channels"), ss = jget(&m->j, an, "samplers");
		for (int i = 0; i < jlen(&m->j, cs) && !r; i++) {
			int c = jat(&m->j, cs, i);
			int tg = jget(&m->j, c, "target");
			long node = jint(&m->j, jget(&m->j, tg, "node"), -1);
			long smp = jint(&m->j, jget(&m->j, c, "sampler"), -1);
			int path = jget(&m->j, tg, "path");
			int tgt = jhas(&m->j, path, "translation") ? RS_TGT_TRANS :
			  jhas(&m->j, path, "rotation") ? RS_TGT_ROT :
			  jhas(&m->j, path, "scale") ? RS_TGT_SCALE : -1;
			if (tgt < 0 || node < 0 || node >= m->nnode || m->nodemap[node] < 0 || smp < 0 ||
			    smp >= jlen(&m->j, ss))
				r = rl_fail(REND_EINVAL, "gltf: animation %d channel %d", a, i);
			else if (rsi_track(m->s, sofs + (unsigned)smp, (unsigned)m->nodemap[node], tgt) < 0)
				r = oom();
		}
		for (int i = 0; i < jlen(&m->j, ss) && !r; i++) {
			int sa = jat(&m->j, ss, i);
			long in = jint(&m->j, jget(&m->j, sa, "input"), -1);
			long out = jint(&m->j, jget(&m->j, sa, "output"), -1);
			int ip = jget(&m->j, sa, "interpolation");
			int interp = jhas(&m->j, ip, "STEP") ? RS_INTERP_STEP :
			  jhas(&m->j, ip, "CUBICSPLINE") ? RS_INTERP_CUBIC : RS_INTERP_LINEAR;
			if (m->accs[in].count < 2)
				r = rl_fail(REND_EINVAL, "gltf: animation %d sampler %d: fewer than 2 keys",
				  a, i);
			else if (rsi_sampler(m->s, m->accs[in].count, (unsigned)ofs[in],
			  (unsigned)ofs[out], interp) < 0)
				r = oom();
		}
		sofs += (unsigned)jlen(&m->j, ss);
	}
	free(ofs);
	return r;
}

int rs_gltf(struct rs_scene *s, const char *path)
{
	struct imp m;
	memset(&m, 0, sizeof(m));
	m.s = s;
	size_t len = 0, blen = 0;
	char *text = slurp(path, &len);
	if (!text)
		return rl_fail(REND_EINVAL, "gltf: cannot read %s", path);
	size_t pl = strlen(path);
	char *bpath = malloc(pl + 4);
	unsigned char *bin = NULL;
	if (bpath && pl > 4) {
		memcpy(bpath, path, pl - 4);
		memcpy(bpath + pl - 4, "bin", 4);
		bin = slurp(bpath, &blen);
	}
	free(bpath);
	int r = 0;
	m.j.s = text;
	m.j.len = (int)len;
	int p = 0;
	if (!bin)
		r = rl_fail(REND_EINVAL, "gltf: no binary buffer next to %s", path);
	else if (len > (size_t)1 << 30 || jvalue(&m.j, &p) != 0 || m.j.t[0].type != J_OBJ)
		r = rl_fail(REND_EINVAL, "gltf: %s is not a JSON object", path);
	m.bin = bin;
	m.binlen = blen;
	int roots = -1;
	if (!r && !rsi_empty(s))
		r = rl_fail(REND_EINVAL, "gltf: the scene builder must be empty");
	int bufs = r ? -1 : jget(&m.j, 0, "buffers");
	if (!r && bufs >= 0 && jlen(&m.j, bufs) != 1)
		r = rl_fail(REND_EINVAL, "gltf: expected one buffer");
	if (!r) {
		roots = jget(&m.j, jat(&m.j, jget(&m.j, 0, "scenes"), 0), "nodes");
		m.nnode = jlen(&m.j, jget(&m.j, 0, "nodes"));
		if (!jlen(&m.j, jget(&m.j, 0, "meshes")) || !jlen(&m.j, roots) || !m.nnode)
			r = rl_fail(REND_EINVAL, "gltf: expected meshes, nodes and a scene with root "
			  "nodes");
	}
	if (!r)
		m.nodemap = malloc((size_t)m.nnode * sizeof(int));
	if (!r && !m.nodemap)
		r = rl_fail(REND_ENOMEM, "gltf");
	for (int i = 0; !r && i < m.nnode; i++)
		m.nodemap[i] = -1;
	if (!r && (read_accessors(&m) || import_materials(&m) || import_cameras(&m) ||
	    import_meshes(&m)))
		r = rend_hip_lasterror();
	if (!r) {
		const float t0[3] = {0, 0, 0}, q0[4] = {0, 0, 0, 1}, s1[3] = {1, 1, 1};
		if (rsi_node(s, -1, -1, -1, 0, t0, q0, s1) != 0)
			r = oom();
	}
	for (int i = 0; !r && i < jlen(&m.j, roots); i++)
		r = import_node(&m, (int)jint(&m.j, jat(&m.j, roots, i), -1), 0, 0) ?
		  rend_hip_lasterror() : 0;
	if (!r && import_animations(&m))
		r = rend_hip_lasterror();
	if (!r && rsi_instances_from_nodes(s))
		r = rl_fail(REND_EINVAL, "gltf: more than 4096 mesh instances");
	free(m.j.t);
	free(m.accs);
	free(m.mesh_flags);
	free(m.mesh_id);
	free(m.nodemap);
	free(bin);
	free(text);
	return r;
}
