/*
 * Internal declarations shared by the librend host modules (C11).
 * Nothing here is part of the public ABI.
 */
#ifndef REND_INT_H
#define REND_INT_H

#include <float.h>
#include <stddef.h>
#include <stdint.h>

#include "rend.h"

/* Instance id bits in a hit id: id = tri << 12 | inst (rend.c:24-25). */
#define INST_ID_BITS  12
#define INST_ID_MASK  0xfffu

/* Flags from the reference's types.h:4-18 that the renderer reads. */
#define RFLAG_EMISSIVE  0x1000u

/* Builder hazard bits (cases where the reference's build is undefined or
 * produces a tree whose boxes do not enclose their primitives). */
#define BVH_HAZ_NOSPLIT   1u  /* no finite SAH cost: reference reads an uninitialised split */
#define BVH_HAZ_HDEGEN    2u  /* forced 1:n-1 horizontal split keeps stale bounds (rend.c:547-551) */
#define BVH_HAZ_VDEGEN    4u  /* empty vertical side: reference leaves a count as a node offset */
#define BVH_HAZ_STACK     8u  /* builder stack would exceed 64 entries (rend.c:639) */

/* ---- scene builder internals (scenegen.c, gltf.c) ----------------------- */

/* Node flags (types.h:4-18) and scene-graph constants (scene.h:63-100). */
#define RS_MESH       0x0001u
#define RS_CAM        0x0002u
#define RS_DYNAMIC    0x0200u
#define RS_EMISSIVE   0x1000u
#define RS_REFRACTIVE 0x2000u
#define RS_TGT_TRANS  0
#define RS_TGT_ROT    1
#define RS_TGT_SCALE  2
#define RS_INTERP_STEP   0
#define RS_INTERP_LINEAR 1
#define RS_INTERP_CUBIC  2
#define RS_DIRTY_TRANSFORM 1
#define RS_DIRTY_CAM       2

struct rs_scene;
int      rsi_node(struct rs_scene *s, int parent, int mesh, int cam, unsigned flags,
                  const float t[3], const float q[4], const float sc[3]);
unsigned rsi_node_flags(struct rs_scene *s, unsigned node, unsigned set);
int      rsi_cam(struct rs_scene *s, float vfov, float focdist, float focangle, int node);
int      rsi_cam_node(struct rs_scene *s, unsigned cam, int node);
unsigned rsi_mtl_flags(const struct rs_scene *s, unsigned id);
int      rsi_empty(const struct rs_scene *s);
int      rsi_mtl_set(struct rs_scene *s, unsigned id, float metallic, float roughness,
                     unsigned flags, const float *col);
long     rsi_animdata(struct rs_scene *s, const float *v, unsigned n);
int      rsi_sampler(struct rs_scene *s, unsigned kcnt, unsigned kofs, unsigned dofs,
                     int interp);
int      rsi_track(struct rs_scene *s, unsigned sampler, unsigned node, int tgt);
int      rsi_instances_from_nodes(struct rs_scene *s);

/* ---- float helpers with the reference's operation order ---------------- */

static inline struct vec3 rv_add(struct vec3 a, struct vec3 b)
{ struct vec3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline struct vec3 rv_sub(struct vec3 a, struct vec3 b)
{ struct vec3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline struct vec3 rv_mul(struct vec3 a, struct vec3 b)
{ struct vec3 r = {a.x * b.x, a.y * b.y, a.z * b.z}; return r; }
static inline struct vec3 rv_scale(struct vec3 v, float s)
{ struct vec3 r = {v.x * s, v.y * s, v.z * s}; return r; }
/* min/max with the reference's macro semantics (util.h:13-14): the second
 * operand wins when the comparison is false, including NaN. */
static inline float rf_min(float a, float b) { return a < b ? a : b; }
static inline float rf_max(float a, float b) { return a > b ? a : b; }
static inline struct vec3 rv_min(struct vec3 a, struct vec3 b)
{ struct vec3 r = {rf_min(a.x, b.x), rf_min(a.y, b.y), rf_min(a.z, b.z)}; return r; }
static inline struct vec3 rv_max(struct vec3 a, struct vec3 b)
{ struct vec3 r = {rf_max(a.x, b.x), rf_max(a.y, b.y), rf_max(a.z, b.z)}; return r; }
static inline float rv_c(struct vec3 v, unsigned a)
{ return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
static inline float rv_dot(struct vec3 a, struct vec3 b)
{ return a.x * b.x + a.y * b.y + a.z * b.z; }

/* ---- BVH8 build (bvh8.c) ------------------------------------------------ */

struct bvh8_stats {
	unsigned int nodes;     /* interior nodes, the reference's return value */
	unsigned int leaves;    /* leaf4 records (BLAS) or instance leaves (TLAS) */
	unsigned int used;      /* bytes of the arena written, from offset 0 */
	unsigned int hazards;   /* BVH_HAZ_* */
};

unsigned int bvh8_build(struct bnode8 *nodes, const struct aabb *aabbs,
                        unsigned int *imap, const struct rtri *tris,
                        unsigned int pcnt, struct vec3 rootmin,
                        struct vec3 rootmax, struct bvh8_stats *st);

/* Bytes reachable from the root of a BLAS (tris != 0) or TLAS arena; -1 on a
 * malformed tree (offset out of range, misaligned, or cycle depth > 64). */
long long bvh8_extent(const struct bnode8 *root, int is_blas, size_t limit);

/* Device repack (dlayout.h): returns the repacked byte size or -1 on a
 * malformed tree / short buffer.  _size only measures. */
long long bvh8_repack_size(const struct bnode8 *root, int is_blas,
                           size_t limit, unsigned *nnodes, unsigned *nleaves);
long long bvh8_repack(void *dst, size_t cap, const struct bnode8 *root,
                      int is_blas, size_t limit);

#endif
