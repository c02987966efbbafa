#ifndef REND_SCENE_H
#define REND_SCENE_H

/*
 * Procedural scene construction for librend.  The reference loads glTF
 * scenes into its scene graph (scene.h) and flattens them into rdata with
 * cpy_rdata/upd_rinsts/upd_rcam (main.c:53-195); this builder produces the
 * same rdata contents from seeded generators (displaced icospheres, tori,
 * terrain grids, quads) and runs the same flattening, so that benchmark
 * scenes need no files.
 *
 * Usage: rs_new, add materials, build meshes (rs_mesh_begin, generators,
 * rs_mesh_end), add instances, set the camera, then rs_commit into an rdata
 * (which calls rend_init, rend_prepstatic, rend_prepdynamic, rend_resaccum).
 */

#include "rend.h"

#ifdef __cplusplus
extern "C" {
#endif

struct rs_scene;

struct rs_scene *rs_new(void);
void             rs_free(struct rs_scene *s);

/* Material (rmtl); flags = 0x1000 (EMISSIVE) makes col the emitted radiance.
 * Returns the material id or -1. */
int  rs_material(struct rs_scene *s, float r, float g, float b,
                 unsigned int flags);

/* Mesh under construction: generators append triangles with the given
 * material.  rs_mesh_end returns the mesh id or -1. */
int  rs_mesh_begin(struct rs_scene *s);
int  rs_mesh_end(struct rs_scene *s);

/* Icosahedron subdivided `subdiv` times (20*4^subdiv triangles), radius r,
 * each vertex moved radially by amp * fbm(freq * p, seed). */
int  rs_icosphere(struct rs_scene *s, unsigned int subdiv, float cx, float cy,
                  float cz, float r, float amp, float freq, unsigned int seed,
                  unsigned int mtl);
/* Torus around the y axis with nu x nv quads (2*nu*nv triangles), tube
 * vertices moved along the tube normal by amp * fbm(freq * p, seed). */
int  rs_torus(struct rs_scene *s, unsigned int nu, unsigned int nv, float R,
              float r, float amp, float freq, unsigned int seed,
              unsigned int mtl);
/* Height field y = amp * fbm(freq * (x, 0.37, z), seed) on an nx x nz cell
 * grid over [x0,x1] x [z0,z1], two triangles per cell. */
int  rs_terrain(struct rs_scene *s, unsigned int nx, unsigned int nz,
                float x0, float x1, float z0, float z1, float amp, float freq,
                unsigned int seed, unsigned int mtl);
/* Indexed triangles: nv vertices (3 floats each), optional vertex normals
 * (NULL: area-weighted normals of these triangles), ntri index triples into
 * this block of vertices, and per-triangle material ids (mtls) or one
 * material for all (mtls == NULL).  Returns 0, or -1 on an index or material
 * out of range. */
int  rs_mesh_data(struct rs_scene *s, const float *v, unsigned int nv, const float *n,
                  const unsigned int *idx, unsigned int ntri, const unsigned int *mtls,
                  unsigned int mtl);
/* Quad p0 p1 p2 p3 (counter-clockwise seen from the front) as 2 triangles. */
int  rs_quad(struct rs_scene *s, const float p0[3], const float p1[3],
             const float p2[3], const float p3[3], unsigned int mtl);

/* Instance of a mesh with world transform T * R(q) * S (scene.c:176-185);
 * q = (x, y, z, w) unit quaternion.  Returns the instance id or -1. */
int  rs_instance(struct rs_scene *s, unsigned int mesh, const float t[3],
                 const float q[4], const float scale[3], unsigned int flags);

/* Camera looking from eye to target; vfov in degrees (upd_rcam,
 * main.c:159-195; calc_cam, scene.c:401-409). */
int  rs_camera(struct rs_scene *s, const float eye[3], const float target[3],
               float vfov, float focdist, float focangle);
int  rs_background(struct rs_scene *s, float r, float g, float b);

/* Flatten into rd at w x h with spp samples per pixel per frame
 * (cpy_rdata, upd_rinsts, upd_rcam; main.c:53-195).  A scene with a scene
 * graph (rs_gltf) is flattened at animation time 0. */
int  rs_commit(struct rs_scene *s, struct rdata *rd, unsigned int w,
               unsigned int h, unsigned int spp);

/* glTF 2.0 scene (a .gltf file and its one external .bin with the same
 * name) into the builder, with the reference importer's subset and rules
 * (gltf.h:4-11, gltf.c, gltfimp.c:330-435): the first scene's node tree
 * under a root node, meshes of triangle primitives with float POSITION and
 * NORMAL and u8/u16/u32 indices, materials (baseColorFactor,
 * metallicFactor, roughnessFactor; emission = emissiveFactor x
 * KHR_materials_emissive_strength), perspective cameras (camera 0 is
 * used), node translation/rotation/scale and their keyframe animations
 * (STEP, LINEAR, CUBICSPLINE).  Every mesh node becomes an instance.
 * Returns 0, or a negative error with rend_hip_errstr() set. */
int  rs_gltf(struct rs_scene *s, const char *path);

/* Advance a scene-graph scene to `time` seconds (scene_updanims,
 * scene_updtransforms, scene_updcams; upd_rinsts and rend_prepdynamic when
 * a mesh node moved, upd_rcam when a camera did: main.c:253-278).  Returns
 * a mask of what changed (1 instances and TLAS, 2 camera) or a negative
 * error. */
int  rs_animate(struct rs_scene *s, struct rdata *rd, float time);

/* Recompute rd->cam for a camera after rs_commit (rendering w x h). */
int  rs_set_rcam(struct rdata *rd, const float eye[3], const float target[3],
                 float vfov, float focdist, float focangle, unsigned int w,
                 unsigned int h);

unsigned int rs_tricount(const struct rs_scene *s);
unsigned int rs_mtlcount(const struct rs_scene *s);

/* Benchmark scenes.  config: 1 sphere-in-box, 2 terrain (1000 x 500 cells),
 * 3 instanced meshes, 5 occluded-lights terrain.  detail < 1 scales triangle
 * counts down for tests (1 = full size).  Returns 0 or a negative error. */
int  rs_config(struct rs_scene *s, int config, unsigned int seed,
               float detail);

/* Deterministic fbm value noise in [-1, 1] (exported for tests). */
float rs_noise(float x, float y, float z, unsigned int seed);

/* World matrix (16 floats, row-major) of instance i as built by rs_instance. */
int  rs_instance_glob(const struct rs_scene *s, unsigned int i, float m[16]);

/* Pointers and counters of an rdata, for bindings that do not mirror its
 * layout (its camera block holds 32-byte aligned vector fields). */
struct rs_view {
	void          *mtls, *tris, *nrms, *insts, *aabbs, *bnodes, *acc, *buf;
	unsigned int   instcnt, tlasofs, width, height, spp, samples, rays;
	int            blknum;
	float          bgcol[3];
	const void    *cam;      /* &rd->cam, camsize bytes */
	unsigned int   camsize;
};
void rs_rdata_view(const struct rdata *rd, struct rs_view *v);
/* Camera constants of rd->cam: eye, ri, up, fwd (3 each), 1/w, 1/h, aspect,
 * focus distance, 2 tan(vfov/2) focdist, lens radius (lane 0 of the vector
 * fields upd_rcam fills, main.c:159-195). */
void rs_rdata_camera(const struct rdata *rd, float out[18]);
/* Sets the frame-protocol fields the reference's main loop writes
 * (main.c:424-426) and the application colour buffer. */
void rs_rdata_setframe(struct rdata *rd, int blknum, unsigned int rays,
                       unsigned int samples, unsigned int *buf);

/* rdata with the alignment its vector fields need; zero-initialised. */
struct rdata *rs_rdata_new(void);
void          rs_rdata_free(struct rdata *rd);
unsigned int  rs_sizeof_rdata(void);

#ifdef __cplusplus
}
#endif

#endif
