#ifndef REND_HIP_H
#define REND_HIP_H

/*
 * Out-of-band C-ABI entry points of librend beyond the reference's rend.h.
 * The reference has no counterpart for these: it returns void/0 everywhere
 * (rend.h:149-162) and traces rays only from inside rend_render.  They give
 * callers explicit ray buffers (the traversal parity and throughput contract),
 * error reporting, the path-tracing render mode and batched multi-camera
 * rendering.  All pointers named d_* are device (HBM) pointers; `stream` is a
 * hipStream_t passed as void* (NULL = the library's own stream).
 *
 * Every function returns 0 on success and a negative REND_E* code on failure;
 * rend_hip_lasterror()/rend_hip_errstr() give the last failure of the
 * calling process.
 */

#include "rend.h"
#include "rend_rays.h"

#ifdef __cplusplus
extern "C" {
#endif

#define REND_OK          0
#define REND_ENODEV     -1   /* no HIP device or HIP runtime error */
#define REND_ENOMEM     -2   /* device or host allocation failed */
#define REND_EINVAL     -3   /* invalid argument or scene */
#define REND_ESTATE     -4   /* call order: scene not prepared/uploaded */
#define REND_EOVERFLOW  -5   /* traversal stack or iteration limit hit */

/* Reciprocal used by the ray/triangle test (and the reciprocal square root
 * of the reference-exact camera rays).  The reference's results depend on
 * the CPU that runs it (Intel and AMD cores implement RCPSS/RSQRTPS
 * differently); the library captures this host's instructions at set-up. */
#define REND_RCP_HOST_TABLE  0   /* this host's RCPSS/RSQRTPS: the reference on this host, bit for bit */
#define REND_RCP_IEEE        1   /* correctly rounded 1/det */
#define REND_RCP_XEON_TABLE  2   /* the tables captured on the build host (an Intel Xeon):
                                    reproduces results computed there */

int          rend_hip_lasterror(void);
const char  *rend_hip_errstr(void);

/* Device properties (compute units, name) and the CPU whose tables the
 * current mode uses (with the result of the set-up check of the rule). */
int          rend_hip_devinfo(int *cus, char *name, unsigned int namelen);
const char  *rend_hip_rcp_cpu(void);

/* Select the reciprocal mode for later traces (default HOST_TABLE). */
int          rend_hip_set_rcp(int mode);

/* Where rend_prepdynamic builds the TLAS (rend.c:3073-3090).  HOST: the
 * reference's SAH-binned build_bvh8, byte for byte, into rd->bnodes, then
 * repacked and uploaded.  GPU: a device LBVH build (Morton order, radix tree,
 * collapsed to 8-wide nodes) from rd->aabbs at the next upload, at most 4096
 * instances; rd->bnodes' TLAS region is not written.  The trees differ, so
 * hits are the same (up to which of two instances at exactly equal distance
 * is reported) and visit counts are not. */
#define REND_TLAS_HOST  0
#define REND_TLAS_GPU   1
int          rend_hip_set_tlas_build(int mode);

/* Where the BLASes of scenes prepared by later rend_prepstatic calls are
 * built (rend_prepstatic, rend.c:3033-3071; build_bvh8, rend.c:458-709).
 * HOST (default): the reference's SAH-binned build, byte for byte, into
 * rd->bnodes, repacked and uploaded.  GPU: a device LBVH build per mesh
 * (Morton order, radix tree, collapsed to 8-wide nodes with leaves of at
 * most 4 triangles) at the next upload, from the triangles rend_prepstatic
 * reads; rd->bnodes then holds, per mesh, only a root whose first child box
 * is the mesh's bounds (what upd_rinsts reads, main.c:117-157).  The leaves
 * carry the reference's FP32 triangle data, so hits are the same up to
 * which of two triangles at exactly equal distance is reported; visit
 * counts are not. */
#define REND_BLAS_HOST  0
#define REND_BLAS_GPU   1
int          rend_hip_set_blas_build(int mode);

/* The device BLAS of the mesh at triangle offset triofs as the traversal
 * reads it (after any pending upload or device build): bytes receives its
 * size, dst (cap bytes, may be NULL) its nodes (struct dnode) and leaves
 * (struct dleaf) in the arena layout of dlayout.h, the root first. */
int          rend_hip_blas_download(struct rdata *rd, unsigned int triofs, void *dst,
                                    unsigned long long cap, unsigned long long *bytes);

/* Re-upload the scene of rd (normally done by rend_prepstatic and
 * rend_prepdynamic). */
int          rend_hip_upload(struct rdata *rd);
/* Upload (or, with REND_TLAS_GPU, build on the device) what rend_prepstatic
 * and rend_prepdynamic changed since the last upload, and wait; the next
 * trace or frame would do this itself. */
int          rend_hip_sync_scene(struct rdata *rd);
/* The device TLAS as the traversal reads it (after any pending upload or
 * device build): nodes receives its node count, dst (cap bytes, may be NULL)
 * the nodes in the arena layout (dlayout.h: struct dnode, 256 bytes). */
int          rend_hip_tlas_download(struct rdata *rd, void *dst, unsigned long long cap,
                                    unsigned int *nodes);

/* Closest hit for n rays of d_rays into d_hits (device buffers). Reference
 * semantics: intersect_tlas (rend.c:1976-2155) per ray, starting t = tmax.
 * stats may be NULL; when given, visit counters are accumulated into it
 * (slower, counting launch). */
int          rend_hip_trace(struct rdata *rd, const struct rray *d_rays,
                            struct rhit *d_hits, unsigned int n,
                            void *stream, struct rtrace_stats *stats);

/* rend_hip_trace without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); it has no counterpart in the reference.
 * Same kernels and results as rend_hip_trace.  The scene must already be on
 * the device (rend_hip_sync_scene after rend_prepstatic/rend_prepdynamic,
 * else REND_ESTATE), n <= 2^31 - 1 rays (else REND_EINVAL), no visit
 * counting.  The traversal stack spills into a static 49 MiB array per
 * device that the library holds until the process ends; every call on a
 * device shares it, so calls that overlap on one device race (calls on
 * different streams are ordered by the library outside graph capture). */
int          rend_hip_trace_noalloc(struct rdata *rd, const struct rray *d_rays,
                                    struct rhit *d_hits, unsigned int n, void *stream);

/* Any hit within (0, tmax): d_occ[i] = 1 when occluded, else 0.  Reference
 * semantics: intersect_any_tlas (rend.c:2883-2998). */
int          rend_hip_occluded(struct rdata *rd, const struct rray *d_rays,
                               unsigned char *d_occ, unsigned int n,
                               void *stream);

/* Host-buffer conveniences: copy in, trace, copy out (synchronous). */
int          rend_hip_trace_host(struct rdata *rd, const struct rray *rays,
                                 struct rhit *hits, unsigned int n,
                                 struct rtrace_stats *stats);
int          rend_hip_occluded_host(struct rdata *rd, const struct rray *rays,
                                    unsigned char *occ, unsigned int n);

/* Traversal launch shape: persistent blocks per CU (0 = occupancy limit,
 * the default) and the number of idle lanes that triggers a wave refill
 * (1..64, default 48). */
int          rend_hip_set_tuning(unsigned int grid_mult, unsigned int refill);
/* Lanes of a wave that wait at leaves (or at instance entries and ray
 * set-ups) before the wave runs that step while other lanes still have
 * interior nodes to visit (1..64; 0 = the default, 8).  Results never
 * depend on it: each ray's visits keep their order. */
int          rend_hip_set_step_batch(unsigned int lanes);
/* Wave-level counters of the last counting trace (rend_hip_trace with
 * stats): wave loop iterations, lanes holding a ray summed over those
 * iterations (lane utilisation = out[1] / (64 out[0])), leaf steps and
 * set-up steps run. */
int          rend_hip_trace_lanes(unsigned long long out[4]);

/* Traversal kernel variant (0..3): LDS stack entries / waves per SIMD of
 * 15/4, 11/5, 8/6 and 5/8 (deeper entries go to a global spill area).  Results are
 * identical for every variant; only speed differs.  Default 0. */
int          rend_hip_set_trace_variant(int variant);

/* out: {blocks per CU, refill, variant} currently in use. */
int          rend_hip_get_tuning(unsigned int out[3]);

/* Traversal step limit per ray (default 4,194,304); a ray that reaches it,
 * or overflows the 64-entry stack, ends with REND_EOVERFLOW reported by the
 * next synchronous call together with that ray's record. */
int          rend_hip_set_iter_limit(unsigned int steps);

/* Builder hazards (bit set) met while building rd's trees; nonzero means the
 * reference's own build would be undefined or degenerate for this scene. */
unsigned int rend_hip_hazards(struct rdata *rd);

/* Bytes of the device BVH arena (BLASes + TLAS), -1 before upload. */
long long    rend_hip_arena_bytes(struct rdata *rd);

/* Frame rendering.  REND_MODE_NORMALS is the reference's active shading
 * (rend.c:3404-3563); REND_MODE_PATHTRACE is the wavefront path tracer
 * (trace3 semantics, rend.c:3205-3291, generalised). */
#define REND_MODE_NORMALS    0
#define REND_MODE_PATHTRACE  1
int          rend_hip_set_mode(int mode);

/* Path-tracing parameters: maximum segments per path, first depth at which
 * Russian roulette applies (0 = never) and paths in flight per wavefront
 * batch (0 = automatic). */
int          rend_hip_pt_config(unsigned int maxdepth, unsigned int rr_depth,
                                unsigned int batch);

/* One view of a batch: camera looking from eye at target (upd_rcam
 * semantics, main.c:159-195), resolution and samples per pixel. */
struct rend_view {
	float         eye[3], target[3];
	float         vfov, focdist, focangle;
	unsigned int  w, h, spp;
};

/* Path traces nviews views of rd's scene in one wavefront launch sequence
 * (captured once as a HIP graph and replayed while the view table is
 * unchanged).  d_acc holds sum(w*h*3) floats, the views' accumulators in
 * order (radiance sums; the caller clears them); sample_base offsets every
 * view's sample index so successive calls draw fresh samples.  A view's
 * random numbers depend only on its camera, resolution and sample indices,
 * so its image is the same whichever views share the call (two views with
 * the same camera and resolution draw the same samples).  rays (may be
 * NULL) receives extension + shadow rays traced. */
int          rend_hip_render_views(struct rdata *rd, const struct rend_view *views,
                                   unsigned int nviews, unsigned int sample_base,
                                   float *d_acc, void *stream,
                                   unsigned long long *rays);

/* Times candidate launch shapes (traversal variant, then paths in flight
 * per batch x traversal refill threshold) on one call of
 * rend_hip_render_views with the given views and keeps the fastest for
 * later calls.  out (may be NULL) receives {batch, refill, ms of the best,
 * ms of the default, variant}. */
int          rend_hip_autotune(struct rdata *rd, const struct rend_view *views,
                               unsigned int nviews, float *d_acc, float out[5]);

/* HIP-graph capture of path-tracing frames (default on; guided frames are
 * always launched directly). */
int          rend_hip_pt_graphs(int on);

/* ---- learned path guiding (no counterpart in the reference) -------------
 * A hash-grid-encoded 3x64 MLP outputs an 8-lobe vMF mixture per shading
 * point; the indirect direction is drawn from mix * mixture + (1 - mix) *
 * cosine BSDF (one-sample MIS, unbiased for 0 <= mix < 1).  The network is
 * trained online from the vertices of every train_stride-th path: `steps`
 * Adam steps of 4K records per frame, spread over its wavefront batches,
 * records drawn in proportion to their radiance targets, learning rate lr
 * for the hash grid and 0.1 lr for the MLP.  The proposal is used from the
 * first trained batch on. */
int          rend_hip_guide_enable(struct rdata *rd, float mix,
                                   unsigned int train_stride, unsigned int steps,
                                   float lr, unsigned int seed);
int          rend_hip_guide_disable(struct rdata *rd);
/* Changes mix and training steps (0 = frozen network) while enabled. */
int          rend_hip_guide_set(struct rdata *rd, float mix, unsigned int steps);
/* Training schedule while enabled: records per Adam step (default 4K),
 * learning rates of the hash grid and of the MLP, and whether records are
 * drawn in proportion to their target (importance != 0) or uniformly. */
int          rend_hip_guide_training(struct rdata *rd, unsigned int records_per_step,
                                     float lr_grid, float lr_mlp, int importance);
/* Which vertices are guided: those at depth < depth (default 1: the
 * primary hits); share = k > 0 evaluates the depth-0 proposal of a single
 * view once per run of k pixels of a row (k = 1: once per pixel, the
 * default) and uses it for all of their samples; share = 0 evaluates it per
 * path. */
int          rend_hip_guide_sampling(struct rdata *rd, unsigned int depth, int share);
/* Training target of a guided vertex: the radiance its path gathers after
 * it (0) or only the next vertex's direct light (1, default: one bounce,
 * lower variance). */
int          rend_hip_guide_target(struct rdata *rd, int next_vertex_direct);
/* Training objective: the KL divergence from the normalised target to the
 * vMF mixture (0, default) or to the density one-sample MIS actually samples,
 * mix * vMF + (1 - mix) * cos/pi (1: lobes need not cover what the BSDF
 * already samples well). */
int          rend_hip_guide_objective(struct rdata *rd, int mis_density);
/* Warm-up: the first `frames` frames after rend_hip_guide_enable train with
 * `steps` Adam steps each (when that is more than the per-frame count), so
 * the untrained proposal improves quickly (default 16 frames of 64 steps;
 * frames = 0 turns it off). */
int          rend_hip_guide_warmup(struct rdata *rd, unsigned int frames, unsigned int steps);
/* out: last training loss, records of the last batch, Adam steps, enabled. */
int          rend_hip_guide_info(struct rdata *rd, float out[4]);
/* Test access on host buffers: all parameters (GD_NPARAM floats, get/set),
 * raw network outputs for positions+normals (6 floats each, 64 outputs per
 * point), and the loss gradient over training records (its hash-grid part
 * accumulated in bf16 pairs on the device, returned as fp32). */
int          rend_hip_guide_params(struct rdata *rd, float *params, int set);
int          rend_hip_guide_forward_host(struct rdata *rd, const float *pn,
                                         unsigned int n, float *out);
int          rend_hip_guide_grad_host(struct rdata *rd, const void *records,
                                      unsigned int n, float *grad, float *loss);
/* `steps` Adam steps on n host training records (GdRecord layout, target
 * set), sampled per step as rend_hip_guide_training sets; loss = last
 * step's loss. */
int          rend_hip_guide_train_host(struct rdata *rd, const void *records,
                                       unsigned int n, unsigned int steps,
                                       float *loss);
/* Mixture records of the rendering-time inference kernel for n
 * positions+normals: per lobe weight, mean direction (3), concentration. */
int          rend_hip_guide_mix_host(struct rdata *rd, const float *pn,
                                     unsigned int n, float *out);
/* Inclusive prefix sums of n non-negative floats by the record-sampling
 * kernels, in fixed point: floor(v * 2^(60 - e)) summed exactly, with e the
 * exponent for which 2^(e-1) <= wsum < 2^e. */
int          rend_hip_guide_cdf_host(const float *values, unsigned int n, float wsum,
                                     unsigned long long *cdf);
/* The proposal's direction sampling routine (the device code, compiled for
 * the host as well) for one mixture record (GD_MIX floats: per lobe weight,
 * mean xyz, concentration) and three uniforms in [0, 1); needs no device. */
int          rend_hip_guide_sample_host(const float *mix, float u0, float u1, float u2,
                                        float dir[3]);
/* D = A (16x32) * B (32x16), row-major fp32 in, bf16 MFMA inside. */
int          rend_hip_mfma_check(const float *A, const float *B, float *D);

/* Renders one frame of rd (what rend_render does for the claiming thread)
 * and reports the rays traced; returns an error code instead of 0. */
int          rend_hip_render_frame(struct rdata *rd, unsigned int *rays);

/* Jittered camera rays of rd->cam for a w x h pixel grid, spp consecutive
 * rays per pixel (pixel-major), written to d_rays (w*h*spp records). */
int          rend_hip_gen_camrays(struct rdata *rd, struct rray *d_rays,
                                  unsigned int w, unsigned int h,
                                  unsigned int spp, unsigned int seed,
                                  void *stream);

/* The reference's own primary rays for a frame that starts after `samples`
 * accumulated samples (make_camray_pckt in rend_render, rend.c:3603-3660,
 * 3293-3402, pinhole and thin-lens cameras): blocks of the rend_staticinit
 * size (rd->width / bszx by rd->height / bszy of them, edge remainders not
 * rendered, as in the reference), 4x2-pixel packets in Morton order, the
 * reference's xorshift128+ jitter stream, the host RSQRTPS and, for a thin
 * lens, this host's libmvec cosf/sinf of the lens angle (tabulated over its
 * 2^23 values on first use: 64 MiB of device memory kept until the process
 * ends), bit for bit.
 * Ray i = (block * packets_per_block + packet) * 8 + lane; tmax = FLT_MAX.
 * d_rays holds rend_hip_camrays_ref_count(rd) records. */
unsigned long long rend_hip_camrays_ref_count(struct rdata *rd);
int          rend_hip_gen_camrays_ref(struct rdata *rd, struct rray *d_rays,
                                      unsigned int samples, void *stream);

/* One cosine-weighted diffuse ray per hit of a previous trace, about the
 * geometric normal facing the incoming ray, origin p + 1e-3 * dir; a missed
 * input gives a ray with tmax = -1. */
int          rend_hip_gen_diffuse(struct rdata *rd, const struct rray *d_prim,
                                  const struct rhit *d_hits,
                                  struct rray *d_out, unsigned int n,
                                  unsigned int seed, void *stream);

/* Copies the device accumulator into rd->acc (kept on the device between
 * frames; rd->buf is refreshed after every frame). */
int          rend_hip_download_acc(struct rdata *rd);

/* The library's HIP stream (hipStream_t as void*). */
void        *rend_hip_stream(void);

/* Device memory helpers for callers without another HIP binding. */
void        *rend_hip_malloc(unsigned long long bytes);
int          rend_hip_free(void *d);
int          rend_hip_memcpy_htod(void *d, const void *h, unsigned long long bytes);
int          rend_hip_memcpy_dtoh(void *h, const void *d, unsigned long long bytes);
int          rend_hip_sync(void);

/* Streaming device copy (bytes rounded down to 16) by a plain copy kernel;
 * used to calibrate hardware byte counters. */
int          rend_hip_copy(const void *d_src, void *d_dst, unsigned long long bytes,
                           void *stream);

/* Event timing on a stream: returns milliseconds between two marks. */
void        *rend_hip_event_create(void);
int          rend_hip_event_record(void *ev, void *stream);
float        rend_hip_event_elapsed(void *start, void *stop);
int          rend_hip_event_destroy(void *ev);

#ifdef __cplusplus
}
#endif

#endif
