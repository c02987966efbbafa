/*
 * Internal state of the librend host library: the process-global device
 * context and the per-rdata device mirror.  Not part of the public ABI.
 */
#ifndef LIBREND_H
#define LIBREND_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "rend.h"
#include "../dlayout.h"

#define RL_CTL_SLOTS  64  /* ring of traversal control blocks (RK_CTL_WORDS) */

struct rl_global {
	int                 ready;
	hipStream_t         stream;
	int                 cus;
	char                name[256];
	int                 rcp_ieee;
	int                 rcp_mode;     /* REND_RCP_HOST_TABLE / _IEEE / _XEON_TABLE */
	uint16_t            rcp_host[4096];   /* this host's RCPSS table (approx.c) */
	uint32_t            rsqrt_host[8192]; /* this host's RSQRTPS tables */
	unsigned long long  approx_bad;   /* inputs that broke the table rule */
	char                approx_cpu[192];
	unsigned            refill;      /* idle lanes that trigger a wave refill */
	unsigned            max_iter;    /* traversal step limit per ray */
	unsigned            grid_mult;   /* persistent blocks per CU, 0 = occupancy */
	int                 variant;     /* traversal occupancy variant */
	unsigned            step_batch;  /* lanes a leaf/setup step waits for, 0 = default */
	int                 tlas_gpu;    /* rend_prepdynamic builds the TLAS on the device */
	int                 blas_gpu;    /* rend_prepstatic leaves the BLAS build to the device */
	unsigned           *ctl;         /* RL_CTL_SLOTS x RK_CTL_WORDS, then RK_DIAG_WORDS */
	unsigned           *diag;        /* error block shared by all traversal launches */
	unsigned            ctl_next;
	unsigned            last_slot;
	unsigned long long *stats;       /* 8 counters of the counting variant */
	unsigned long long  lanes[4];    /* wave-level counters of the last counting launch */
	void               *spill;       /* traversal stack spill area */
	size_t              spill_bytes;
	/* Every traversal launch and path-tracing frame shares the control
	 * ring and the spill area, so launches on different streams are put
	 * in one order: a launch on stream B waits for an event recorded after
	 * the last launch on stream A. */
	hipEvent_t          order_ev;
	hipStream_t         order_stream;
	int                 order_pending;
	int                 scratch_err;  /* result of the static spill area lookup */
};

struct rl_frame;  /* frame.c */
struct rl_pt;     /* pt.c */

/* Learned proposal state (guide.c). */
struct rl_guide {
	float    *master, *grad, *m1, *m2;  /* fp32 parameters, gradient, Adam moments */
	float    *part;                     /* per-block MLP gradient partial sums */
	void     *grid16, *w16;             /* fp16 grid / bf16 weight mirrors */
	float    *mix;                      /* per queued path mixture */
	void     *rec;                      /* training records of a batch */
	float    *scal;                     /* [0] weight sum, [1] loss */
	size_t    cap;                      /* paths the mixture buffer holds */
	size_t    rec_cap;                  /* records the record buffer holds */
	float     gmin[3], ginv[3];
	float    *dense;                    /* record targets, dense */
	unsigned long long *cdf, *tsum;     /* fixed-point target prefix sums, tile sums */
	size_t    cdf_cap;
	float     mix_prob, lr_grid, lr_mlp;
	unsigned  stride, steps, step, last_records;
	unsigned  chunk;                    /* records per training step */
	unsigned  depth;                    /* vertices at depth < depth are guided */
	unsigned  share;                    /* pixels per shared depth-0 proposal (0: none) */
	int       target;                   /* 1: train on next-vertex direct light */
	int       objective;                /* 1: fit the MIS density, not the vMF alone */
	unsigned  warm_frames, warm_steps;  /* Adam steps of the first frames after enable */
	unsigned  frames;                   /* frames trained since enable */
	int       importance;               /* draw records in proportion to target */
	int       enabled;
};

struct rl_scene {
	struct rdata        *rd;
	unsigned             maxmtls, maxtris, maxinsts;
	unsigned char       *arena;      /* device: BLASes then TLAS */
	size_t               arena_cap, arena_used, blas_bytes, tlas_bytes;
	unsigned long long   tlas;       /* byte offset of the TLAS root */
	unsigned             nblas;
	unsigned            *blas_triofs;
	unsigned long long  *blas_dev;   /* byte offset of each BLAS root */
	unsigned long long  *blas_len;   /* its bytes in the arena */
	void                *insts;      /* device dinst[] */
	size_t               insts_cap;
	unsigned             ninsts;
	void                *nrms, *mtls, *tris;
	size_t               nrms_cap, mtls_cap, tris_cap;
	float               *acc;        /* device accumulator, 3 floats/pixel */
	unsigned            *buf;        /* device ARGB8888 */
	size_t               acc_cap;
	int                  acc_dirty;
	int                  static_ok, dynamic_ok;
	int                  static_dirty, dynamic_dirty, tlas_built;
	int                  tlas_dev;   /* the pending TLAS is built on the device */
	int                  blas_dev_build; /* BLASes are built on the device at upload */
	void                *aabbs_dev;  /* instance world AABBs for the device build */
	unsigned            *tlas_info;  /* device build result: nodes, error */
	size_t               aabbs_cap;
	void                *tlas_stage; /* pinned host staging of the device build */
	size_t               tlas_stage_cap;
	unsigned             hazards;
	unsigned long long   gen;        /* bumped by every upload of scene data */
	struct rl_frame     *frame;
	struct rl_pt        *pt;
	struct rl_guide     *guide;
};

extern struct rl_global G;

int              rl_fail(int code, const char *fmt, ...);
int              rl_hipfail(int herr, const char *what);
int              rl_device(void);
struct rl_scene *rl_scene(const struct rdata *rd);
int              rl_trace_grid(int closest, int count, int rcp_ieee, unsigned *grid);
int              rl_trace_async(struct rl_scene *s, const void *d_rays,
                                void *d_hits, void *d_occ, unsigned n,
                                int closest, int count, hipStream_t st);
int              rl_check_errors(hipStream_t st);
int              rl_stream_order(hipStream_t st);
void             rl_approx_ready(void);
unsigned         rl_trace_flags(unsigned extra);
int              rl_stream_mark(hipStream_t st);
unsigned         rl_blkszx(void);
unsigned         rl_blkszy(void);
void             rl_frame_release(struct rl_scene *s);
void             rl_pt_release(struct rl_scene *s);
void             rl_guide_release(struct rl_scene *s);
struct PtArgs;
void             rl_guide_args(struct rl_scene *s, struct PtArgs *a, size_t cap);
unsigned         rl_guide_frame_steps(struct rl_guide *g);
int              rl_guide_buffers(struct rl_scene *s, size_t cap);
int              rl_guide_eval(struct rl_scene *s, const struct PtArgs *a,
                                const void *rays, const void *hits,
                                const unsigned *counts, unsigned share, hipStream_t st);
int              rl_guide_train(struct rl_scene *s, const struct PtArgs *a,
                                const void *paths, unsigned *counts,
                                unsigned steps, hipStream_t st);
int              rl_sync_scene(struct rl_scene *s);
int              rl_render_pt(struct rl_scene *s, unsigned rw, unsigned rh,
                              unsigned *rays);
int              rl_acc_ready(struct rl_scene *s, hipStream_t st);
void             rl_camargs(const struct rcam *c, struct CamArgs *a);

/* Kernel C-ABI (kern/ HIP sources). */
int rk_trace_variants(void);
int rk_trace_lds_depth(int variant);
int rk_trace_occupancy(int closest, int count, int rcp_ieee, int variant);
unsigned long long rl_approx_capture(uint16_t rcp[4096], uint32_t rsq[8192], char *cpu,
                                     size_t cpulen);
uint32_t rl_rcp_rule(const uint16_t *tab, uint32_t u);
uint32_t rl_rsqrt_rule(const uint32_t *tab, uint32_t u);
int rk_trace_set_host_rcp(const uint16_t *tab);
int rk_scratch_init(void);
unsigned long long rk_scratch_bytes(void);
int rk_trace_launch_noalloc(const void *arena, unsigned long long tlas,
                            const void *insts, const void *rays, void *hits,
                            unsigned int n, unsigned int *ctl, int rcp_ieee,
                            unsigned int cus, unsigned int refill,
                            unsigned int max_iter, unsigned int *diag, unsigned int flags,
                            void *stream);
int rk_trace_launch(const void *arena, unsigned long long tlas,
                    const void *insts, const void *rays, void *hits,
                    void *occ, unsigned int n, unsigned int *ctl,
                    void *spill, unsigned long long *stats, int closest,
                    int count, int rcp_ieee, int variant, unsigned int grid,
                    unsigned int refill, unsigned int max_iter,
                    const unsigned int *n_dev, unsigned int *diag,
                    unsigned int flags, void *stream);

#endif
