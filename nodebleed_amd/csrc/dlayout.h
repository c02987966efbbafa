/*
 * Device (HBM) data layout of a scene, shared by the host C code that
 * repacks and uploads it and the HIP kernels that read it.  Plain C types
 * only.
 *
 * Arena: one byte-addressed buffer holding every BLAS and the TLAS.  Within
 * one tree, child words keep the reference's encoding (rend.h:13, 24-27):
 * NODE_LEAF | offset of a dleaf (BLAS), NODE_LEAF | instance id (TLAS), or
 * the offset of a child dnode, relative to that tree's root.  Nodes sit at
 * 256-byte aligned offsets so a node is exactly two 128-byte cache lines;
 * leaves follow all nodes of their tree.
 */
#ifndef DLAYOUT_H
#define DLAYOUT_H

#include <stdint.h>

#define DNODE_LEAF  0x80000000u

/* 256 bytes.  Bounds and child words are the reference's bnode8 values;
 * inv[o] replaces perm8: bits [3c, 3c+3) give child c's traversal rank for
 * ray-direction octant o (rank 7 = nearest, visited first). */
struct dnode {
	float     minx[8];
	float     maxx[8];
	float     miny[8];
	float     maxy[8];
	float     minz[8];
	float     maxz[8];
	uint32_t  child[8];
	uint32_t  inv[8];
};

/* 192 bytes, identical to the reference's leaf4 (rend.h:32-45). */
struct dleaf {
	float     v0x[4], v0y[4], v0z[4];
	float     e0x[4], e0y[4], e0z[4];
	float     e1x[4], e1y[4], e1z[4];
	uint32_t  id[4];
	uint32_t  tricnt;
	uint32_t  pad[7];
};

/* 64 bytes: inverse transform rows (rinst.globinv), BLAS root byte offset in
 * the arena, first triangle, flags. */
struct dinst {
	float     inv[12];
	uint64_t  blas;
	uint32_t  triofs;
	uint32_t  flags;
};

/* Camera constants of one view, taken from the caller's rcam (rend.h:89-117,
 * filled by upd_rcam, main.c:159-195). */
struct CamArgs {
	float  eye[3], ri[3], up[3], fwd[3];
	float  rw, rh, aspect, fdist, fovfdist, focrad;
	int    dof;
};

/* Traversal stack: the first DEPTH entries per lane are kept in LDS (DEPTH
 * depends on the kernel variant, REND_LDS_STACK_MIN..RK_DEPTH0), deeper ones
 * spill to a global buffer of (REND_STACK_MAX - DEPTH) entries per thread. */
#define REND_STACK_MAX      64   /* reference bound, rend.c:834 assert */
/* Device arena limit: trees are addressed by 32-bit byte offsets from the
 * arena's base (4 GiB less one 256-byte node, so every aligned offset fits) */
#define RL_ARENA_MAX        0xffffff00ull
#define REND_LDS_STACK_MIN  5
#define RK_DEPTH0           15   /* variant 0 (4 waves per SIMD) */
#define RK_TRACE_VARIANTS   4
/* rk_trace_launch flags: replace exactly-zero instance-space direction
   components by a signed 1e-20 (path-tracing callers only) */
#define RK_TRACE_NUDGE      1u
/* read the build host's (Intel Xeon) RCPSS table instead of this host's */
#define RK_TRACE_XEON       2u
/* bits 8..15 of the flags: lanes a leaf or setup step waits for (0: the
   default RK_STEP_BATCH) */
#define RK_STEP_BATCH       8u
#define REND_TRACE_BLOCK 256

/* Per-launch control block of the traversal kernel: RK_NSEG ray segments,
 * one per XCD (workgroups are dispatched to the 8 XCDs round-robin), each
 * with its own claim counter on a separate 128-byte line; waves claim
 * RK_CHUNK rays at a time from their XCD's segment and steal from the other
 * segments once it is drained.  The error word follows.  A wave's lanes hold
 * rays from within one chunk (two refills of 48 lanes), so the chunk also
 * sets how coherent a wave's rays are when the ray order is (scanline
 * primaries, path queues in pixel order); measured over 64..256 it is best
 * at 96: 1080p terrain diffuse rays 361 -> 352 us against 256, instanced
 * primaries 1139 -> 872 us, instanced path-traced frames 411 -> 291 ms. */
#define RK_NSEG        8
#define RK_SEG_STRIDE  32
#define RK_CHUNK       96
#define RK_CTL_ERR     (RK_NSEG * RK_SEG_STRIDE)
#define RK_CTL_WORDS   (RK_CTL_ERR + RK_SEG_STRIDE)

/* Allocation-free closest-hit launches (rend_hip_trace_noalloc) spill into
 * a static device array sized for at most RK_NOALLOC_BLOCKS persistent
 * blocks of the 15-entry LDS stack variant: 512 x 256 threads x 49 entries
 * x 8 bytes = 49 MiB per device. */
#define RK_NOALLOC_BLOCKS    512u
#define RK_NOALLOC_DEPTH     RK_DEPTH0
#define RK_NOALLOC_MAX_RAYS  0x7fffffffu

/* Error bits written by kernels into the device status word. */
/* Diagnostic block shared by every traversal launch (after the control
 * ring): [0] sticky error bits, [1] claim flag, [4..11] the first failing
 * ray's record, [12] its index, [13] its iteration count, [14] its error. */
#define RK_DIAG_WORDS 16
#define RK_ERR_STACK  1u
#define RK_ERR_ITER   2u

#endif
