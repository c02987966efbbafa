#ifndef REND_RAYS_H
#define REND_RAYS_H

/*
 * Ray and hit records of librend's ray-buffer entry points (rend_hip.h).
 * Kept free of x86 headers so that HIP device code can include it.
 */

/* One ray: origin, maximum distance, direction.  32 bytes. */
struct rray {
	float  ox, oy, oz, tmax;
	float  dx, dy, dz, pad;
};

/* Closest hit: t (tmax when missed), barycentrics u, v and
 * id = tri << 12 | inst (rend.c:922-923), 0xffffffff when missed.  16 bytes.*/
struct rhit {
	float         t, u, v;
	unsigned int  id;
};

/* Per-launch traversal visit counters (algorithmic traffic accounting). */
struct rtrace_stats {
	unsigned long long  nodes;   /* bnode8 visits (TLAS and BLAS) */
	unsigned long long  leaves;  /* leaf4 visits */
	unsigned long long  insts;   /* instance transforms */
	unsigned long long  rays;
};

#endif
