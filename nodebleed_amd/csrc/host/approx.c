/*
 * The host CPU's approximate reciprocal (RCPSS) and reciprocal square root
 * (RSQRTPS), captured at library set-up for the GPU kernels.
 *
 * The reference computes the Moeller-Trumbore 1/det with _mm_rcp_ps
 * (rend.c:872, 1930) and normalises camera directions with _mm256_rsqrt_ps
 * (rend.c:3370, 3397), so its hits and rays depend on the CPU it runs on:
 * Intel cores read 11 leading mantissa bits, AMD Zen cores 12, with
 * different results.  Both follow one rule (proven for every 2^32 input by
 * gen_rcp_table.c and gen_rsqrt_table.c on an Intel Xeon and an AMD EPYC):
 * the result is a table of the 12 leading mantissa bits (at one exponent, or
 * two for the square root's exponent parity) with the exponent scaled, and
 * fixed results for zero, denormal, infinite and NaN inputs.  The tables are
 * read from this host's instructions and the rule is checked on a sample of
 * inputs; rend_hip_rcp_cpu() reports the CPU and whether the check held.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "librend.h"

static uint32_t rcpss(uint32_t u)
{
	float f, r;
	memcpy(&f, &u, 4);
	r = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(f)));
	memcpy(&u, &r, 4);
	return u;
}

static uint32_t rsqrtss(uint32_t u)
{
	float f, r;
	memcpy(&f, &u, 4);
	r = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(f)));
	memcpy(&u, &r, 4);
	return u;
}

/* The rules of gen_rcp_table.c / gen_rsqrt_table.c over 12-bit tables. */
uint32_t rl_rcp_rule(const uint16_t *tab, uint32_t u)
{
	uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
	if (e == 0)
		return s | 0x7f800000u;
	if (e == 255)
		return m ? (u | 0x00400000u) : s;
	if (e >= 253)
		return s;
	return s | ((253u - e) << 23) | ((uint32_t)tab[m >> 11] << 11);
}

uint32_t rl_rsqrt_rule(const uint32_t *tab, uint32_t u)
{
	uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
	if (e == 255)
		return m ? (u | 0x00400000u) : (s ? 0xffc00000u : 0u);
	if (e == 0)
		return s | 0x7f800000u;
	if (s)
		return 0xffc00000u;
	uint32_t p = e & 1u;
	int sh = ((int)e - (p ? 127 : 128)) / 2;
	return (uint32_t)((int32_t)tab[(p << 12) | (m >> 11)] - sh * (1 << 23));
}

static uint32_t mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352du;
	x ^= x >> 15;
	x *= 0x846ca68bu;
	x ^= x >> 16;
	return x;
}

/* Captures both tables from this host and checks the rules on every
 * mantissa at two exponents and on 2^20 hashed inputs.  Returns the number
 * of inputs that broke a rule (0 on every CPU the generators proved). */
unsigned long long rl_approx_capture(uint16_t rcp[4096], uint32_t rsq[8192], char *cpu,
                                     size_t cpulen)
{
	for (uint32_t i = 0; i < 4096; i++) {
		rcp[i] = (uint16_t)((rcpss(0x3f800000u | (i << 11)) >> 11) & 0xfffu);
		for (uint32_t p = 0; p < 2; p++)
			rsq[(p << 12) | i] = rsqrtss(((p ? 127u : 128u) << 23) | (i << 11));
	}
	unsigned long long bad = 0;
	for (uint32_t m = 0; m < (1u << 23); m += 7) {
		for (uint32_t e = 127; e <= 128; e++) {
			uint32_t u = (e << 23) | m;
			bad += rcpss(u) != rl_rcp_rule(rcp, u);
			bad += rsqrtss(u) != rl_rsqrt_rule(rsq, u);
		}
	}
	for (uint32_t i = 0; i < (1u << 20); i++) {
		uint32_t u = mix32(i * 0x9e3779b9u + 1u);
		bad += rcpss(u) != rl_rcp_rule(rcp, u);
		bad += rsqrtss(u) != rl_rsqrt_rule(rsq, u);
	}
	snprintf(cpu, cpulen, "unknown");
	FILE *f = fopen("/proc/cpuinfo", "r");
	if (f) {
		char line[512];
		while (fgets(line, sizeof(line), f))
			if (!strncmp(line, "model name", 10)) {
				char *c = strchr(line, ':');
				if (c) {
					snprintf(cpu, cpulen, "%s", c + 2);
					cpu[strcspn(cpu, "\n")] = 0;
				}
				break;
			}
		fclose(f);
	}
	return bad;
}
