/*
 * Captures the host CPU's approximate reciprocal square root (x86 RSQRTPS)
 * as two 4096-entry tables (odd and even biased exponent) and proves, over
 * all 2^32 single-precision inputs, that
 *
 *   rsqrt(x) = T[p][m >> 11] - (((E - B_p) / 2) << 23)   x > 0 normal,
 *                                                        p = E & 1, B_1 = 127,
 *                                                        B_0 = 128
 *            = sign | 0x7f800000    zero or denormal (the sign of x kept)
 *            = 0                    +inf
 *            = 0xffc00000           x < 0 (normal or -inf)
 *            = x | 0x00400000       NaN (quieted)
 *
 * where E is the biased exponent, m the 23-bit mantissa and T[p][i] the
 * instruction's result for the input with exponent B_p and mantissa i << 11
 * (Intel cores read 11 leading mantissa bits, AMD Zen cores 12).
 *
 * The reference normalises camera-ray directions with _mm256_rsqrt_ps
 * (rend.c:3370, rend.c:3397), so bit-exact primary rays need this exact
 * table on the GPU.  The tool writes a C header on stdout and exits non-zero
 * if any input disagrees.  Build: gcc -O2 -mavx2 gen_rsqrt_table.c
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t rsqrtss_bits(uint32_t u)
{
	float f, r;
	memcpy(&f, &u, 4);
	r = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(f)));
	memcpy(&u, &r, 4);
	return u;
}

static uint32_t table_rsqrt(const uint32_t *tab, uint32_t u)
{
	uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
	if (e == 255)
		return m ? (u | 0x00400000u) : (s ? 0xffc00000u : 0u);
	if (e == 0)
		return s | 0x7f800000u;
	if (s)
		return 0xffc00000u;
	uint32_t p = e & 1u;
	int b = p ? 127 : 128;
	int sh = ((int)e - b) / 2;  /* exact: e - b is even */
	return (uint32_t)((int32_t)tab[(p << 12) | (m >> 11)] - sh * (1 << 23));
}

int main(void)
{
	static uint32_t tab[8192];
	for (uint32_t p = 0; p < 2; p++)
		for (uint32_t i = 0; i < 4096; i++)
			tab[(p << 12) | i] = rsqrtss_bits(((p ? 127u : 128u) << 23) | (i << 11));

	/* Exhaustive check, 8 lanes at a time. */
	uint64_t bad = 0;
	for (uint64_t b = 0; b < (1ull << 32); b += 8) {
		__m256i v = _mm256_add_epi32(_mm256_set1_epi32((int)(uint32_t)b),
		  _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7));
		__m256 r = _mm256_rsqrt_ps(_mm256_castsi256_ps(v));
		uint32_t got[8];
		_mm256_storeu_si256((__m256i *)got, _mm256_castps_si256(r));
		for (int k = 0; k < 8; k++) {
			uint32_t u = (uint32_t)b + (uint32_t)k;
			if (got[k] != table_rsqrt(tab, u) && bad++ < 8)
				fprintf(stderr, "mismatch %08x: cpu %08x table %08x\n",
				  u, got[k], table_rsqrt(tab, u));
		}
	}
	if (bad) {
		fprintf(stderr, "%llu mismatches\n", (unsigned long long)bad);
		return 1;
	}

	char model[256] = "unknown";
	FILE *f = fopen("/proc/cpuinfo", "r");
	if (f) {
		char line[512];
		while (fgets(line, sizeof(line), f))
			if (!strncmp(line, "model name", 10)) {
				char *c = strchr(line, ':');
				if (c) {
					snprintf(model, sizeof(model), "%s", c + 2);
					model[strcspn(model, "\n")] = 0;
				}
				break;
			}
		fclose(f);
	}
	printf("/* Generated by gen_rsqrt_table.c: x86 RSQRTPS tables.\n"
	       " * Captured on: %s\n"
	       " * Verified bit-exact against the CPU instruction for all 2^32 inputs\n"
	       " * (rule in gen_rsqrt_table.c).  Do not edit. */\n"
	       "#ifndef RSQRT_TABLE_H\n#define RSQRT_TABLE_H\n"
	       "#define RSQRT_TABLE_CPU \"%s\"\n"
	       "/* [p][i]: result for the input with biased exponent 127 (p = 1) or\n"
	       " * 128 (p = 0) and mantissa i << 11; index (p << 12) | i. */\n"
	       "#define RSQRT_TABLE_INIT { \\\n", model, model);
	for (int i = 0; i < 8192; i++)
		printf("0x%08x,%s", tab[i], (i % 8 == 7) ? " \\\n" : " ");
	printf("}\n#endif\n");
	return 0;
}
