/*
 * Captures the host CPU's approximate reciprocal (x86 RCPSS) as a 4096-entry
 * mantissa table and proves, over all 2^32 single-precision inputs, that
 *
 *   rcp(x) = sign | (253 - E) << 23 | T[m >> 11]      for 1 <= E <= 252
 *          = sign | 0x7f800000                       for E == 0 (zero, denormal)
 *          = sign (signed zero)                      for 253 <= E <= 254, or +-inf
 *          = x | 0x00400000                          for NaN (quieted)
 *
 * where E is the biased exponent and m the 23-bit mantissa of x.  Intel
 * cores read 11 leading mantissa bits (pairs of equal entries), AMD Zen
 * cores 12; both satisfy the rule.
 *
 * The reference's Moeller-Trumbore test computes idet with _mm_rcp_ps
 * (rend.c:872, rend.c:1930), so bit-exact hit t/u/v need this exact table on
 * the GPU.  The tool writes a C header on stdout and exits non-zero if any
 * input disagrees.  Build: gcc -O2 -mavx2 gen_rcp_table.c
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t rcpss_bits(uint32_t u)
{
	float f, r;
	memcpy(&f, &u, 4);
	r = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(f)));
	memcpy(&u, &r, 4);
	return u;
}

static uint32_t table_rcp(const uint32_t *tab, uint32_t u)
{
	uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
	if (e == 0)
		return s | 0x7f800000u;
	if (e == 255)
		return m ? (u | 0x00400000u) : s;
	if (e >= 253)
		return s;
	return s | ((253u - e) << 23) | tab[m >> 11];
}

int main(void)
{
	static uint32_t tab[4096];
	for (uint32_t i = 0; i < 4096; i++) {
		uint32_t r = rcpss_bits(0x3f800000u | (i << 11));
		if (((r >> 23) & 0xffu) != 126u) {
			fprintf(stderr, "unexpected exponent at %u: %08x\n", i, r);
			return 1;
		}
		tab[i] = r & 0x7fffffu;
	}

	/* Exhaustive check, 8 lanes at a time. */
	uint64_t bad = 0;
	for (uint64_t b = 0; b < (1ull << 32); b += 8) {
		__m256i v = _mm256_add_epi32(_mm256_set1_epi32((int)(uint32_t)b),
		  _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7));
		__m256 r = _mm256_rcp_ps(_mm256_castsi256_ps(v));
		uint32_t got[8];
		_mm256_storeu_si256((__m256i *)got, _mm256_castps_si256(r));
		for (int k = 0; k < 8; k++) {
			uint32_t u = (uint32_t)b + (uint32_t)k;
			if (got[k] != table_rcp(tab, u) && bad++ < 8)
				fprintf(stderr, "mismatch %08x: cpu %08x table %08x\n",
				  u, got[k], table_rcp(tab, u));
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

	printf("/* Generated by gen_rcp_table.c: x86 RCPSS mantissa table.\n"
	  " * Captured on: %s\n"
	  " * Verified bit-exact against the CPU instruction for all 2^32 inputs\n"
	  " * (rule in gen_rcp_table.c).  Do not edit. */\n", model);
	printf("#ifndef RCP_TABLE_H\n#define RCP_TABLE_H\n");
	printf("#define RCP_TABLE_CPU \"%s\"\n", model);
	printf("/* Result mantissa of RCPSS for inputs 1 + i/4096 (exponent 126). */\n");
	printf("#define RCP_TABLE_INIT { \\\n");
	for (int i = 0; i < 4096; i++)
		printf("0x%06x,%s", tab[i], (i % 8 == 7) ? " \\\n" : " ");
	printf("}\n");
	printf("#ifndef RCP_TABLE_NO_HOST_ARRAY\n"
	  "static const unsigned int rcp_mant_table[4096] = RCP_TABLE_INIT;\n"
	  "#endif\n#endif\n");
	return 0;
}
