/*
 * Captures the x86 RCPPS approximate reciprocal as a 2048-entry table and
 * verifies a table-driven emulation of it over all 2^32 single-precision
 * inputs.
 *
 * The reference fuses depth with _mm256_rcp_ps / _mm_rcp_ps
 * (src/fusion/update_functions.cpp:1246, 1373, 1386).  On the capture host
 * the instruction's output depends only on the sign, the exponent and the
 * top 11 mantissa bits of its input, so a table indexed by those 11 bits,
 * plus an exponent shift, reproduces it bit for bit.  The table is written
 * as a C header that the HIP kernels include, so the GPU reproduces the
 * host CPU's reciprocal rather than an IEEE division.
 *
 * Build and run on an x86-64 host with SSE:
 *   gcc -O2 -msse2 gen_rcp_table.c -o gen_rcp_table && ./gen_rcp_table rcp_table.h
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint32_t table[2048];

/* Table-driven emulation; the HIP kernels implement the same mapping. */
static uint32_t emulate(uint32_t u)
{
    uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
    if (e == 0) return s | 0x7f800000u;                 /* zero / denormal -> inf */
    if (e == 255) return m ? (u | 0x00400000u) : s;     /* NaN -> quiet, inf -> 0 */
    uint32_t t = table[m >> 12];
    int re = (int)((t >> 23) & 0xffu) + 127 - (int)e;
    if (re <= 0) return s;                              /* result flushed to zero */
    return s | ((uint32_t)re << 23) | (t & 0x7fffffu);
}

int main(int argc, char **argv)
{
    for (uint32_t i = 0; i < 2048; i++)
        table[i] = f2u(_mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(u2f(0x3f800000u | (i << 12))))));

    /* exhaustive check, 4 lanes at a time */
    uint64_t mismatches = 0;
    uint32_t first_bad = 0;
    for (uint64_t b = 0; b < (1ull << 32); b += 4) {
        __m128i v = _mm_set_epi32((int)(b + 3), (int)(b + 2), (int)(b + 1), (int)b);
        __m128 r = _mm_rcp_ps(_mm_castsi128_ps(v));
        uint32_t out[4];
        _mm_storeu_si128((__m128i *)out, _mm_castps_si128(r));
        for (int k = 0; k < 4; k++) {
            uint32_t in = (uint32_t)(b + k);
            if (emulate(in) != out[k]) {
                if (!mismatches) first_bad = in;
                mismatches++;
            }
        }
    }
    /* the 256-bit form must agree with the 128-bit one (the AVX2 build uses it) */
    uint64_t avx_mismatch = 0;
#ifdef __AVX__
    for (uint32_t i = 0; i < 2048; i++) {
        float o[8];
        _mm256_storeu_ps(o, _mm256_rcp_ps(_mm256_set1_ps(u2f(0x3f800000u | (i << 12) | 0x5a5u))));
        if (f2u(o[0]) != table[i]) avx_mismatch++;
    }
#endif
    fprintf(stderr, "rcpps emulation mismatches over 2^32 inputs: %llu (first 0x%08x); avx mismatches %llu\n",
            (unsigned long long)mismatches, first_bad, (unsigned long long)avx_mismatch);
    if (mismatches || avx_mismatch) return 1;

    FILE *f = argc > 1 ? fopen(argv[1], "w") : stdout;
    if (!f) return 2;
    fprintf(f, "/* Generated by gen_rcp_table.c: RCPPS output for inputs 1.m, m = i << 12.\n"
               "   Exhaustively verified against the capture host's RCPPS over all 2^32 inputs. */\n");
    fprintf(f, "#pragma once\n#include <stdint.h>\nstatic const uint32_t FF_RCP_TABLE[2048] = {\n");
    for (uint32_t i = 0; i < 2048; i++)
        fprintf(f, "0x%08xu,%s", table[i], (i % 8 == 7) ? "\n" : " ");
    fprintf(f, "};\n");
    if (f != stdout) fclose(f);
    return 0;
}
