/*
 * ora_halton.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Halton sequence with Faure-permuted digits, 256 dimensions, restated from
 * its definition: sampling/halton.py:27-87 (Faure permutations, digit
 * grouping with permutation tables of at most 500 entries, 32-bit digit
 * sums) and the evaluation code that script generates
 * (sampling/halton.cpp:35-51 and the per-base functions up to base 1619).
 * The tables are regenerated here, not copied.
 *
 * Also the per-pixel Sampler (sampling/image_sampler.hpp:16-38) and
 * hash_u32 (utils/hash.hpp:6-16).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORA_HALTON_DIMS 256

typedef struct {
    uint32_t base, pow_base, n_terms;
    uint32_t div[8], power[8];
    double scale;           /* (1 - 2^-24) / max_power, rounded to double */
    uint16_t* perm;         /* pow_base entries */
} halton_base;

static halton_base g_bases[ORA_HALTON_DIMS];
static int g_ready = 0;

static void build_tables(void) {
    /* first 256 primes */
    uint32_t primes[ORA_HALTON_DIMS];
    uint32_t cand = 1;
    for (int i = 0; i < ORA_HALTON_DIMS; ++i) {
        for (;;) {
            ++cand;
            int prime = 1;
            for (uint32_t k = 2; k * k <= cand; ++k)
                if (cand % k == 0) { prime = 0; break; }
            if (prime) break;
        }
        primes[i] = cand;
    }
    const uint32_t maxb = primes[ORA_HALTON_DIMS - 1];
    /* Faure permutations for every b in [0, maxb] */
    uint16_t** faure = (uint16_t**)calloc(maxb + 1, sizeof(uint16_t*));
    for (uint32_t b = 0; b <= maxb; ++b) {
        uint32_t n = b < 2 ? 1 : b;
        faure[b] = (uint16_t*)malloc(n * sizeof(uint16_t));
        if (b < 2) { faure[b][0] = 0; continue; }
        if (b == 2) { faure[b][0] = 0; faure[b][1] = 1; continue; }
        if (b & 1) {
            const uint32_t c = (b - 1) / 2;
            for (uint32_t i = 0; i < b; ++i) {
                if (i == c) { faure[b][i] = (uint16_t)c; continue; }
                const uint32_t f = faure[b - 1][i - (i > c ? 1 : 0)];
                faure[b][i] = (uint16_t)(f + (f >= c ? 1 : 0));
            }
        } else {
            const uint32_t c = b / 2;
            for (uint32_t i = 0; i < b; ++i)
                faure[b][i] = (uint16_t)(i < c ? 2 * faure[c][i] : 2 * faure[c][i - c] + 1);
        }
    }
    for (int d = 1; d < ORA_HALTON_DIMS; ++d) {
        halton_base* hb = &g_bases[d];
        const uint32_t base = primes[d];
        uint32_t digits = 1;
        uint64_t pow_base = base;
        while (pow_base * base <= 500) { pow_base *= base; ++digits; }
        uint64_t max_power = pow_base;
        while (max_power * pow_base < (1ull << 32)) max_power *= pow_base;
        hb->base = base;
        hb->pow_base = (uint32_t)pow_base;
        hb->perm = (uint16_t*)malloc(pow_base * sizeof(uint16_t));
        for (uint32_t j = 0; j < pow_base; ++j) {
            uint32_t idx = j, result = 0;
            for (uint32_t k = 0; k < digits; ++k) {
                const uint32_t rem = idx % base;
                idx /= base;
                result = result * base + faure[base][rem];
            }
            hb->perm[j] = (uint16_t)result;
        }
        uint64_t power = max_power / pow_base, div = 1;
        uint32_t t = 0;
        hb->div[t] = 1; hb->power[t] = (uint32_t)power; ++t;
        while (power / pow_base > 1) {
            div *= pow_base; power /= pow_base;
            hb->div[t] = (uint32_t)div; hb->power[t] = (uint32_t)power; ++t;
        }
        hb->div[t] = (uint32_t)(div * pow_base); hb->power[t] = 1; ++t;
        hb->n_terms = t;
        hb->scale = 0x1.fffffep-1 / (double)max_power;
    }
    for (uint32_t b = 0; b <= maxb; ++b) free(faure[b]);
    free(faure);
    g_ready = 1;
}

static inline float halton2(uint32_t index) {
    index = (index << 16) | (index >> 16);
    index = ((index & 0x00ff00ffu) << 8) | ((index & 0xff00ff00u) >> 8);
    index = ((index & 0x0f0f0f0fu) << 4) | ((index & 0xf0f0f0f0u) >> 4);
    index = ((index & 0x33333333u) << 2) | ((index & 0xccccccccu) >> 2);
    index = ((index & 0x55555555u) << 1) | ((index & 0xaaaaaaaau) >> 1);
    return (float)index * (1.0f / 4294967296.0f);
}

float ora_halton(uint32_t dim, uint32_t index) {
    if (!g_ready) build_tables();
    if (dim == 0) return halton2(index);
    if (dim >= ORA_HALTON_DIMS) return 0.0f;
    const halton_base* hb = &g_bases[dim];
    uint32_t sum = 0;
    for (uint32_t t = 0; t < hb->n_terms; ++t)
        sum += (uint32_t)hb->perm[(index / hb->div[t]) % hb->pow_base] * hb->power[t];
    return (float)((double)sum * hb->scale);
}

void ora_halton_batch(uint32_t dim, const uint32_t* index, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = ora_halton(dim, index[i]);
}

/* Table export for tests (pow_base, n_terms, perm) */
int ora_halton_table(uint32_t dim, uint32_t* pow_base, uint32_t* base, const uint16_t** perm) {
    if (!g_ready) build_tables();
    if (dim == 0 || dim >= ORA_HALTON_DIMS) return 0;
    *pow_base = g_bases[dim].pow_base;
    *base = g_bases[dim].base;
    *perm = g_bases[dim].perm;
    return 1;
}

uint32_t ora_hash_u32(uint32_t n, uint32_t seed) {
    uint32_t hash = n;
    for (uint32_t i = 0; i < 3; ++i) {
        hash *= 1936502639u;
        hash ^= hash >> 16;
        hash += seed;
    }
    return hash;
}

uint32_t ora_sampler_offset(uint32_t x, uint32_t y, uint32_t n, uint32_t seed) {
    return ora_hash_u32(x ^ ((y >> 16) | (y << 16)), seed) + n;
}

static const uint32_t k_dim_order[11] = {10, 7, 6, 5, 4, 2, 9, 8, 3, 1, 0};

float ora_sampler_get(uint32_t offset, uint32_t dim) {
    if (dim < 11) return ora_halton(k_dim_order[dim], offset);
    return ora_halton(dim, offset);
}

/* Sobol::sample (sampling/sobol.hpp:40-53) with the generator matrices
 * passed in (Sobol::Matrices::matrices, sampling/sobol.cpp:36: 52 columns
 * per dimension, dimension-major).  The loop walks the flat table from
 * dim*52 while index bits remain, as the reference does; columns past the
 * end of the table (n_cols) read as zero. */
double ora_sobol(const uint64_t* m, uint32_t n_cols, uint32_t dim, uint64_t index, uint64_t scramble) {
    uint64_t r = scramble & ~(~0ull << 52);
    for (uint32_t i = dim * 52u; index; index >>= 1, ++i)
        if ((index & 1ull) && i < n_cols) r ^= m[i];
    return (double)r * (1.0 / (double)(1ull << 52));
}

void ora_sobol_batch(const uint64_t* m, uint32_t n_cols, uint32_t dim, const uint64_t* index, int64_t n,
                     uint64_t scramble, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = ora_sobol(m, n_cols, dim, index[i], scramble);
}
