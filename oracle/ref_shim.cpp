/*
 * ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * C entry points into the parts of the reference (psychopath_cpp) that
 * compile from their own sources with no third-party dependency.  Built by
 * oracle/Makefile (target `ref`) straight from the reference tree into
 * oracle/_ref/libpsyref.so; used only to pin the oracle in tests.
 *   Halton::sample           sampling/halton.cpp:2154
 *   Sobol::sample, matrices  sampling/sobol.hpp:40-53, sampling/sobol.cpp:36
 *   hash_u32                 utils/hash.hpp:6-16
 *   SIMD::invert_44_matrix   utils/simd.hpp:110-205
 *   spectrum_xyz_to_p        color/spectrum_grid.h:21-135
 */
#include <stdint.h>
#include "numtype.h"
#include "halton.hpp"
#include "sobol.hpp"
#include "hash.hpp"
#include "simd.hpp"
#include "spectrum_grid.h"

extern "C" {
float psyref_halton(uint32_t dim, uint32_t index) { return Halton::sample(dim, index); }
double psyref_sobol(uint32_t dim, uint64_t index, uint64_t scramble) {
    return Sobol::sample(dim, index, scramble);
}
uint64_t psyref_sobol_matrix(uint32_t i) { return Sobol::Matrices::matrices[i]; }
uint32_t psyref_sobol_dims(void) { return Sobol::Matrices::num_dimensions; }
uint32_t psyref_hash_u32(uint32_t n, uint32_t seed) { return hash_u32(n, seed); }
float psyref_invert44(float* m16) {
    alignas(16) float m[16];
    for (int i = 0; i < 16; ++i) m[i] = m16[i];
    const float r = SIMD::invert_44_matrix(m);
    for (int i = 0; i < 16; ++i) m16[i] = m[i];
    return r;
}
float psyref_spectrum_xyz_to_p(float lambda, const float* xyz) { return spectrum_xyz_to_p(lambda, xyz); }
}
