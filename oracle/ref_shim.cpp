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
// The grid tables themselves, for callers that hand them to the oracle and
// the GPU library (which embed none).
void psyref_spectrum_tables(const int32_t** cells, int* w, int* h, const float** points, int* n_points,
                            int* n_samples, float* smin, float* smax, const float** mat_xy_to_uv) {
    static_assert(sizeof(spectrum_grid_cell_t) == 8 * sizeof(int32_t), "cell layout");
    static_assert(sizeof(spectrum_data_point_t) == (4 + 81) * sizeof(float), "point layout");
    *cells = (const int32_t*)spectrum_grid;
    *w = spectrum_grid_width;
    *h = spectrum_grid_height;
    *points = (const float*)spectrum_data_points;
    *n_points = (int)(sizeof(spectrum_data_points) / sizeof(spectrum_data_points[0]));
    *n_samples = spectrum_num_samples;
    *smin = spectrum_sample_min;
    *smax = spectrum_sample_max;
    *mat_xy_to_uv = spectrum_mat_xy_to_uv;
}
float psyref_color_to_spectrum(const float* rgb, float lambda) {
    // Color_to_spectrum (color/color.hpp:494-496) without the Color type:
    // Color_to_XYZ then XYZ_to_spectrum, in the reference's float order
    float xyz[3];
    xyz[0] = (rgb[0] * 0.4339f) + (rgb[1] * 0.3762f) + (rgb[2] * 0.1899f);
    xyz[1] = (rgb[0] * 0.2126f) + (rgb[1] * 0.7152f) + (rgb[2] * 0.0722f);
    xyz[2] = (rgb[0] * 0.0177f) + (rgb[1] * 0.1095f) + (rgb[2] * 0.8728f);
    return spectrum_xyz_to_p(lambda, xyz) * (1.0f / equal_energy_reflectance);
}
}
