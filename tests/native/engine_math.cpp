// Host build of arithmetic the kernels share with the host
// (glazier_amd/csrc/ghip_common.h): the engine's own f32 exponential and
// logarithm, the volume's connectivity test and the plain pass's 24-bit
// pick, and the major-axis length (glazier_amd/csrc/geometry.h), so the CPU
// tests can hold them to the C library, to numpy and to the oracle.
#include "../../glazier_amd/csrc/ghip_common.h"
#include "../../glazier_amd/csrc/geometry.h"

extern "C" {
void engine_exp2_f32(const float *x, float *y, long n) {
    for (long i = 0; i < n; ++i) y[i] = ghip::exp2_f32(x[i]);
}
void engine_log2_f32(const float *x, float *y, long n) {
    for (long i = 0; i < n; ++i) y[i] = ghip::log2_f32(x[i]);
}
// The eight-neighbour pick of each (key, site) pair.
void engine_pick8(const uint32_t *key, const uint32_t *site, uint32_t *out, long n) {
    for (long i = 0; i < n; ++i) out[i] = ghip::site_pick8(key[i], site[i]);
}
// Major axis length of each of n moment sets (nine sums each) of count[i] sites.
void engine_length(const int64_t *m, const int64_t *count, int three_d, double *out, long n) {
    for (long i = 0; i < n; ++i) out[i] = ghip::length_of(m + 9 * i, count[i], three_d != 0);
}
// One byte per ring mask in [lo, hi): whether its members are one piece.
void engine_ring_table(uint32_t lo, uint32_t hi, uint8_t *out) {
    for (uint32_t m = lo; m < hi; ++m) out[m - lo] = ghip::cube_one_piece(m) ? 1 : 0;
}
}
