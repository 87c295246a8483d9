// Deterministic synthetic inputs for the benchmark workloads and the parity
// tests.  Every generator uses only std::mt19937 / std::mt19937_64, whose output
// sequences are fixed by the C++ standard, and derives values with explicit
// arithmetic (no std::*_distribution, whose algorithms differ between standard
// libraries).  Host-only code: built with g++ into libtrtc_workloads.so and
// loaded through ctypes.
//
// Workloads (sizes are chosen by the caller):
//   sort keys u32 uniform ................ gen_u32_uniform(seed)
//   sort keys u32 low-entropy (a&b&c) .... gen_u32_low_entropy(seed)
//   sort keys f32 N(0,1) with -0.0/+0.0 .. gen_f32_normal_signed_zeros(...)
//   scan input int32 in [lo, hi] ......... gen_i32_range(seed, lo, hi)
//   reduce input f32 in [0, 1) ........... gen_f32_unit(seed)
//   segment sizes log-uniform ............ gen_segment_sizes(seed, ...)
//   Poisson arrivals ..................... gen_exponential(seed, rate)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>

extern "C" {

// key = raw 32-bit output of mt19937(seed)
void wl_gen_u32_uniform(uint32_t seed, uint64_t n, uint32_t* out)
{
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; i++) out[i] = (uint32_t)g();
}

// key = a & b & c of three consecutive draws: every bit is set with p = 1/8,
// so the 8-bit digit histograms are heavily skewed toward small digits.
void wl_gen_u32_low_entropy(uint32_t seed, uint64_t n, uint32_t* out)
{
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t a = (uint32_t)g();
        uint32_t b = (uint32_t)g();
        uint32_t c = (uint32_t)g();
        out[i] = a & b & c;
    }
}

// N(0,1) by the Box-Muller transform from mt19937_64(seed_normal):
//   u1 = ((r1 >> 11) + 1) * 2^-53 in (0, 1],  u2 = (r2 >> 11) * 2^-53 in [0, 1)
//   z  = sqrt(-2 ln u1) * cos(2 pi u2), rounded to float.
// Then, with an independent mt19937(seed_zero) stream r: r % 100 == 0 writes
// -0.0f and (when plant_pos_zero) r % 100 == 1 writes +0.0f, so that about 1%
// of the keys are -0.0 and 1% are +0.0 and ties between them occur.
void wl_gen_f32_normal_signed_zeros(uint64_t seed_normal, uint32_t seed_zero,
                                    int plant_pos_zero, uint64_t n, float* out)
{
    std::mt19937_64 g(seed_normal);
    std::mt19937 z(seed_zero);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r1 = g();
        uint64_t r2 = g();
        double u1 = (double)((r1 >> 11) + 1) * (1.0 / 9007199254740992.0);
        double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);
        double v = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
        float f = (float)v;
        uint32_t r = (uint32_t)z();
        if (r % 100 == 0) f = -0.0f;
        else if (plant_pos_zero && r % 100 == 1) f = 0.0f;
        out[i] = f;
    }
}

// value = (int)(u % (hi - lo + 1)) + lo with u = mt19937(seed)
void wl_gen_i32_range(uint32_t seed, int32_t lo, int32_t hi, uint64_t n, int32_t* out)
{
    std::mt19937 g(seed);
    // span == 0 encodes the full 2^32 range
    uint32_t span = (uint32_t)((int64_t)hi - (int64_t)lo + 1);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t u = (uint32_t)g();
        uint32_t r = span ? u % span : u;
        out[i] = (int32_t)(r + (uint32_t)lo);
    }
}

// value = (u >> 8) * 2^-24, exactly representable in float, in [0, 1)
void wl_gen_f32_unit(uint32_t seed, uint64_t n, float* out)
{
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t u = (uint32_t)g();
        out[i] = (float)(u >> 8) * (1.0f / 16777216.0f);
    }
}

// size = floor(exp(ln(lo) + U * (ln(hi) - ln(lo)))), U = (r >> 11) * 2^-53 from
// mt19937_64(seed); clamped to [lo, hi].  Returns the total.
uint64_t wl_gen_segment_sizes(uint64_t seed, uint64_t lo, uint64_t hi,
                              uint64_t count, uint64_t* sizes)
{
    std::mt19937_64 g(seed);
    double llo = std::log((double)lo), lhi = std::log((double)hi);
    uint64_t total = 0;
    for (uint64_t i = 0; i < count; i++) {
        double u = (double)(g() >> 11) * (1.0 / 9007199254740992.0);
        uint64_t s = (uint64_t)std::floor(std::exp(llo + u * (lhi - llo)));
        if (s < lo) s = lo;
        if (s > hi) s = hi;
        sizes[i] = s;
        total += s;
    }
    return total;
}

// Exponential inter-arrival times -ln(U)/rate with U = ((r >> 11) + 1) * 2^-53
// in (0, 1] from mt19937_64(seed): the gaps of a Poisson process.
void wl_gen_exponential(uint64_t seed, double rate, uint64_t n, double* out)
{
    std::mt19937_64 g(seed);
    for (uint64_t i = 0; i < n; i++) {
        double u = (double)((g() >> 11) + 1) * (1.0 / 9007199254740992.0);
        out[i] = -std::log(u) / rate;
    }
}

// Uniform draws in [0, 1) from mt19937_64(seed): used to pick request kinds
// and sizes in the served stream.
void wl_gen_unit_f64(uint64_t seed, uint64_t n, double* out)
{
    std::mt19937_64 g(seed);
    for (uint64_t i = 0; i < n; i++)
        out[i] = (double)(g() >> 11) * (1.0 / 9007199254740992.0);
}

// 64-bit FNV-1a over a byte buffer: used to pin generator outputs.
uint64_t wl_fnv1a64(const void* data, uint64_t nbytes)
{
    const unsigned char* p = (const unsigned char*)data;
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < nbytes; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

}  // extern "C"
