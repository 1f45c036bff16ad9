// rc_vec.h -- float3 arithmetic with the evaluation order of the reference's
// vec3 (Real Caustics/src/vec3.h), shared by host set-up code and gfx950
// kernels.  Everything is compiled with -ffp-contract=off, so a*b+c stays two
// rounded operations on both sides and the CPU oracle reproduces GPU bits.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RC_HD __host__ __device__ __forceinline__

namespace rc {

struct F3 {
    float x, y, z;
};

// Low 10 bits of v spread to every third bit (3-D Morton codes).
RC_HD uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

RC_HD F3 f3(float x, float y, float z) { return F3{x, y, z}; }
RC_HD F3 f3s(float a) { return F3{a, a, a}; }
RC_HD F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RC_HD F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RC_HD F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
RC_HD F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RC_HD F3 operator*(F3 a, float t) { return F3{a.x * t, a.y * t, a.z * t}; }
// vec3 / float multiplies by the reciprocal (vec3.h:154-157)
RC_HD F3 div_recip(F3 a, float t) {
    float s = 1.0f / t;
    return a * s;
}
// vec3 /= float divides each component (vec3.h:79-85)
RC_HD F3 div_each(F3 a, float t) { return F3{a.x / t, a.y / t, a.z / t}; }
RC_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RC_HD float absdot(F3 a, F3 b) { return fabsf(dot(a, b)); }
RC_HD F3 cross(F3 u, F3 v) { return F3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
RC_HD float len2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RC_HD F3 normalize(F3 a) { return div_recip(a, sqrtf(len2(a))); }
RC_HD float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
RC_HD float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
RC_HD float maxc(F3 a) { return smax(smax(a.x, a.y), a.z); }
RC_HD F3 reflect(F3 v, F3 n) { return v - n * (2.0f * dot(v, n)); }
RC_HD bool refract(F3 wi, F3 n, float eta, F3* wo) {
    float cos_i = dot(n, wi);
    float sin2_i = smax(0.0f, 1.0f - cos_i * cos_i);
    float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return false;
    float cos_t = sqrtf(1.0f - sin2_t);
    *wo = (-wi) * eta + n * (eta * cos_i - cos_t);
    return true;
}

// Orthonormal-ish frame from an (unnormalised) z (Frame.h:12-29).
struct Frame {
    F3 x, y, z;
};
RC_HD Frame frame_from_z(F3 z) {
    Frame f;
    f.z = z;
    F3 tx = (fabsf(z.x) > 0.99f) ? f3(0.0f, 1.0f, 0.0f) : f3(1.0f, 0.0f, 0.0f);
    f.y = normalize(cross(z, tx));
    f.x = cross(f.y, z);
    return f;
}
RC_HD F3 to_world(const Frame& f, F3 a) { return (f.x * a.x + f.y * a.y) + f.z * a.z; }
RC_HD F3 to_local(const Frame& f, F3 a) { return F3{dot(a, f.x), dot(a, f.y), dot(a, f.z)}; }

// ---------------------------------------------------------------------
// The engine's own sin / cos / acos: Taylor series of the functions,
// reduced to |r| <= pi/4 by a two-part pi/2 (k * 1.5703125 is exact).
// The same float operations run on the host oracle, so both agree bitwise.
// ---------------------------------------------------------------------
RC_HD float own_sincos(float x, int cosine) {
    const float two_over_pi = 0.6366197466850281f;
    const float c1 = 1.5703125f;
    const float c2 = 0.0004838267923332751f;
    float fk = x * two_over_pi;
    int k = (int)(fk + (fk >= 0.0f ? 0.5f : -0.5f));
    float kf = (float)k;
    float r = (x - kf * c1) - kf * c2;
    float r2 = r * r;
    float ps = -0.00019841270113829523f + r2 * 2.7557318844628753e-06f;
    ps = 0.008333333767950535f + r2 * ps;
    ps = -0.1666666716337204f + r2 * ps;
    float s = r + (r * r2) * ps;
    float pc = 2.4801587642286904e-05f + r2 * -2.755731998149713e-07f;
    pc = -0.0013888889225199819f + r2 * pc;
    pc = 0.0416666679084301f + r2 * pc;
    pc = -0.5f + r2 * pc;
    float c = 1.0f + r2 * pc;
    int q = (k + cosine) & 3;
    return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}
// Both values of own_sincos from one reduction (identical results).
RC_HD void own_sincos2(float x, float* sin_out, float* cos_out) {
    const float two_over_pi = 0.6366197466850281f;
    const float c1 = 1.5703125f;
    const float c2 = 0.0004838267923332751f;
    float fk = x * two_over_pi;
    int k = (int)(fk + (fk >= 0.0f ? 0.5f : -0.5f));
    float kf = (float)k;
    float r = (x - kf * c1) - kf * c2;
    float r2 = r * r;
    float ps = -0.00019841270113829523f + r2 * 2.7557318844628753e-06f;
    ps = 0.008333333767950535f + r2 * ps;
    ps = -0.1666666716337204f + r2 * ps;
    float s = r + (r * r2) * ps;
    float pc = 2.4801587642286904e-05f + r2 * -2.755731998149713e-07f;
    pc = -0.0013888889225199819f + r2 * pc;
    pc = 0.0416666679084301f + r2 * pc;
    pc = -0.5f + r2 * pc;
    float c = 1.0f + r2 * pc;
    int q = k & 3;
    *sin_out = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
    int qc = (k + 1) & 3;
    *cos_out = qc == 0 ? s : (qc == 1 ? c : (qc == 2 ? -s : -c));
}
RC_HD float own_sin(float x) { return own_sincos(x, 0); }
RC_HD float own_cos(float x) { return own_sincos(x, 1); }
RC_HD float own_asin_poly(float y) {
    float y2 = y * y;
    float p = 0.008390335366129875f;
    p = 0.009761609137058258f + y2 * p;
    p = 0.01155180111527443f + y2 * p;
    p = 0.013964843936264515f + y2 * p;
    p = 0.017352763563394547f + y2 * p;
    p = 0.022372158244252205f + y2 * p;
    p = 0.0303819440305233f + y2 * p;
    p = 0.0446428582072258f + y2 * p;
    p = 0.07500000298023224f + y2 * p;
    p = 0.1666666716337204f + y2 * p;
    return y + (y * y2) * p;
}
RC_HD float own_acos(float x) {
    const float pio2 = 1.5707963705062866f, pi = 3.1415927410125732f;
    float ax = fabsf(x);
    if (ax <= 0.5f) {
        float a = own_asin_poly(ax);
        return x >= 0.0f ? pio2 - a : pio2 + a;
    }
    float z = (1.0f - ax) * 0.5f;
    float a = 2.0f * own_asin_poly(sqrtf(z));
    return x >= 0.0f ? a : pi - a;
}

// ---------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the sample space.
// Dimension d of path p, iteration i, stream s = word (d & 3) of
// Philox(ctr = {d >> 2, p, i, s}, key = seed), mapped to [0,1) with 24 bits.
// ---------------------------------------------------------------------
struct U4 {
    uint32_t a, b, c, d;
};
RC_HD U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    return U4{c0, c1, c2, c3};
}
RC_HD float u24(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

// exp(x) for x <= 0 in float operations only (the same bits on the host and
// the GPU): Cody-Waite reduction by ln 2 (two-part constant) and the degree-6
// Taylor polynomial on |r| <= ln2 / 2; 0 below -80, where the result would
// approach the denormal range.
RC_HD float own_exp_neg(float x) {
    if (x < -80.0f) return 0.0f;
    float k = rintf(x * 1.44269502f);
    float r = (x - k * 0.693145752f) - k * 1.42860677e-06f;
    float p = 1.38888889e-03f;
    p = p * r + 8.33333377e-03f;
    p = p * r + 4.16666679e-02f;
    p = p * r + 1.66666672e-01f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)k);
}
// Progressive camera samples stratified across iterations (option
// camera_qmc): sample i of a pixel takes point i of the (0,2)-sequence (van der
// Corput, Sobol's second dimension) for the jitter and the base-3 radical
// inverse for the first bounce's first dimension (the reflect / refract
// choice at a dielectric), each shifted modulo 1 by a per-pixel random offset
// that does not change across iterations (Cranley-Patterson rotation), so
// every iteration is still uniformly distributed.  Values are 24-bit, in the
// word form u24() reads.
enum { STREAM_QMC_SHIFT = 4 };
RC_HD uint32_t qmc_vdc(uint32_t i) {
    i = (i << 16) | (i >> 16);
    i = ((i & 0x00FF00FFu) << 8) | ((i & 0xFF00FF00u) >> 8);
    i = ((i & 0x0F0F0F0Fu) << 4) | ((i & 0xF0F0F0F0u) >> 4);
    i = ((i & 0x33333333u) << 2) | ((i & 0xCCCCCCCCu) >> 2);
    i = ((i & 0x55555555u) << 1) | ((i & 0xAAAAAAAAu) >> 1);
    return i;
}
RC_HD uint32_t qmc_sobol2(uint32_t i) {
    uint32_t r = 0, v = 1u << 31;
    for (; i; i >>= 1, v ^= v >> 1)
        if (i & 1u) r ^= v;
    return r;
}
RC_HD uint32_t qmc_ri3_24(uint32_t i) {  // floor(2^24 * base-3 radical inverse of i)
    uint64_t num = 0, den = 1;
    for (; i; i /= 3u) {
        num = num * 3u + i % 3u;
        den *= 3u;
    }
    return (uint32_t)((num << 24) / den);
}
RC_HD uint32_t qmc_word(uint32_t v24, uint32_t shift_word) { return ((v24 + (shift_word >> 8)) & 0xFFFFFFu) << 8; }

// Mutation scale of the adaptive MLT light pass: (1/e)^(1/size), and 1 (a
// jump of up to the whole unit interval) when the adapted size is not positive.
RC_HD float mlt_step(float size) { return size > 0.0f ? own_exp_neg(-1.0f / size) : 1.0f; }
RC_HD uint32_t pick_word(U4 w, int i) { return i == 0 ? w.a : (i == 1 ? w.b : (i == 2 ? w.c : w.d)); }

// Light paths: dimensions 0..4 are the light sample, bounce k uses Philox
// block 2 + k ({BSDF u0, u1, u2, Russian roulette}).  Camera paths:
// dimensions 0, 1 are the pixel jitter, bounce k uses block 1 + k.
// The adaptive MLT light pass adds a stream of mutation draws and one of
// initial chain states.
enum { STREAM_LIGHT = 0, STREAM_CAMERA = 1, STREAM_MLT_DELTA = 2, STREAM_MLT_INIT = 3 };
RC_HD int bounce_block(int stream, int path_length) { return (stream == STREAM_LIGHT ? 2 : 1) + (path_length & 7); }

struct PathRng {
    uint32_t k0, k1, path, iter, stream;
    RC_HD U4 block(uint32_t b) const { return philox(b, path, iter, stream, k0, k1); }
    RC_HD float u(int d) const { return u24(pick_word(block((uint32_t)d >> 2), d & 3)); }
};

// BF16 <-> FP32 with round-to-nearest-even (finite values).
RC_HD uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
RC_HD float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// Spatial hash of an integer photon-map cell.
RC_HD uint32_t cell_hash(int x, int y, int z) {
    return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}
// Bucket of a cell: the low six bits are (x mod 4, y mod 4, z mod 4) and the
// rest a hash of the enclosing 4x4x4 block, so the cells of any run of four
// along each axis (a whole merge neighbourhood) occupy distinct buckets.
// Needs mask >= 63.
RC_HD uint32_t cell_bucket(int x, int y, int z, uint32_t mask) {
    uint32_t low = (uint32_t)(x & 3) | ((uint32_t)(y & 3) << 2) | ((uint32_t)(z & 3) << 4);
    return ((cell_hash(x >> 2, y >> 2, z >> 2) << 6) | low) & mask;
}

}  // namespace rc
