/*
 * oracle.c -- CPU restatement of the Real Caustics vertex-merging estimator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Built twice (see oracle/Makefile): REAL = float reproduces the reference's
 * float arithmetic in the reference's evaluation order (compiled with
 * -ffp-contract=off, SSE2, no fast-math); REAL = double is the FP64 oracle of
 * the same estimator.  Citations are "file:line" in "Real Caustics/src".
 *
 * What pins each part (tests/test_oracle_*.py):
 *   Philox            Random123 known-answer vectors.
 *   own sin/cos/acos  C library to a few ulp over the ranges used.
 *   Fresnel, GGX      closed forms: normal incidence ((n-1)/(n+1))^2, TIR,
 *                     energy of D*cos, G2 range/symmetry, sampler histogram
 *                     vs pdf, reciprocity, hand-derived mirror/refraction cases.
 *   BVH + triangle    brute-force closest hit over all triangles; builder
 *                     invariants (every triangle in one leaf, child box inside
 *                     parent box).
 *   KD gather         brute-force radius sets.
 *   lights, camera    closed forms (corner interpolation, cosine lobe moments).
 *   dVM / MIS         hand-derived one- and two-bounce values.
 *   whole image       closed-form mirror/diffuse scene and f32 vs f64 agreement.
 *   serial MLT loop   the serial xorshift + adaptive MLT light pass
 *                     (rng_serial = 1, Vertex_Merging.cpp:220-282) has no run
 *                     of the reference to compare against (the reference
 *                     source does not compile with this image's libstdc++);
 *                     it is pinned by what the Metropolis construction fixes
 *                     (tests/test_oracle_mlt.py): the chain's states, rejected
 *                     mutations repeating the state, are distributed as the
 *                     visible uniform paths, and with every path visible no
 *                     mutation happens and the photon list is the plain one.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_DOUBLE
typedef double real;
#else
typedef float real;
#endif
#define K(x) ((real)(x))

/* Global constants (main.cpp:42-50), rounded to REAL. */
static const real PI = K(3.14159265359);
static const real PI2 = K(6.28318530718);
static const real INV_PI = K(0.31830988618);
static const real PI_OVER_2 = K(1.57079632679);
static const real PI_OVER_4 = K(0.78539816339);
static const real INV_E = K(0.36787944117);
static const real INF = (real)INFINITY;

/* ------------------------------------------------------------------ */
/* vec3 with the reference's operator semantics (vec3.h)               */
/* ------------------------------------------------------------------ */
typedef struct { real x, y, z; } V;
static inline V v3(real x, real y, real z) { V r = {x, y, z}; return r; }
static inline V vs(real a) { return v3(a, a, a); }
static inline V vadd(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V vsub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V vmul(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V vscale(V a, real t) { return v3(a.x * t, a.y * t, a.z * t); }
/* operator/(vec3, float) multiplies by the reciprocal (vec3.h:154-157). */
static inline V vdivs(V a, real t) { real s = K(1) / t; return vscale(a, s); }
/* operator/=(float) divides each component (vec3.h:79-85). */
static inline V vdiv_each(V a, real t) { return v3(a.x / t, a.y / t, a.z / t); }
static inline V vneg(V a) { return v3(-a.x, -a.y, -a.z); }
static inline real vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline real vabsdot(V a, V b) { return (real)fabs(vdot(a, b)); }
static inline V vcross(V u, V v) {
    return v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline real vlen2(V a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline real vlen(V a) { return (real)sqrt(vlen2(a)); }
static inline V vnormalize(V a) { return vdivs(a, vlen(a)); }
static inline real vcomp(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
/* std::min / std::max semantics. */
static inline real smin(real a, real b) { return (b < a) ? b : a; }
static inline real smax(real a, real b) { return (a < b) ? b : a; }
static inline real vmaxc(V a) { return smax(smax(a.x, a.y), a.z); }
/* reflect (vec3.h:185-188) */
static inline V vreflect(V v, V n) { return vsub(v, vscale(n, K(2) * vdot(v, n))); }
/* refract (vec3.h:190-203) */
static int vrefract(V wi, V n, real eta, V *wo) {
    real cos_i = vdot(n, wi);
    real sin2_i = smax(K(0), K(1) - cos_i * cos_i);
    real sin2_t = eta * eta * sin2_i;
    if (sin2_t >= K(1)) return 0;
    real cos_t = (real)sqrt(K(1) - sin2_t);
    *wo = vadd(vscale(vneg(wi), eta), vscale(n, eta * cos_i - cos_t));
    return 1;
}

/* ------------------------------------------------------------------ */
/* The engine's own sin/cos/acos.  Taylor series of the mathematics,   */
/* evaluated in float without contraction so the CPU and the GPU give  */
/* identical bits.  The FP64 build and the "libm" option use libm.     */
/* ------------------------------------------------------------------ */
static int g_libm = 0;

static float own_sincos(float x, int cosine) {
    const float two_over_pi = 0.6366197466850281f;
    const float c1 = 1.5703125f;              /* 9 significant bits: k*c1 is exact */
    const float c2 = 0.0004838267923332751f;  /* float(pi/2 - c1) */
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

static float own_asin_poly(float y) {
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

static float own_acos(float x) {
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

#ifdef ORC_DOUBLE
static inline real rsin(real x) { return sin(x); }
static inline real rcos(real x) { return cos(x); }
static inline real racos(real x) { return acos(x); }
#else
static inline real rsin(real x) { return g_libm ? sinf(x) : own_sincos(x, 0); }
static inline real rcos(real x) { return g_libm ? cosf(x) : own_sincos(x, 1); }
static inline real racos(real x) { return g_libm ? acosf(x) : own_acos(x); }
#endif

double orc_math(int which, double x) {
    float xf = (float)x;
    if (which == 0) return own_sincos(xf, 0);
    if (which == 1) return own_sincos(xf, 1);
    return own_acos(xf);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                  */
/* ------------------------------------------------------------------ */
void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Per-path sample space of the plain estimator.  Dimension d of path p in
 * iteration i and stream s is word (d & 3) of Philox(ctr = {d >> 2, p, i, s},
 * key = seed), mapped to [0,1) with 24 bits, so every Philox block holds four
 * consecutive dimensions.  Light paths: dimensions 0..4 are the light sample
 * (pick, two for the position, two for the direction), bounce k uses block
 * 2 + k: its first three words are the BSDF sample and the fourth the Russian
 * roulette draw.  Camera paths: dimensions 0 and 1 are the pixel jitter and
 * bounce k uses block 1 + k the same way.  The primary sample vector the
 * reference indexes (MLT_Sampler.h:29-40) is filled from these dimensions. */
enum { STREAM_LIGHT = 0, STREAM_CAMERA = 1 };

typedef struct { uint32_t key[2]; uint32_t path, iter, stream; } PathKey;

static int bounce_block(int stream, int path_length) { return (stream == STREAM_LIGHT ? 2 : 1) + (path_length & 7); }

static real u24(uint32_t x) { return (real)(x >> 8) * K(5.9604644775390625e-08); }

static real path_u(const PathKey *k, int d) {
    uint32_t ctr[4] = {(uint32_t)d >> 2, k->path, k->iter, k->stream}, out[4];
    orc_philox(ctr, k->key, out);
    return u24(out[d & 3]);
}

/* ------------------------------------------------------------------ */
/* Serial xorshift stream of the reference (random_generators.h:18-50) */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t a; } Xorshift;
static uint32_t xorshift32(Xorshift *s) {
    uint32_t x = s->a;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return s->a = x;
}
static uint32_t random_int(Xorshift *s, uint32_t range) {
    uint32_t x, r;
    do {
        x = xorshift32(s);
        r = x % range;
    } while (x - r > (~range + 1u));
    return r;
}
static real random_float_0_1(Xorshift *s) {
    return (real)((float)random_int(s, 10000000u) / 9999999.f);
}

/* ------------------------------------------------------------------ */
/* Primary sample vector (MLT_Sampler.h / MLT_Sampler.cpp)            */
/* ------------------------------------------------------------------ */
#define NSAMP_MAX 64
typedef struct {
    real value[NSAMP_MAX], backup[NSAMP_MAX];
    int n, index;
    real mutation_size;
} Sampler;

static void sampler_init_serial(Sampler *s, int max_path_length, Xorshift *x) {
    s->n = (max_path_length + 2) * 3 + 2;
    s->index = 0;
    s->mutation_size = K(1);
    for (int i = 0; i < s->n; ++i) { real e = random_float_0_1(x); s->value[i] = e; s->backup[i] = e; }
}
static void sampler_init_uniform(Sampler *s, Xorshift *x) {
    for (int i = 0; i < s->n; ++i) { s->backup[i] = s->value[i]; s->value[i] = random_float_0_1(x); }
}
static void sampler_mutate(Sampler *s, Xorshift *x) {
    for (int i = 0; i < s->n; ++i) {
        s->backup[i] = s->value[i];
        real delta;
        real e = random_float_0_1(x);
#ifdef ORC_DOUBLE
        real step = pow(INV_E, K(1) / s->mutation_size);
#else
        real step = powf(INV_E, K(1) / s->mutation_size);
#endif
        if (random_float_0_1(x) < K(0.5)) delta = e * step; else delta = -e * step;
        s->value[i] += delta;
        s->value[i] -= (real)floor(s->value[i]);
    }
}
static void sampler_reject(Sampler *s) { for (int i = 0; i < s->n; ++i) s->value[i] = s->backup[i]; }
static void sampler_plain(Sampler *s, const PathKey *k) {
    s->index = 0;
    int first = 0;
    if (k->stream == STREAM_LIGHT) {
        for (int d = 0; d < 5; ++d) s->value[d] = path_u(k, d);
        first = 5;
    }
    for (int b = 0; b < 8; ++b) {
        uint32_t ctr[4] = {(uint32_t)bounce_block((int)k->stream, b), k->path, k->iter, k->stream}, out[4];
        orc_philox(ctr, k->key, out);
        for (int j = 0; j < 3; ++j) s->value[first + 3 * b + j] = u24(out[j]);
    }
    s->n = first + 24;
}
static real get1d(Sampler *s) {
    real v = s->index < s->n ? s->value[s->index] : K(0);
    s->index++;
    return v;
}

/* Source of the draws that the reference takes from its serial stream:
 * Russian roulette and camera jitter. */
typedef struct { int serial; Xorshift *xs; PathKey key; const uint32_t *words; } Draws;
static real draw_rr(Draws *d, int path_length) {
    if (d->words) return u24(d->words[4 * bounce_block(STREAM_LIGHT, path_length) + 3]);
    return d->serial ? random_float_0_1(d->xs) : path_u(&d->key, 4 * bounce_block((int)d->key.stream, path_length) + 3);
}

/* ------------------------------------------------------------------ */
/* Frame (Frame.h:12-29) and sampling (Sampling.h:15-46)               */
/* ------------------------------------------------------------------ */
typedef struct { V x, y, z; } Frame;
static Frame frame_from_z(V z) {
    Frame f;
    f.z = z;
    V tmpx = (fabs(z.x) > K(0.99)) ? v3(0, 1, 0) : v3(1, 0, 0);
    f.y = vnormalize(vcross(z, tmpx));
    f.x = vcross(f.y, z);
    return f;
}
static inline V to_world(const Frame *f, V a) {
    return vadd(vadd(vscale(f->x, a.x), vscale(f->y, a.y)), vscale(f->z, a.z));
}
static inline V to_local(const Frame *f, V a) { return v3(vdot(a, f->x), vdot(a, f->y), vdot(a, f->z)); }

static void sample_uniform_disk(real ex, real ey, real *dx, real *dy) {
    real ux = ex * K(2) - K(1), uy = ey * K(2) - K(1);
    if (ux == K(0) && uy == K(0)) { *dx = 0; *dy = 0; return; }
    real theta, r;
    if (fabs(ux) > fabs(uy)) { r = ux; theta = PI_OVER_4 * (uy / ux); }
    else { r = uy; theta = PI_OVER_2 - PI_OVER_4 * (ux / uy); }
    *dx = rcos(theta) * r;
    *dy = rsin(theta) * r;
}
static V sample_cos_hemisphere(real ex, real ey) {
    real dx, dy;
    sample_uniform_disk(ex, ey, &dx, &dy);
    real z = (real)sqrt(smax(K(0), K(1) - dx * dx - dy * dy));
    return v3(dx, dy, z);
}

/* ------------------------------------------------------------------ */
/* Materials and BSDFs (materials.h:59-146, BXDF.cpp, BXDF.h:461-505)  */
/* ------------------------------------------------------------------ */
enum { EV_REFLECTION = 1, EV_TRANSMISSION = 2, EV_DIFFUSE = 4, EV_GLOSSY = 8, EV_SPECULAR = 16 };
/* BK_RGLASS: glass of roughness >= 0.01 as a GGX microfacet dielectric
 * (option compat_glass = 0; the reference's glass is always delta,
 * materials.h:104).  Walter, Marschner, Li, Torrance 2007, with the
 * reference's GGX sampling, D and height-correlated G2 (BXDF.cpp:286-311). */
enum { BK_SPEC_REFL, BK_FRESNEL_SPEC, BK_FMG, BK_FSMG, BK_MRG, BK_RGLASS };

typedef struct { char type; V color; real ior, roughness, specular; int rough_glass; } Material;

typedef struct {
    int kind, radiance, valid;
    V R;
    real a2, s, ior, cont;
    V wo;
    Frame fr;
} Bsdf;

/* Material conversion (Scene.cpp:45-68 and the constructors of materials.h). */
static Material material_from_python(const unsigned char *pm) {
    Material m;
    float ior, rough, spec, col[3];
    memcpy(&ior, pm + 4, 4);
    memcpy(&rough, pm + 8, 4);
    memcpy(&spec, pm + 12, 4);
    memcpy(col, pm + 16, 12);
    m.type = (char)pm[0];
    m.rough_glass = 0;
    m.color = v3(col[0], col[1], col[2]);
    m.ior = ior;
    if (m.type == 'P') { m.roughness = (real)rough * (real)rough; m.specular = (real)spec * K(0.08f); }
    else if (m.type == 'G') { m.roughness = (real)rough; m.specular = 0; }
    else if (m.type == 'M') { m.roughness = rough; m.specular = 0; }
    else { m.type = 'P'; m.color = v3(1, 1, 1); m.roughness = 1; m.specular = K(0.0) * K(0.08f); }
    return m;
}

static real cont_of(V c) { return smax(K(0.05f), smax(smax(c.x, c.y), c.z)); }

static void bsdf_init(Bsdf *b, const Material *m, V shade_n, V dir, int radiance) {
    b->radiance = radiance;
    b->a2 = 0; b->s = 0; b->ior = 1;
    if (m->type == 'P') {
        b->R = m->color;
        b->s = m->specular;
        if (m->roughness < K(0.01f)) { b->kind = BK_FSMG; b->cont = 1; }
        else { b->kind = BK_FMG; b->a2 = m->roughness * m->roughness; b->cont = 1; }
    } else if (m->type == 'G') {
        b->kind = m->rough_glass ? BK_RGLASS : BK_FRESNEL_SPEC; b->R = m->color; b->ior = m->ior;
        b->cont = cont_of(m->color);
        if (m->rough_glass) b->a2 = m->roughness * m->roughness;
    } else {
        b->R = m->color;
        if (m->roughness < K(0.01f)) { b->kind = BK_SPEC_REFL; }
        else { b->kind = BK_MRG; b->a2 = m->roughness * m->roughness; }
        b->cont = cont_of(m->color);
    }
    b->fr = frame_from_z(shade_n);
    b->wo = to_local(&b->fr, vneg(dir));
    b->valid = !(fabs(b->wo.z) < K(1e-6f));
}

static int bsdf_is_delta(const Bsdf *b) { return b->kind == BK_SPEC_REFL || b->kind == BK_FRESNEL_SPEC; }

static real fresnel_dielectric_eta(real cos_theta, real eta) {
    real g = eta * eta - K(1) + cos_theta * cos_theta;
    if (g > K(0)) {
        g = (real)sqrt(g);
        real A = (g - cos_theta) / (g + cos_theta);
        real B = (cos_theta * (g + cos_theta) - K(1)) / (cos_theta * (g - cos_theta) + K(1));
        return K(0.5) * A * A * (K(1) + B * B);
    }
    return K(1);
}
static real fresnel_dielectric_reflectance(real cos_theta, real reflectance) {
    if (reflectance >= K(1)) return K(1);
    real n = (K(1) + (real)sqrt(reflectance)) / (K(1) - (real)sqrt(reflectance));
    real g = n * n - K(1) + cos_theta * cos_theta;
    if (g > K(0)) {
        g = (real)sqrt(g);
        real A = (g - cos_theta) / (g + cos_theta);
        real B = (cos_theta * (g + cos_theta) - K(1)) / (cos_theta * (g - cos_theta) + K(1));
        return K(0.5) * A * A * (K(1) + B * B);
    }
    return K(1);
}
static real ggx_D(real ndoth, real a2) {
    real den = ndoth * ndoth * (a2 - K(1)) + K(1);
    return a2 / (PI * den * den);
}
static real ggx_G2(real ndotw, real ndotl, real a2) {
    real A = ndotw * (real)sqrt(a2 + (K(1) - a2) * ndotl * ndotl);
    real B = ndotl * (real)sqrt(a2 + (K(1) - a2) * ndotw * ndotw);
    return K(2) * ndotl * ndotw / (A + B);
}
static V ggx_half(real a2, real e1, real e2) {
    real theta = racos((real)sqrt((K(1) - e1) / (e1 * (a2 - K(1)) + K(1))));
    real phi = e2 * PI2;
    real st = rsin(theta);
    return v3(st * rcos(phi), st * rsin(phi), rcos(theta));
}
static int same_side(const Bsdf *b, V wi, V gn) {
    return !(vdot(wi, gn) * vdot(to_world(&b->fr, b->wo), gn) <= K(0));
}

typedef struct { V wi; real fwd, rev, cos_out, cont; int event; } Sample;

/* BxDF::sample for each class; returns the factor, zero on failure. */
static V bsdf_sample(const Bsdf *b, V gn, real u0, real u1, real u2, Sample *o) {
    const V zero = vs(0);
    V wo = b->wo;
    switch (b->kind) {
    case BK_SPEC_REFL: { /* BXDF.cpp:25-53 */
        o->event = EV_REFLECTION | EV_GLOSSY | EV_SPECULAR;
        V wil = v3(-wo.x, -wo.y, wo.z);
        o->cos_out = (real)fabs(wil.z);
        if (o->cos_out < K(1e-6f)) return zero;
        o->wi = to_world(&b->fr, wil);
        if (!same_side(b, o->wi, gn)) return zero;
        o->fwd = 1; o->rev = 1; o->cont = b->cont;
        return b->R;
    }
    case BK_FRESNEL_SPEC: { /* BXDF.cpp:136-205 */
        real eta = b->ior;
        if (wo.z > 0) eta = K(1) / b->ior;
        real f = fresnel_dielectric_eta((real)fabs(wo.z), K(1) / eta);
        if (u0 < f) {
            o->event = EV_GLOSSY | EV_SPECULAR | EV_REFLECTION;
            V wil = v3(-wo.x, -wo.y, wo.z);
            o->cos_out = (real)fabs(wil.z);
            if (o->cos_out < K(1e-6f)) return zero;
            o->wi = to_world(&b->fr, wil);
            if (!same_side(b, o->wi, gn)) return zero;
            o->fwd = f; o->rev = f; o->cont = b->cont;
            return b->R;
        }
        o->event = EV_TRANSMISSION | EV_GLOSSY | EV_SPECULAR;
        V wil;
        if (!vrefract(wo, wo.z > 0 ? v3(0, 0, 1) : v3(0, 0, -1), eta, &wil)) return zero;
        o->cos_out = (real)fabs(wil.z);
        if (o->cos_out < K(1e-6f)) return zero;
        o->wi = to_world(&b->fr, wil);
        if (vdot(o->wi, gn) * vdot(to_world(&b->fr, wo), gn) >= K(0)) return zero;
        o->fwd = K(1) - f; o->rev = K(1) - f; o->cont = b->cont;
        if (b->radiance) return vscale(vscale(b->R, eta), eta);
        return b->R;
    }
    case BK_MRG: { /* BXDF.cpp:302-337 */
        o->event = EV_GLOSSY | EV_REFLECTION;
        V h = ggx_half(b->a2, u1, u2);
        if (wo.z < 0) h.z = -h.z;
        V wil = vreflect(vneg(wo), h);
        o->cos_out = (real)fabs(wil.z);
        if (o->cos_out < K(1e-6f)) return zero;
        o->wi = to_world(&b->fr, wil);
        if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        o->fwd = o->rev = d * chn * K(0.25f) / vabsdot(h, wo);
        o->cont = b->cont;
        real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
        return vdivs(vscale(vscale(b->R, g2), vabsdot(wo, h)), (real)fabs(wo.z) * chn);
    }
    case BK_RGLASS: {
        /* a GGX microfacet normal on wo's side; the Fresnel term of wo at the
         * microfacet picks reflection or refraction through it; weight
         * G2 (wo.m) / (|wo.n| |m.n|), times eta^2 for refraction in radiance
         * mode (Fresnel_Specular's weights as the roughness goes to zero) */
        V m = ggx_half(b->a2, u1, u2);
        if (wo.z < 0) m.z = -m.z;
        real wom = vdot(wo, m);
        if (!(wom > K(0))) return zero;
        real eta = b->ior;
        if (wo.z > 0) eta = K(1) / b->ior;
        real f = fresnel_dielectric_eta(wom, K(1) / eta);
        real mz = (real)fabs(m.z);
        real d = ggx_D(mz, b->a2);
        o->cont = b->cont;
        if (u0 < f) {
            o->event = EV_GLOSSY | EV_REFLECTION;
            V wil = vreflect(vneg(wo), m);
            o->cos_out = (real)fabs(wil.z);
            if (o->cos_out < K(1e-6f)) return zero;
            o->wi = to_world(&b->fr, wil);
            if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
            o->fwd = o->rev = ((f * d) * mz) * K(0.25f) / wom;
            real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
            return vscale(vscale(b->R, g2), wom / ((real)fabs(wo.z) * mz));
        }
        o->event = EV_GLOSSY | EV_TRANSMISSION;
        V wil;
        if (!vrefract(wo, m, eta, &wil)) return zero;
        o->cos_out = (real)fabs(wil.z);
        if (o->cos_out < K(1e-6f) || wil.z * wo.z >= K(0)) return zero;
        o->wi = to_world(&b->fr, wil);
        if (vdot(o->wi, gn) * vdot(to_world(&b->fr, wo), gn) >= K(0)) return zero;
        /* pdfs with the Jacobian of the refracted direction (Walter et al. eq. 17) */
        real wim = vdot(wil, m);
        real sq = wim + eta * wom;
        real den = sq * sq;
        real t = ((K(1) - f) * d) * mz;
        o->fwd = (t * (real)fabs(wim)) / den;
        o->rev = (t * ((eta * eta) * wom)) / den;
        real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
        V w = vscale(vscale(b->R, g2), wom / ((real)fabs(wo.z) * mz));
        if (b->radiance) w = vscale(vscale(w, eta), eta);
        return w;
    }
    case BK_FMG: { /* BXDF.cpp:384-489 */
        real f = fresnel_dielectric_reflectance((real)fabs(wo.z), b->s);
        if (u0 < f) {
            o->event = EV_GLOSSY | EV_REFLECTION;
            V h = ggx_half(b->a2, u1, u2);
            if (wo.z < 0) h.z = -h.z;
            V wil = vreflect(vneg(wo), h);
            o->cos_out = (real)fabs(wil.z);
            if (o->cos_out < K(1e-6f)) return zero;
            o->wi = to_world(&b->fr, wil);
            if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
            real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
            o->cont = smax(K(0.05f), f);
            real fd = (real)fabs(wil.z) * INV_PI, rd = (real)fabs(wo.z) * INV_PI;
            real chn = (real)fabs(h.z);
            real d = ggx_D(chn, b->a2);
            real sp = d * chn * K(0.25f) / vabsdot(wo, h);
            if (f > K(1)) { o->fwd = o->rev = sp; }
            else { o->fwd = fd * (K(1) - f) + sp * f; o->rev = rd * (K(1) - f) + sp * f; }
            return vs(g2 * vabsdot(wo, h) / ((real)fabs(wo.z) * chn));
        }
        o->event = EV_DIFFUSE;
        V wil = sample_cos_hemisphere(u1, u2);
        o->cos_out = wil.z;
        if (o->cos_out < K(1e-6f)) return zero;
        if (wo.z < 0) wil.z = -wil.z;
        o->wi = to_world(&b->fr, wil);
        if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
        o->cont = smax(K(0.05f), vmaxc(vscale(b->R, K(1) - f)));
        real fd = (real)fabs(wil.z) * INV_PI, rd = (real)fabs(wo.z) * INV_PI;
        V h = vnormalize(vadd(wo, wil));
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        real sp = d * chn * K(0.25f) / vabsdot(wo, h);
        if (f > K(1)) { o->fwd = o->rev = sp; }
        else { o->fwd = fd * (K(1) - f) + sp * f; o->rev = rd * (K(1) - f) + sp * f; }
        return b->R;
    }
    default: { /* BK_FSMG, BXDF.cpp:568-632 */
        real f = fresnel_dielectric_reflectance((real)fabs(wo.z), b->s);
        if (u0 < f) {
            o->event = EV_SPECULAR | EV_REFLECTION | EV_GLOSSY;
            V wil = v3(-wo.x, -wo.y, wo.z);
            o->cos_out = (real)fabs(wil.z);
            if (o->cos_out < K(1e-6f)) return zero;
            o->wi = to_world(&b->fr, wil);
            if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
            o->fwd = f; o->rev = f;
            o->cont = smax(K(0.05f), f);
            return vs(1);
        }
        o->event = EV_DIFFUSE;
        V wil = sample_cos_hemisphere(u1, u2);
        o->cos_out = wil.z;
        if (o->cos_out < K(1e-6f)) return zero;
        if (wo.z < 0) wil.z = -wil.z;
        o->wi = to_world(&b->fr, wil);
        if (!same_side(b, o->wi, gn) || wil.z * wo.z <= K(0)) return zero;
        o->fwd = (real)fabs(wil.z) * INV_PI;
        o->rev = (real)fabs(wo.z) * INV_PI;
        o->cont = smax(K(0.05f), vmaxc(vscale(b->R, K(1) - f)));
        return b->R;
    }
    }
}

/* The rough dielectric's half vector for a pair (wo, wi): wo + wi for a
 * reflection, -(wi + eta wo) for a refraction (the generalised half vector,
 * Walter et al. eq. 16), turned to wo's side; with its cosines, Fresnel term
 * and D. */
typedef struct { int refl; real eta, wom, wim, f, mz, d; } RoughPair;
static RoughPair rough_pair(const Bsdf *b, V wil) {
    RoughPair r;
    V wo = b->wo;
    r.eta = b->ior;
    if (wo.z > 0) r.eta = K(1) / b->ior;
    r.refl = wil.z * wo.z > K(0);
    V mh = r.refl ? vadd(wo, wil) : vneg(vadd(wil, vscale(wo, r.eta)));
    mh = vnormalize(mh);
    if (mh.z * wo.z < K(0)) mh = vneg(mh);
    r.wom = vdot(wo, mh);
    r.wim = vdot(wil, mh);
    r.f = fresnel_dielectric_eta(r.wom, K(1) / r.eta);
    r.mz = (real)fabs(mh.z);
    r.d = ggx_D(r.mz, b->a2);
    return r;
}

/* BxDF::evaluate */
static V bsdf_eval(const Bsdf *b, V wi, V gn) {
    const V zero = vs(0);
    V wo = b->wo;
    switch (b->kind) {
    case BK_SPEC_REFL:
    case BK_FRESNEL_SPEC:
        return zero;
    case BK_MRG: { /* BXDF.cpp:339-356 */
        if (!same_side(b, wi, gn)) return zero;
        V wil = to_local(&b->fr, wi);
        if (wil.z * wo.z <= K(0)) return zero;
        real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
        V h = vnormalize(vadd(wo, wil));
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        return vdivs(vscale(vscale(vscale(b->R, d), g2), K(0.25f)), (real)fabs(wo.z));
    }
    case BK_RGLASS: { /* f |cos wi| of the rough dielectric */
        V wil = to_local(&b->fr, wi);
        RoughPair r = rough_pair(b, wil);
        real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
        if (r.refl) {
            if (!same_side(b, wi, gn) || !(r.wom > K(0))) return zero;
            return vscale(b->R, ((r.f * r.d) * g2) * K(0.25f) / (real)fabs(wo.z));
        }
        real sq = r.wim + r.eta * r.wom;
        real den = sq * sq;
        if (!(vdot(wi, gn) * vdot(to_world(&b->fr, wo), gn) < K(0)) || !(r.wom > K(0)) || !(r.wim < K(0)))
            return zero;
        /* the eta^2 of radiance transport, as the sampler's weight (the merge evaluates in radiance mode) */
        real v = (((K(1) - r.f) * r.d) * g2) * (r.wom * (real)fabs(r.wim));
        if (b->radiance) v = v * (r.eta * r.eta);
        v = v / ((real)fabs(wo.z) * den);
        return vscale(b->R, v);
    }
    case BK_FMG: { /* BXDF.cpp:491-519 */
        if (!same_side(b, wi, gn)) return zero;
        real f = fresnel_dielectric_reflectance((real)fabs(wo.z), b->s);
        V wil = to_local(&b->fr, wi);
        if (wil.z * wo.z <= K(0)) return zero;
        V h = vnormalize(vadd(wil, wo));
        V diffuse = vscale(b->R, INV_PI);
        real g2 = ggx_G2((real)fabs(wo.z), (real)fabs(wil.z), b->a2);
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        real spec = f * d * g2 * K(0.25f) / (real)fabs(wo.z);
        return v3(diffuse.x + spec, diffuse.y + spec, diffuse.z + spec);
    }
    default: { /* BK_FSMG, BXDF.cpp:634-648 */
        if (!same_side(b, wi, gn)) return zero;
        if (to_local(&b->fr, wi).z * wo.z <= K(0)) return zero;
        return vscale(b->R, INV_PI);
    }
    }
}

/* BxDF::pdf */
static void bsdf_pdf(const Bsdf *b, V wi, real *fwd, real *rev) {
    V wo = b->wo;
    switch (b->kind) {
    case BK_SPEC_REFL:
    case BK_FRESNEL_SPEC:
        *fwd = 0; *rev = 0; return;
    case BK_MRG: { /* BXDF.cpp:358-367 */
        V wil = to_local(&b->fr, wi);
        V h = vnormalize(vadd(wo, wil));
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        *fwd = *rev = d * chn * K(0.25f) / vabsdot(h, wo);
        return;
    }
    case BK_RGLASS: {
        RoughPair r = rough_pair(b, to_local(&b->fr, wi));
        /* directions no microfacet normal on wo's side sends wo to have density 0 */
        if (!(r.wom > K(0)) || (!r.refl && !(r.wim < K(0)))) { *fwd = *rev = 0; return; }
        if (r.refl) {
            *fwd = *rev = ((r.f * r.d) * r.mz) * K(0.25f) / r.wom;
            return;
        }
        real sq = r.wim + r.eta * r.wom;
        real den = sq * sq;
        real t = ((K(1) - r.f) * r.d) * r.mz;
        *fwd = (t * (real)fabs(r.wim)) / den;
        *rev = (t * ((r.eta * r.eta) * r.wom)) / den;
        return;
    }
    case BK_FMG: { /* BXDF.cpp:521-548 */
        V wil = to_local(&b->fr, wi);
        real fd = (real)fabs(wil.z) * INV_PI, rd = (real)fabs(wo.z) * INV_PI;
        V h = vnormalize(vadd(wo, wil));
        real chn = (real)fabs(h.z);
        real d = ggx_D(chn, b->a2);
        real sp = d * chn * K(0.25f) / vabsdot(h, wo);
        real f = fresnel_dielectric_reflectance((real)fabs(wo.z), b->s);
        if (f > K(1)) { *fwd = *rev = sp; }
        else { *fwd = fd * (K(1) - f) + sp * f; *rev = rd * (K(1) - f) + sp * f; }
        return;
    }
    default: { /* BK_FSMG, BXDF.cpp:650-667 */
        real f = fresnel_dielectric_reflectance((real)fabs(wo.z), b->s);
        if (f > K(1)) { *fwd = *rev = 0; return; }
        V wil = to_local(&b->fr, wi);
        *fwd = (real)fabs(wil.z) * INV_PI * (K(1) - f);
        *rev = (real)fabs(wo.z) * INV_PI * (K(1) - f);
        return;
    }
    }
}

/* ------------------------------------------------------------------ */
/* Geometry: boxes, triangles, meshes, two-level BVH                   */
/* ------------------------------------------------------------------ */
typedef struct { V mn, mx; } Box;
static Box box_empty(void) { Box b = {vs(INF), vs(-INF)}; return b; }
static Box box_union(Box a, Box b) {
    Box r = {v3(smin(a.mn.x, b.mn.x), smin(a.mn.y, b.mn.y), smin(a.mn.z, b.mn.z)),
             v3(smax(a.mx.x, b.mx.x), smax(a.mx.y, b.mx.y), smax(a.mx.z, b.mx.z))};
    return r;
}
static Box box_union_pt(Box a, V p) {
    Box r = {v3(smin(a.mn.x, p.x), smin(a.mn.y, p.y), smin(a.mn.z, p.z)),
             v3(smax(a.mx.x, p.x), smax(a.mx.y, p.y), smax(a.mx.z, p.z))};
    return r;
}
static real box_area(Box b) {
    V d = vsub(b.mx, b.mn);
    return K(2) * (d.x * d.y + d.x * d.z + d.y * d.z);
}
static int box_max_extent(Box b) {
    V d = vsub(b.mx, b.mn);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}
static V box_offset(Box b, V p) {
    V o = vsub(p, b.mn);
    o.x /= b.mx.x - b.mn.x;
    o.y /= b.mx.y - b.mn.y;
    o.z /= b.mx.z - b.mn.z;
    return o;
}

typedef struct { V o, d; } Ray;

typedef struct { long long box_tests, tri_tests; } Counters;

/* aabb::hit (aabb.cpp:6-42) */
static int box_hit(const Box *b, const Ray *r, real tmin, real tmax, Counters *c) {
    c->box_tests++;
    real invD = K(1) / r->d.x;
    real t0 = (b->mn.x - r->o.x) * invD, t1 = (b->mx.x - r->o.x) * invD;
    if (invD < 0) { real t = t0; t0 = t1; t1 = t; }
    tmin = t0 > tmin ? t0 : tmin;
    tmax = t1 < tmax ? t1 : tmax;
    if (tmax < tmin) return 0;
    invD = K(1) / r->d.y;
    t0 = (b->mn.y - r->o.y) * invD; t1 = (b->mx.y - r->o.y) * invD;
    if (invD < 0) { real t = t0; t0 = t1; t1 = t; }
    tmin = t0 > tmin ? t0 : tmin;
    tmax = t1 < tmax ? t1 : tmax;
    if (tmax < tmin) return 0;
    invD = K(1) / r->d.z;
    t0 = (b->mn.z - r->o.z) * invD; t1 = (b->mx.z - r->o.z) * invD;
    if (invD < 0) { real t = t0; t0 = t1; t1 = t; }
    tmin = t0 > tmin ? t0 : tmin;
    tmax = t1 < tmax ? t1 : tmax;
    if (tmax < tmin) return 0;
    return 1;
}

typedef struct {
    V gn, sn, pos, dir;
    real t, u, v;
    unsigned mat;
    int tri_global;
} Isect;

typedef struct { Box b; uint32_t u0, u1; } Node; /* 32-byte CacheBVHNode layout */

typedef struct {
    uint32_t ntris, nverts, tri_offset;
    V *wv, *wn;        /* world-space vertices and normalized world normals */
    uint32_t *tri;     /* v0 v1 v2 material */
    Node *nodes;
    uint32_t nnodes;
    int *tris_indices; /* 1-based, as BVH/BVH_mesh.cpp:231 */
    Box bbox;
    real cost;
} Mesh;

/* Triangle::hit (Triangle.cpp:20-88) on pre-transformed vertices: the
 * world transform of Transform.cpp:3-31 gives the same floats once or per test. */
static int tri_hit(const Mesh *m, uint32_t t, const Ray *r, real tmin, real tmax, Isect *h, Counters *c) {
    c->tri_tests++;
    const uint32_t *T = m->tri + 4 * t;
    V p0 = m->wv[T[0]], p1 = m->wv[T[1]], p2 = m->wv[T[2]];
    V e1 = vsub(p1, p0), e2 = vsub(p2, p0);
    V pvec = vcross(r->d, e2);
    real det = vdot(e1, pvec);
    if (fabs(det) < K(0.0000001f)) return 0;
    real inv = K(1) / det;
    V tvec = vsub(r->o, p0);
    real u = vdot(tvec, pvec) * inv;
    if (u < 0 || u > 1) return 0;
    V qvec = vcross(tvec, e1);
    real v = vdot(r->d, qvec) * inv;
    if (v < 0 || u + v > 1) return 0;
    real tt = vdot(e2, qvec) * inv;
    if (tt > tmin && tt < tmax) {
        h->t = tt; h->u = u; h->v = v;
        h->gn = vnormalize(vcross(e1, e2));
        h->sn = vadd(vadd(vscale(m->wn[T[0]], K(1) - u - v), vscale(m->wn[T[1]], u)), vscale(m->wn[T[2]], v));
        h->pos = vadd(r->o, vscale(r->d, tt));
        h->dir = r->d;
        h->mat = T[3];
        h->tri_global = (int)(m->tri_offset + t);
        return 1;
    }
    return 0;
}

/* ---- Mesh SAH build (BVH/BVH_mesh.cpp:28-155, 199-261, 286-304) ---- */
typedef struct { uint32_t index; Box b; V c; int cost; } Item;

typedef struct {
    Node *nodes; uint32_t nn, cap;
    int *list; uint32_t nl, lcap;
} Flat;

static uint32_t flat_node(Flat *f) {
    if (f->nn == f->cap) { f->cap = f->cap ? 2 * f->cap : 64; f->nodes = realloc(f->nodes, f->cap * sizeof(Node)); }
    return f->nn++;
}
static void flat_push(Flat *f, int v) {
    if (f->nl == f->lcap) { f->lcap = f->lcap ? 2 * f->lcap : 64; f->list = realloc(f->list, f->lcap * sizeof(int)); }
    f->list[f->nl++] = v;
}
static void make_leaf(Flat *f, uint32_t id, Item *it, int start, int end, Box bb) {
    f->nodes[id].b = bb;
    f->nodes[id].u0 = 0x80000000u | (uint32_t)(end - start);
    f->nodes[id].u1 = f->nl - 1; /* startIndex: position before the first entry (1-based list) */
    for (int i = start; i < end; ++i) flat_push(f, (int)it[i].index);
}
static int bucket_of(Box cb, V c, int dim) {
    int b = (int)(12 * vcomp(box_offset(cb, c), dim));
    if (b == 12) b = 11;
    return b;
}
/* std::partition as libstdc++ implements it for bidirectional iterators. */
static int partition_items(Item *it, int first, int last, Box cb, int dim, real min_bucket) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if ((real)bucket_of(cb, it[first].c, dim) <= min_bucket) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!((real)bucket_of(cb, it[last].c, dim) <= min_bucket)) --last; else break;
        }
        Item t = it[first]; it[first] = it[last]; it[last] = t;
        ++first;
    }
}

/* One recursive builder for both levels: world = 1 uses the mesh-cost SAH
 * of BVH/BVH_world.cpp:25-159, world = 0 the triangle SAH of BVH_mesh.cpp. */
static void build_rec(Flat *f, uint32_t id, Item *it, int start, int end, int world) {
    Box bb = box_empty();
    for (int i = start; i < end; ++i) bb = box_union(bb, it[i].b);
    int n = end - start;
    if (n == 1) { make_leaf(f, id, it, start, end, bb); return; }
    Box cb = box_empty();
    for (int i = start; i < end; ++i) cb = box_union_pt(cb, it[i].c);
    int dim = box_max_extent(cb);
    int mid = (start + end) / 2;
    if (vcomp(cb.mx, dim) == vcomp(cb.mn, dim)) { make_leaf(f, id, it, start, end, bb); return; }
    if (n == 2) {
        if (!(vcomp(it[start].c, dim) < vcomp(it[end - 1].c, dim))) {
            Item t = it[start]; it[start] = it[end - 1]; it[end - 1] = t;
        }
    } else {
        int count[12] = {0};
        Box bbox[12];
        real tcost[12];
        for (int i = 0; i < 12; ++i) { bbox[i] = box_empty(); tcost[i] = 0; }
        for (int i = start; i < end; ++i) {
            int b = bucket_of(cb, it[i].c, dim);
            count[b]++;
            bbox[b] = box_union(bbox[b], it[i].b);
            tcost[b] += (real)it[i].cost;
        }
        real cost[11];
        for (int i = 0; i < 11; ++i) {
            Box b0 = box_empty(), b1 = box_empty();
            int c0 = 0, c1 = 0, t0 = 0, t1 = 0;
            for (int j = 0; j <= i; ++j) { b0 = box_union(b0, bbox[j]); c0 += count[j]; t0 = (int)((real)t0 + tcost[j]); }
            for (int j = i + 1; j < 12; ++j) { b1 = box_union(b1, bbox[j]); c1 += count[j]; t1 = (int)((real)t1 + tcost[j]); }
            if (world)
                cost[i] = (real)(0.16874 + (double)((box_area(b0) * (real)t0 + box_area(b1) * (real)t1) / box_area(bb)));
            else
                cost[i] = (real)(0.16874 + (double)(((real)c0 * box_area(b0) + (real)c1 * box_area(b1)) / box_area(bb)));
        }
        real min_cost = cost[0], min_bucket = 0, leaf_cost = 0;
        if (world) {
            for (int i = 0; i < 11; ++i) {
                leaf_cost += tcost[i];
                if (cost[i] < min_cost) { min_cost = cost[i]; min_bucket = (real)i; }
            }
        } else {
            for (int i = 0; i < 11; ++i)
                if (cost[i] < min_cost) { min_cost = cost[i]; min_bucket = (real)i; }
            leaf_cost = (real)n;
        }
        if (min_cost < leaf_cost) mid = partition_items(it, start, end, cb, dim, min_bucket);
        else { make_leaf(f, id, it, start, end, bb); return; }
    }
    f->nodes[id].b = bb;
    uint32_t l = flat_node(f);
    build_rec(f, l, it, start, mid, world);
    uint32_t r = flat_node(f);
    build_rec(f, r, it, mid, end, world);
    f->nodes[id].u0 = l;
    f->nodes[id].u1 = r;
}

static real cost_of_traversing(const Node *nodes, uint32_t i) {
    const Node *n = &nodes[i];
    if (n->u0 & 0x80000000u) return (real)(n->u0 & 0x7fffffffu);
    real a = box_area(n->b), la = box_area(nodes[n->u0].b), ra = box_area(nodes[n->u1].b);
    return (real)(0.16874 + (double)((cost_of_traversing(nodes, n->u0) * la + cost_of_traversing(nodes, n->u1) * ra) / a));
}

/* Mesh::hit (mesh.cpp:70-104): prunes with the record's distance, not tmax. */
static int mesh_hit(const Mesh *m, const Ray *r, real tmin, Isect *rec, Counters *c) {
    uint32_t stack[64];
    int sp = 0, hit = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node *n = &m->nodes[stack[--sp]];
        if (box_hit(&n->b, r, tmin, rec->t, c)) {
            if (n->u0 & 0x80000000u) {
                /* CacheBVHNode_mesh::hit (BVH_mesh.cpp:265-283) */
                real closest = rec->t;
                Isect tmp;
                int end = (int)(n->u1 + (n->u0 & 0x7fffffffu) + 1);
                for (int i = (int)n->u1 + 1; i < end; ++i)
                    if (tri_hit(m, (uint32_t)m->tris_indices[i], r, tmin, closest, &tmp, c)) {
                        hit = 1; closest = tmp.t; *rec = tmp;
                    }
            } else {
                stack[sp++] = n->u1;
                stack[sp++] = n->u0;
            }
        }
    }
    return hit;
}

/* ------------------------------------------------------------------ */
/* Lights (Lights.cpp, Light_list.cpp)                                 */
/* ------------------------------------------------------------------ */
enum { L_AREA, L_SUN, L_POINT };
typedef struct {
    int kind;
    V color, pos, bl, h, v, normal;
    real power, radius;
    Frame fr;
} Light;

/* matrix_3x3(vec3 r) (matrix.cpp:77-89); the reference's unqualified cos/sin
 * on float arguments resolve to the double C functions on this platform. */
typedef struct { V i, j, k; } M3;
static M3 rot_matrix(V r) {
    double x = r.x, y = r.y, z = r.z;
    M3 m;
    m.i = v3((real)(cos(z) * cos(y)), (real)(sin(z) * cos(y)), (real)(-sin(y)));
    m.j = v3((real)(cos(z) * sin(y) * sin(x) - sin(z) * cos(x)), (real)(sin(z) * sin(y) * sin(x) + cos(z) * cos(x)),
             (real)(cos(y) * sin(x)));
    m.k = v3((real)(cos(z) * sin(y) * cos(x) + sin(z) * sin(x)), (real)(sin(z) * sin(y) * cos(x) - cos(z) * sin(x)),
             (real)(cos(y) * cos(x)));
    return m;
}
static V mul_vm(V v, M3 m) { return vadd(vadd(vscale(m.i, v.x), vscale(m.j, v.y)), vscale(m.k, v.z)); }
static M3 transpose3(M3 m) {
    M3 t;
    t.i = v3(m.i.x, m.j.x, m.k.x);
    t.j = v3(m.i.y, m.j.y, m.k.y);
    t.k = v3(m.i.z, m.j.z, m.k.z);
    return t;
}
/* rotate_vec with degrees (matrix.h:101-106) */
static V rotate_deg(V v, V p, V rot) {
    M3 m = rot_matrix(v3(rot.x * PI / K(180.f), rot.y * PI / K(180.f), rot.z * PI / K(180.f)));
    return vadd(mul_vm(vsub(v, p), m), p);
}
static V project_onto_plane(V point, V center, V normal) {
    V tp = vsub(point, center);
    return vadd(vsub(tp, vscale(normal, vdot(tp, normal) / vlen2(normal))), center);
}

/* ------------------------------------------------------------------ */
/* Scene                                                               */
/* ------------------------------------------------------------------ */
enum { GUIDE_BINS = 1024, GUIDE_MAX_LIGHTS = 8 };

typedef struct {
    int samples;
    real base_radius, radius_alpha;
    uint32_t nmeshes;
    Mesh *meshes;
    Node *wnodes;
    uint32_t nwnodes;
    int *mesh_indices;
    Box bbox;
    Material *mats;
    uint32_t nmats;
    unsigned char *py_lights;
    uint32_t nlights;
    Light *lights;
    real *cdf, integral;
    int W, H;
    V cam_o, cam_bl, cam_h, cam_v;
    int max_path_length;
    /* options */
    int rng_serial, compat_point, compat_pdf, compat_rows, camera_stride, gather_brute;
    int compat_glass; /* 0: glass of roughness >= 0.01 is the rough dielectric */
    uint64_t seed;
    long long photons;
    int prepared;
    real scene_radius;
    Xorshift xs;
    Sampler serial_smp; /* the iteration's MLT_Sampler, shared by both passes */
    Counters ctr;
    long long visible_paths;
    long long mlt_mutations, mlt_accepted; /* adaptive MLT light pass of the last iteration */
    long long mlt_chains;                  /* rng_serial = 2: chains of the engine's MLT light pass */
    int camera_qmc;                        /* stratified camera jitter and first-bounce choice */
    /* learned emission proposal: per-cell CDFs given by the caller */
    real *guide_cdf;            /* [slot][cell][GUIDE_BINS + 1] */
    int guide_cells, guide_slot[GUIDE_MAX_LIGHTS];
    real guide_alpha;
    uint32_t guide_words[4];    /* this path's {dim 3, dim 4, mixture, bin} words */
    int threads;                /* > 1: plain light and camera passes on this many OpenMP threads */
} Scene;

void *orc_create(int samples, float base_radius, float radius_reduction, const long long *meshes_pointers,
                 unsigned number_of_meshes, const float *mesh_matrices, const unsigned *nverts,
                 const unsigned *ntris, unsigned camera_x, unsigned camera_y, const float *cam15,
                 const void *lights, unsigned number_of_lights, const void *materials,
                 unsigned number_of_materials, const int *const *materials_indices) {
    Scene *s = (Scene *)calloc(1, sizeof(Scene));
    s->samples = samples;
    s->base_radius = base_radius;
    s->radius_alpha = radius_reduction;
    s->max_path_length = 6; /* main.cpp:112 */
    s->camera_stride = 1;
    s->mlt_chains = 262144;
    s->compat_point = 1; s->compat_pdf = 1; s->compat_rows = 1; s->compat_glass = 1;
    s->xs.a = 132141u;
    s->nmeshes = number_of_meshes;
    s->meshes = (Mesh *)calloc(number_of_meshes ? number_of_meshes : 1, sizeof(Mesh));
    uint32_t offset = 0;
    for (uint32_t i = 0; i < number_of_meshes; ++i) {
        /* Scene::init_meshes (Scene.cpp:14-43) + Mesh::Mesh (mesh.cpp:12-26) */
        const unsigned char *mb = (const unsigned char *)(intptr_t)meshes_pointers[i];
        const unsigned char *mpoly, *mloop, *mvert;
        memcpy(&mpoly, mb + 200, sizeof(void *));
        memcpy(&mloop, mb + 208, sizeof(void *));
        memcpy(&mvert, mb + 256, sizeof(void *));
        Mesh *m = &s->meshes[i];
        m->ntris = ntris[i];
        m->nverts = nverts[i];
        m->tri_offset = offset;
        offset += ntris[i];
        const float *M = mesh_matrices + 16 * i;
        /* Inverse (matrix.cpp:4-63) in double, then float */
        double a[4][4];
        for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) a[r][c] = M[4 * r + c];
        double A2323 = a[2][2] * a[3][3] - a[2][3] * a[3][2], A1323 = a[2][1] * a[3][3] - a[2][3] * a[3][1];
        double A1223 = a[2][1] * a[3][2] - a[2][2] * a[3][1], A0323 = a[2][0] * a[3][3] - a[2][3] * a[3][0];
        double A0223 = a[2][0] * a[3][2] - a[2][2] * a[3][0], A0123 = a[2][0] * a[3][1] - a[2][1] * a[3][0];
        double det = a[0][0] * (a[1][1] * A2323 - a[1][2] * A1323 + a[1][3] * A1223) -
                     a[0][1] * (a[1][0] * A2323 - a[1][2] * A0323 + a[1][3] * A0223) +
                     a[0][2] * (a[1][0] * A1323 - a[1][1] * A0323 + a[1][3] * A0123) -
                     a[0][3] * (a[1][0] * A1223 - a[1][1] * A0223 + a[1][2] * A0123);
        det = 1.0 / det;
        float inv[3][3];
        inv[0][0] = (float)(det * (a[1][1] * A2323 - a[1][2] * A1323 + a[1][3] * A1223));
        inv[0][1] = (float)(det * -(a[0][1] * A2323 - a[0][2] * A1323 + a[0][3] * A1223));
        inv[1][0] = (float)(det * -(a[1][0] * A2323 - a[1][2] * A0323 + a[1][3] * A0223));
        inv[1][1] = (float)(det * (a[0][0] * A2323 - a[0][2] * A0323 + a[0][3] * A0223));
        inv[2][0] = (float)(det * (a[1][0] * A1323 - a[1][1] * A0323 + a[1][3] * A0123));
        inv[2][1] = (float)(det * -(a[0][0] * A1323 - a[0][1] * A0323 + a[0][3] * A0123));
        {
            double A2313 = a[1][2] * a[3][3] - a[1][3] * a[3][2], A1313 = a[1][1] * a[3][3] - a[1][3] * a[3][1];
            double A1213 = a[1][1] * a[3][2] - a[1][2] * a[3][1], A0313 = a[1][0] * a[3][3] - a[1][3] * a[3][0];
            double A0213 = a[1][0] * a[3][2] - a[1][2] * a[3][0], A0113 = a[1][0] * a[3][1] - a[1][1] * a[3][0];
            inv[0][2] = (float)(det * (a[0][1] * A2313 - a[0][2] * A1313 + a[0][3] * A1213));
            inv[1][2] = (float)(det * -(a[0][0] * A2313 - a[0][2] * A0313 + a[0][3] * A0213));
            inv[2][2] = (float)(det * (a[0][0] * A1313 - a[0][1] * A0313 + a[0][3] * A0113));
        }
        m->wv = (V *)malloc((m->nverts ? m->nverts : 1) * sizeof(V));
        m->wn = (V *)malloc((m->nverts ? m->nverts : 1) * sizeof(V));
        for (uint32_t v = 0; v < m->nverts; ++v) {
            float co[3];
            int16_t no[3];
            memcpy(co, mvert + 20 * (size_t)v, 12);
            memcpy(no, mvert + 20 * (size_t)v + 12, 6);
            /* Transform::ApplyToPoint (Transform.cpp:3-11) */
            real x = co[0], y = co[1], z = co[2];
            m->wv[v] = v3((real)M[0] * x + (real)M[1] * y + (real)M[2] * z + (real)M[3],
                          (real)M[4] * x + (real)M[5] * y + (real)M[6] * z + (real)M[7],
                          (real)M[8] * x + (real)M[9] * y + (real)M[10] * z + (real)M[11]);
            /* normal_short_to_float (Triangle.cpp:11-16) + ApplyToNormal (Transform.cpp:23-31) */
            real nx = (real)no[0] * (real)(1.0f / 32767.0f), ny = (real)no[1] * (real)(1.0f / 32767.0f),
                 nz = (real)no[2] * (real)(1.0f / 32767.0f);
            V wn = v3((real)inv[0][0] * nx + (real)inv[1][0] * ny + (real)inv[2][0] * nz,
                      (real)inv[0][1] * nx + (real)inv[1][1] * ny + (real)inv[2][1] * nz,
                      (real)inv[0][2] * nx + (real)inv[1][2] * ny + (real)inv[2][2] * nz);
            m->wn[v] = vnormalize(wn);
        }
        m->tri = (uint32_t *)malloc((m->ntris ? m->ntris : 1) * 4 * sizeof(uint32_t));
        const int *mi = materials_indices[i];
        for (uint32_t t = 0; t < m->ntris; ++t) {
            int32_t loopstart;
            int16_t mat_nr;
            memcpy(&loopstart, mpoly + 12 * (size_t)t, 4);
            memcpy(&mat_nr, mpoly + 12 * (size_t)t + 8, 2);
            for (int k = 0; k < 3; ++k) memcpy(&m->tri[4 * t + k], mloop + 8 * ((size_t)loopstart + k), 4);
            m->tri[4 * t + 3] = (uint32_t)mi[mat_nr];
        }
    }
    /* Camera (Camera.h:21-27), corners TR, BR, BL, TL */
    s->W = (int)camera_x;
    s->H = (int)camera_y;
    s->cam_o = v3(cam15[0], cam15[1], cam15[2]);
    V br = v3(cam15[6], cam15[7], cam15[8]), bl = v3(cam15[9], cam15[10], cam15[11]),
      tl = v3(cam15[12], cam15[13], cam15[14]);
    s->cam_bl = bl;
    s->cam_h = vsub(br, bl);
    s->cam_v = vsub(tl, bl);
    /* lights are converted after the BVH build (main.cpp:163) */
    s->nlights = number_of_lights;
    s->py_lights = (unsigned char *)malloc(64 * (number_of_lights ? number_of_lights : 1));
    memcpy(s->py_lights, lights, 64 * (size_t)number_of_lights);
    s->nmats = number_of_materials;
    s->mats = (Material *)malloc(sizeof(Material) * (number_of_materials ? number_of_materials : 1));
    for (uint32_t i = 0; i < number_of_materials; ++i)
        s->mats[i] = material_from_python((const unsigned char *)materials + 28 * (size_t)i);
    return s;
}

void orc_free(void *sp) {
    Scene *s = (Scene *)sp;
    if (!s) return;
    for (uint32_t i = 0; i < s->nmeshes; ++i) {
        free(s->meshes[i].wv); free(s->meshes[i].wn); free(s->meshes[i].tri);
        free(s->meshes[i].nodes); free(s->meshes[i].tris_indices);
    }
    free(s->meshes); free(s->wnodes); free(s->mesh_indices); free(s->mats);
    free(s->py_lights); free(s->lights); free(s->cdf); free(s->guide_cdf);
    free(s);
}

int orc_set_option(void *sp, const char *key, double v) {
    Scene *s = (Scene *)sp;
    if (!strcmp(key, "rng_serial")) s->rng_serial = (int)v;
    else if (!strcmp(key, "mlt_chains")) s->mlt_chains = (long long)v;
    else if (!strcmp(key, "camera_qmc")) s->camera_qmc = (int)v;
    else if (!strcmp(key, "seed")) s->seed = (uint64_t)v;
    else if (!strcmp(key, "photons")) s->photons = (long long)v;
    else if (!strcmp(key, "compat_point_light")) s->compat_point = (int)v;
    else if (!strcmp(key, "compat_light_pdf")) s->compat_pdf = (int)v;
    else if (!strcmp(key, "compat_glass")) s->compat_glass = (int)v;
    else if (!strcmp(key, "compat_rows")) s->compat_rows = (int)v;
    else if (!strcmp(key, "libm")) g_libm = (int)v;
    else if (!strcmp(key, "camera_stride")) s->camera_stride = v < 1 ? 1 : (int)v;
    else if (!strcmp(key, "max_path_length")) s->max_path_length = (int)v;
    else if (!strcmp(key, "gather")) s->gather_brute = (int)v;
    else if (!strcmp(key, "threads")) s->threads = v < 1 ? 1 : (int)v;
    else return -1;
    return 0;
}

double orc_get(void *sp, const char *key) {
    Scene *s = (Scene *)sp;
    if (!strcmp(key, "scene_radius")) return s->scene_radius;
    if (!strcmp(key, "base_radius")) return s->base_radius;
    if (!strcmp(key, "visible_paths")) return (double)s->visible_paths;
    if (!strcmp(key, "mlt_mutations")) return (double)s->mlt_mutations;
    if (!strcmp(key, "mlt_accepted")) return (double)s->mlt_accepted;
    if (!strcmp(key, "bbox_min_x")) return s->bbox.mn.x;
    if (!strcmp(key, "bbox_min_y")) return s->bbox.mn.y;
    if (!strcmp(key, "bbox_min_z")) return s->bbox.mn.z;
    if (!strcmp(key, "bbox_max_x")) return s->bbox.mx.x;
    if (!strcmp(key, "bbox_max_y")) return s->bbox.mx.y;
    if (!strcmp(key, "bbox_max_z")) return s->bbox.mx.z;
    if (!strcmp(key, "world_nodes")) return s->nwnodes;
    return NAN;
}

static void init_lights(Scene *s) {
    /* Scene::init_lights (Scene.cpp:70-108) and the light constructors. */
    s->lights = (Light *)calloc(s->nlights ? s->nlights : 1, sizeof(Light));
    for (uint32_t i = 0; i < s->nlights; ++i) {
        const unsigned char *p = s->py_lights + 64 * (size_t)i;
        float col[3], pos[3], w, h, rot[3], power, radius;
        memcpy(col, p + 4, 12); memcpy(pos, p + 16, 12); memcpy(&w, p + 28, 4); memcpy(&h, p + 32, 4);
        memcpy(rot, p + 40, 12); memcpy(&power, p + 52, 4); memcpy(&radius, p + 60, 4);
        Light *L = &s->lights[i];
        V c = v3(col[0], col[1], col[2]);
        L->power = power;
        L->pos = v3(pos[0], pos[1], pos[2]);
        L->radius = radius;
        if (p[0] == 'P') {
            L->kind = L_POINT;
            L->color = c;
        } else if (p[0] == 'S') {
            /* Sun_Light (Lights.cpp:58-119); rotation in radians */
            L->kind = L_SUN;
            L->color = vscale(c, power);
            V r = v3(rot[0], rot[1], rot[2]);
            M3 m = rot_matrix(r);
            L->normal = mul_vm(vsub(v3(0, 0, -1), vs(0)), m);
            L->normal = vadd(L->normal, vs(0));
            M3 inv = transpose3(m);
            Box bb = s->bbox;
            L->pos = vadd(vscale(vneg(L->normal), vlen(vsub(bb.mx, bb.mn))), vscale(vadd(bb.mx, bb.mn), K(0.5f)));
            V pts[8] = {bb.mn, v3(bb.mx.x, bb.mn.y, bb.mn.z), v3(bb.mn.x, bb.mx.y, bb.mn.z),
                        v3(bb.mx.x, bb.mx.y, bb.mn.z), v3(bb.mn.x, bb.mn.y, bb.mx.z),
                        v3(bb.mx.x, bb.mn.y, bb.mx.z), v3(bb.mn.x, bb.mx.y, bb.mx.z), bb.mx};
            V mx = v3(-INF, -INF, 0), mn = v3(INF, INF, 0);
            for (int k = 0; k < 8; ++k) {
                V loc = mul_vm(vsub(project_onto_plane(pts[k], L->pos, L->normal), L->pos), inv);
                if (loc.x > mx.x) mx.x = loc.x;
                if (loc.y > mx.y) mx.y = loc.y;
                if (loc.x < mn.x) mn.x = loc.x;
                if (loc.y < mn.y) mn.y = loc.y;
            }
            V brl = v3(mx.x, mn.y, 0), tll = v3(mn.x, mx.y, 0);
            L->bl = vadd(mul_vm(mn, m), L->pos);
            V brc = vadd(mul_vm(brl, m), L->pos), tlc = vadd(mul_vm(tll, m), L->pos);
            L->h = vsub(brc, L->bl);
            L->v = vsub(tlc, L->bl);
        } else {
            /* Area_Light (Lights.cpp:18-34); rotation in degrees */
            L->kind = L_AREA;
            L->color = vscale(c, power);
            V rr = v3(rot[0], rot[1], rot[2]);
            V bl = vsub(L->pos, v3((real)w / K(2.f), (real)h / K(2.f), K(0.f)));
            V hh = v3(w, 0, 0), vv = v3(0, h, 0), n = v3(0, 0, -1);
            V tl = vadd(bl, vv), br = vadd(bl, hh);
            bl = rotate_deg(bl, L->pos, rr);
            tl = rotate_deg(tl, L->pos, rr);
            br = rotate_deg(br, L->pos, rr);
            n = rotate_deg(n, v3(0, 0, 0), rr);
            L->bl = bl;
            L->h = vsub(br, bl);
            L->v = vsub(tl, bl);
            L->normal = n;
            L->fr = frame_from_z(n);
        }
    }
    /* Lights_list::update_pdf_distribution (Light_list.cpp:32-44) */
    s->cdf = (real *)calloc(s->nlights + 1, sizeof(real));
    s->cdf[0] = 0;
    for (uint32_t i = 1; i < s->nlights + 1; ++i) s->cdf[i] = s->cdf[i - 1] + s->lights[i - 1].power;
    s->integral = s->cdf[s->nlights];
    for (uint32_t i = 1; i < s->nlights + 1; ++i) s->cdf[i] /= s->integral;
}

int orc_prepare(void *sp) {
    Scene *s = (Scene *)sp;
    for (uint32_t i = 0; i < s->nmats; ++i)
        s->mats[i].rough_glass = s->mats[i].type == 'G' && !s->compat_glass && s->mats[i].roughness >= K(0.01f);
    /* Scene::update_bvh (Scene.cpp:110-152) */
    Item *witems = (Item *)malloc(sizeof(Item) * (s->nmeshes ? s->nmeshes : 1));
    for (uint32_t i = 0; i < s->nmeshes; ++i) {
        Mesh *m = &s->meshes[i];
        Item *it = (Item *)malloc(sizeof(Item) * (m->ntris ? m->ntris : 1));
        for (uint32_t t = 0; t < m->ntris; ++t) {
            /* Triangle::bounding_box (Triangle.cpp:90-107) */
            V a = m->wv[m->tri[4 * t]], b = m->wv[m->tri[4 * t + 1]], c = m->wv[m->tri[4 * t + 2]];
            it[t].index = t;
            it[t].b.mn = v3(smin(smin(a.x, b.x), c.x), smin(smin(a.y, b.y), c.y), smin(smin(a.z, b.z), c.z));
            it[t].b.mx = v3(smax(smax(a.x, b.x), c.x), smax(smax(a.y, b.y), c.y), smax(smax(a.z, b.z), c.z));
            it[t].c = vdivs(vadd(it[t].b.mn, it[t].b.mx), K(2.0f));
            it[t].cost = 0;
        }
        Flat f = {0};
        flat_push(&f, 0); /* tris_indices[0] is unused (1-based list) */
        uint32_t root = flat_node(&f);
        if (m->ntris) build_rec(&f, root, it, 0, (int)m->ntris, 0);
        else { f.nodes[root].b = box_empty(); f.nodes[root].u0 = 0x80000000u; f.nodes[root].u1 = 0; }
        m->nodes = f.nodes; m->nnodes = f.nn; m->tris_indices = f.list;
        m->bbox = f.nodes[0].b;
        m->cost = cost_of_traversing(m->nodes, 0);
        free(it);
        witems[i].index = i;
        witems[i].b = m->bbox;
        witems[i].c = vdivs(vadd(m->bbox.mn, m->bbox.mx), K(2.0f));
        witems[i].cost = (int)m->cost;
    }
    Flat f = {0};
    flat_push(&f, 0);
    uint32_t root = flat_node(&f);
    if (s->nmeshes) build_rec(&f, root, witems, 0, (int)s->nmeshes, 1);
    else { f.nodes[root].b = box_empty(); f.nodes[root].u0 = 0x80000000u; f.nodes[root].u1 = 0; }
    s->wnodes = f.nodes; s->nwnodes = f.nn; s->mesh_indices = f.list;
    s->bbox = f.nodes[0].b;
    free(witems);
    init_lights(s);
    /* main.cpp:166-168 */
    s->scene_radius = vlen(vsub(s->bbox.mn, vscale(vadd(s->bbox.mn, s->bbox.mx), K(0.5f))));
    s->base_radius *= s->scene_radius;
    s->prepared = 1;
    return 0;
}

/* Scene::hit (Scene.cpp:154-189) with CacheBVHNode_world::hit (BVH_world.cpp:267-283):
 * each world leaf starts a fresh record, exactly as the reference does. */
static int scene_hit(Scene *s, const Ray *r, real tmin, Isect *h) {
    uint32_t stack[64];
    int sp = 0, hit = 0;
    h->t = INF;
    stack[sp++] = 0;
    while (sp) {
        const Node *n = &s->wnodes[stack[--sp]];
        if (box_hit(&n->b, r, tmin, h->t, &s->ctr)) {
            if (n->u0 & 0x80000000u) {
                Isect tmp;
                int any = 0;
                tmp.t = INF;
                int end = (int)(n->u1 + (n->u0 & 0x7fffffffu) + 1);
                for (int i = (int)n->u1 + 1; i < end; ++i)
                    if (mesh_hit(&s->meshes[s->mesh_indices[i]], r, tmin, &tmp, &s->ctr)) { any = 1; *h = tmp; }
                hit |= any;
            } else {
                stack[sp++] = n->u1;
                stack[sp++] = n->u0;
            }
        }
    }
    return hit;
}

static int scene_hit_brute(Scene *s, const Ray *r, real tmin, Isect *h) {
    int hit = 0;
    h->t = INF;
    for (uint32_t i = 0; i < s->nmeshes; ++i)
        for (uint32_t t = 0; t < s->meshes[i].ntris; ++t) {
            Isect tmp;
            if (tri_hit(&s->meshes[i], t, r, tmin, h->t, &tmp, &s->ctr)) { *h = tmp; hit = 1; }
        }
    return hit;
}

void orc_first_hit(void *sp, const float *rays, long long n, float tmin, int brute, int *tri_out, double *t_out) {
    Scene *s = (Scene *)sp;
    for (long long i = 0; i < n; ++i) {
        Ray r = {v3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]), v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5])};
        Isect h;
        int hit = brute ? scene_hit_brute(s, &r, (real)tmin, &h) : scene_hit(s, &r, (real)tmin, &h);
        tri_out[i] = hit ? h.tri_global : -1;
        t_out[i] = hit ? (double)h.t : INFINITY;
    }
}

/* find_interval (utils.h:142-157) */
static int find_interval(const real *cdf, int size, real u) {
    int first = 0, len = size;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        if (cdf[middle] < u) { first = middle + 1; len -= half + 1; }
        else len = half;
    }
    int r = first - 1;
    if (r < 0) r = 0;
    if (r > size - 2) r = size - 2;
    return r;
}

/* Guided direction (the engine's learned proposal): with probability alpha
 * the plain words, otherwise a bin from the CDF of the position cell (cx, cy)
 * and a point inside it; returns q = alpha + (1 - alpha) * 1024 * P(bin). */
static real guide_direction(const Scene *s, int light, real ax, real ay, real *ex, real *ey) {
    uint32_t wx = s->guide_words[0], wy = s->guide_words[1];
    int G = s->guide_cells;
    int cx = (int)(ax * (real)G), cy = (int)(ay * (real)G);
    if (cx > G - 1) cx = G - 1;
    if (cy > G - 1) cy = G - 1;
    const real *cdf = s->guide_cdf + ((size_t)(s->guide_slot[light] * G + cy) * G + cx) * (GUIDE_BINS + 1);
    int bin = (int)(wx >> 27) | ((int)(wy >> 27) << 5);
    if (u24(s->guide_words[2]) >= s->guide_alpha) {
        real uc = (real)((s->guide_words[3] >> 8) + 1u) * K(5.9604644775390625e-08);
        int k = find_interval(cdf, GUIDE_BINS + 1, uc);
        wx = ((uint32_t)(k & 31) << 27) | (wx >> 5);
        wy = ((uint32_t)(k >> 5) << 27) | (wy >> 5);
        bin = k;
    }
    *ex = u24(wx);
    *ey = u24(wy);
    real p = cdf[bin + 1] - cdf[bin];
    return s->guide_alpha + (K(1.0f) - s->guide_alpha) * (real)GUIDE_BINS * p;
}

/* Lights_list::sample_lights (Light_list.cpp:10-30) and the per-light sample_light. */
static int sample_light(Scene *s, Sampler *smp, V *origin, V *dir, V *tp) {
    real rng = get1d(smp);
    int idx = find_interval(s->cdf, (int)s->nlights + 1, rng);
    const Light *L = &s->lights[idx];
    real pick = s->compat_pdf ? L->power / ((real)(int)s->nlights * s->integral) : L->power / s->integral;
    int guided = s->guide_cdf && idx < GUIDE_MAX_LIGHTS && s->guide_slot[idx] >= 0;
    real q = K(1.0f);
    V t;
    if (L->kind == L_AREA) {
        /* Area_Light::sample_light (Lights.cpp:41-55) */
        real ax = get1d(smp), ay = get1d(smp);
        *origin = vadd(vadd(L->bl, vscale(L->h, ax)), vscale(L->v, ay));
        real ex = get1d(smp), ey = get1d(smp);
        if (guided) q = guide_direction(s, idx, ax, ay, &ex, &ey);
        V local = sample_cos_hemisphere(ex, ey);
        *dir = to_world(&L->fr, local);
        t = vscale(L->color, PI);
    } else if (L->kind == L_SUN) {
        /* Sun_Light::sample_light (Lights.cpp:126-136) */
        real ax = get1d(smp), ay = get1d(smp);
        smp->index += 2;
        *origin = vadd(vadd(L->bl, vscale(L->h, ax)), vscale(L->v, ay));
        *dir = L->normal;
        t = L->color;
    } else if (s->compat_point) {
        /* Point_Light::sample_light returns 0 (Lights.h:78-81); the ray then
         * has a zero direction and finds no hit. */
        *origin = vs(0);
        *dir = vs(0);
        t = vs(0);
    } else {
        /* Point / sphere light: isotropic emission with the flux convention
         * of the area light (throughput = color * power * pi). */
        real u1 = get1d(smp), u2 = get1d(smp);
        real z = K(1) - K(2) * u1;
        real rr = (real)sqrt(smax(K(0), K(1) - z * z));
        real phi = PI2 * u2;
        V n = v3(rr * rcos(phi), rr * rsin(phi), z);
        real e1 = get1d(smp), e2 = get1d(smp);
        if (L->radius > 0) {
            if (guided) q = guide_direction(s, idx, u1, u2, &e1, &e2);
            *origin = vadd(L->pos, vscale(n, L->radius));
            Frame fr = frame_from_z(n);
            *dir = to_world(&fr, sample_cos_hemisphere(e1, e2));
        } else {
            *origin = L->pos;
            *dir = n;
        }
        t = vscale(vscale(L->color, L->power), PI);
    }
    *tp = vdiv_each(t, pick);
    if (q != K(1.0f)) *tp = vdiv_each(*tp, q);
    return idx;
}

int orc_set_guide(void *sp, const float *cdf, int slots, int cells, const int *slot_of_light, double alpha) {
    Scene *s = (Scene *)sp;
    free(s->guide_cdf);
    s->guide_cdf = NULL;
    if (!cdf || slots <= 0) return 0;
    size_t n = (size_t)slots * cells * cells * (GUIDE_BINS + 1);
    s->guide_cdf = (real *)malloc(n * sizeof(real));
    if (!s->guide_cdf) return -1;
    for (size_t i = 0; i < n; ++i) s->guide_cdf[i] = (real)cdf[i];
    s->guide_cells = cells;
    for (int i = 0; i < GUIDE_MAX_LIGHTS; ++i) s->guide_slot[i] = slot_of_light[i];
    /* alpha on the 2^-24 grid of the uniform draws */
    s->guide_alpha = (real)(float)(ceil(alpha * 16777216.0) / 16777216.0);
    return 0;
}

void orc_light_sample(void *sp, const float *u5, double *out) {
    Scene *s = (Scene *)sp;
    Sampler smp;
    smp.n = 5; smp.index = 0;
    for (int i = 0; i < 5; ++i) smp.value[i] = u5[i];
    V o, d, t;
    int idx = sample_light(s, &smp, &o, &d, &t);
    double r[10] = {o.x, o.y, o.z, d.x, d.y, d.z, t.x, t.y, t.z, (double)idx};
    memcpy(out, r, sizeof r);
}

/* ------------------------------------------------------------------ */
/* Photon list                                                          */
/* ------------------------------------------------------------------ */
typedef struct { double *p; long long n, cap; } PList;
static void plist_add(PList *l, V pos, V dir, V tp, real dvm, long long uid) {
    if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 4096; l->p = realloc(l->p, (size_t)l->cap * ORC_PHOTON_STRIDE * sizeof(double)); }
    double *r = l->p + (size_t)l->n * ORC_PHOTON_STRIDE;
    r[0] = pos.x; r[1] = pos.y; r[2] = pos.z; r[3] = dir.x; r[4] = dir.y; r[5] = dir.z;
    r[6] = tp.x; r[7] = tp.y; r[8] = tp.z; r[9] = dvm; r[10] = (double)uid;
    l->n++;
}

/* Vertex_Merging::trace_light_path (Vertex_Merging.cpp:140-194) with
 * Generate_light_sample (:18-25) and Sample_Scattering_light (:48-96). */
static int trace_light_path(Scene *s, Sampler *smp, Draws *dr, uint32_t path, PList *out) {
    V o, d, tp;
    sample_light(s, smp, &o, &d, &tp);
    int diffuse = 0, last = EV_DIFFUSE, visible = 0;
    real dvm = 0;
    for (int pl = 0;; ++pl) {
        Ray r = {vadd(o, vscale(d, K(0.001f))), d};
        Isect h;
        if (!scene_hit(s, &r, K(0.001f), &h)) break;
        Bsdf b;
        bsdf_init(&b, &s->mats[h.mat], h.sn, h.dir, 0);
        if (!b.valid) break;
        dvm /= (real)fabs(fabs(b.wo.z));
        int stored = 0;
        if (!bsdf_is_delta(&b) && (last & EV_GLOSSY)) {
            plist_add(out, h.pos, vneg(h.dir), tp, dvm, (long long)path * 8 + pl);
            visible = 1;
            stored = 1;
        }
        if (pl > s->max_path_length || diffuse > 2) break;
        real u0 = get1d(smp), u1 = get1d(smp), u2 = get1d(smp);
        Sample sm;
        V f = bsdf_sample(&b, h.gn, u0, u1, u2, &sm);
        if (f.x == 0 && f.y == 0 && f.z == 0) break;
        if (pl > 2) {
            if (draw_rr(dr, pl) > sm.cont) break;
            f = vdiv_each(f, sm.cont);
        }
        o = h.pos;
        d = sm.wi;
        if (sm.event & EV_DIFFUSE) diffuse++;
        tp = vmul(tp, f);
        last = sm.event;
        if (!stored) dvm *= (sm.cos_out / sm.fwd) * sm.rev;
        else dvm = (sm.cos_out / sm.fwd) * (K(1) + sm.rev * dvm);
    }
    return visible;
}

static long long light_paths_count(const Scene *s) {
    return s->photons > 0 ? s->photons : (long long)s->W * s->H;
}

void orc_iteration_params(void *sp, int iteration, double *radius_o, double *r2_o, double *norm_o) {
    Scene *s = (Scene *)sp;
    /* Vertex_Merging.cpp:201-212 */
    real count = (real)light_paths_count(s);
    real radius = s->base_radius;
#ifdef ORC_DOUBLE
    radius /= pow((real)(iteration + 1), K(0.5) * (K(1) - s->radius_alpha));
#else
    radius /= powf((real)(iteration + 1), K(0.5f) * (K(1.0f) - s->radius_alpha));
#endif
    radius = smax(radius, K(1e-7f));
    real r2 = radius * radius;
    real norm = K(1.0f) / (PI * r2 * count);
    if (radius_o) *radius_o = radius;
    if (r2_o) *r2_o = r2;
    if (norm_o) *norm_o = norm;
}

/* ------------------------------------------------------------------ */
/* Chain-parallel adaptive MLT light pass (rng_serial = 2): the engine's   */
/* k_light_pass_mlt restated.  Chain c takes the indices c, c + C, ...,    */
/* C = clamp(N / 64, 1, mlt_chains).  Primary samples are Philox words in  */
/* the engine's layout (blocks 0-1: light sample, block 2 + k: bounce k    */
/* with its Russian-roulette word); a chain's state is the words of its    */
/* last visible uniform sample, or of its last accepted mutation; a        */
/* mutation moves each 24-bit value by +-e * exp(-1/size) modulo 1 (by     */
/* +-e when size <= 0), e and the sign from the mutation stream.  The      */
/* chain and size arithmetic is float, as on the GPU, so the f32 build     */
/* reproduces the engine's photons; photons keep their throughput (the     */
/* merge carries visible / N).                                             */
/* ------------------------------------------------------------------ */
enum { STREAM_MLT_DELTA = 2, STREAM_MLT_INIT = 3, MLT_BLOCKS = 10, MLT_WORDS = 4 * MLT_BLOCKS };

static float mlt_exp_neg(float x) {
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
static float mlt_step(float size) { return size > 0.0f ? mlt_exp_neg(-1.0f / size) : 1.0f; }

static void mlt_words(const uint32_t key[2], uint32_t path, uint32_t iter, uint32_t stream, uint32_t *w) {
    for (uint32_t b = 0; b < MLT_BLOCKS; ++b) {
        uint32_t ctr[4] = {b, path, iter, stream};
        orc_philox(ctr, key, w + 4 * b);
    }
}
static uint32_t mlt_moved(uint32_t w, uint32_t e_word, uint32_t sign_word, float step) {
    float dlt = (float)u24(e_word) * step;
    dlt -= floorf(dlt);
    uint32_t d24 = (uint32_t)(dlt * 16777216.0f) & 0xFFFFFFu;
    uint32_t v = w >> 8;
    v = (float)u24(sign_word) < 0.5f ? v + d24 : v - d24;
    return (v & 0xFFFFFFu) << 8;
}
static void mlt_propose(const uint32_t key[2], uint32_t path, uint32_t iter, const uint32_t *state, float step,
                        uint32_t *prop) {
    for (uint32_t b = 0; b < MLT_BLOCKS; ++b) {
        uint32_t c0[4] = {2 * b, path, iter, STREAM_MLT_DELTA}, c1[4] = {2 * b + 1, path, iter, STREAM_MLT_DELTA};
        uint32_t r0[4], r1[4];
        orc_philox(c0, key, r0);
        orc_philox(c1, key, r1);
        const uint32_t *s4 = state + 4 * b;
        prop[4 * b + 0] = mlt_moved(s4[0], r0[0], r0[1], step);
        prop[4 * b + 1] = mlt_moved(s4[1], r0[2], r0[3], step);
        prop[4 * b + 2] = mlt_moved(s4[2], r1[0], r1[1], step);
        prop[4 * b + 3] = mlt_moved(s4[3], r1[2], r1[3], step);
    }
}
/* The trace's dimensions from the words: light pick, position, direction
 * (block 0 words a-d, block 1 word a), then three per bounce (block 2 + k). */
static void sampler_from_words(Sampler *smp, const uint32_t *w) {
    smp->index = 0;
    for (int d = 0; d < 4; ++d) smp->value[d] = u24(w[d]);
    smp->value[4] = u24(w[4]);
    for (int b = 0; b < 8; ++b)
        for (int j = 0; j < 3; ++j) smp->value[5 + 3 * b + j] = u24(w[4 * (2 + b) + j]);
    smp->n = 29;
}

static void light_pass_mlt_chains(Scene *s, int iteration, PList *out) {
    long long N = light_paths_count(s);
    const uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    long long C = N / 64 < s->mlt_chains ? N / 64 : s->mlt_chains;
    if (C < 1) C = 1;
    long long visible = 0, mutations = 0, accepted = 0;
    uint32_t state[MLT_WORDS], w[MLT_WORDS], prop[MLT_WORDS];
    for (long long c = 0; c < C && c < N; ++c) {
        float size = 1.0f; /* MLT_Sampler(max_path_length, 1.0f) */
        int acc = 1, mut = 0; /* Vertex_Merging.cpp:227-228 */
        mlt_words(key, (uint32_t)c, (uint32_t)iteration, STREAM_MLT_INIT, state);
        for (long long p = c; p < N; p += C) {
            Draws dr = {0, NULL, {{key[0], key[1]}, (uint32_t)p, (uint32_t)iteration, STREAM_LIGHT}, w};
            Sampler smp;
            mlt_words(key, (uint32_t)p, (uint32_t)iteration, STREAM_LIGHT, w);
            sampler_from_words(&smp, w);
            if (trace_light_path(s, &smp, &dr, (uint32_t)p, out)) {
                ++visible;
                memcpy(state, w, sizeof state);
                continue;
            }
            mlt_propose(key, (uint32_t)p, (uint32_t)iteration, state, mlt_step(size), prop);
            sampler_from_words(&smp, prop);
            dr.words = prop;
            ++mut;
            ++mutations;
            if (trace_light_path(s, &smp, &dr, (uint32_t)p, out)) {
                ++acc;
                ++accepted;
                memcpy(state, prop, sizeof state);
            }
            size += ((float)acc / (float)mut - 0.234f) / (float)mut;
        }
    }
    s->visible_paths = visible;
    s->mlt_mutations = mutations;
    s->mlt_accepted = accepted;
}

static void light_pass(Scene *s, int iteration, PList *out) {
    long long N = light_paths_count(s);
    s->visible_paths = 0;
    if (s->rng_serial == 2) {
        light_pass_mlt_chains(s, iteration, out);
        return;
    }
#ifdef _OPENMP
    if (!s->rng_serial && s->threads > 1 && !s->guide_cdf) {
        /* Paths are independent (counter-based draws): thread t takes the
         * contiguous range [N t / T, N (t + 1) / T) with its own photon list and
         * counters on a shallow copy of the scene, and the lists are joined in
         * thread order -- the serial loop's list, photon for photon. */
        const int T = s->threads;
        PList *lists = (PList *)calloc((size_t)T, sizeof(PList));
        long long *vis = (long long *)calloc((size_t)T, sizeof(long long));
        Counters *ctrs = (Counters *)calloc((size_t)T, sizeof(Counters));
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();
            const long long lo = N * t / nt, hi = N * (t + 1) / nt;
            Scene local = *s;
            local.ctr.box_tests = local.ctr.tri_tests = 0;
            Draws dr = {0, NULL, {{(uint32_t)s->seed, (uint32_t)(s->seed >> 32)}, 0, (uint32_t)iteration, STREAM_LIGHT}};
            long long v = 0;
            for (long long p = lo; p < hi; ++p) {
                dr.key.path = (uint32_t)p;
                Sampler smp;
                sampler_plain(&smp, &dr.key);
                v += trace_light_path(&local, &smp, &dr, (uint32_t)p, &lists[t]);
            }
            vis[t] = v;
            ctrs[t] = local.ctr;
        }
        for (int t = 0; t < T; ++t) {
            for (long long i = 0; i < lists[t].n; ++i) {
                if (out->n == out->cap) {
                    out->cap = out->cap ? 2 * out->cap : 4096;
                    out->p = realloc(out->p, (size_t)out->cap * ORC_PHOTON_STRIDE * sizeof(double));
                }
                memcpy(out->p + (size_t)out->n * ORC_PHOTON_STRIDE, lists[t].p + (size_t)i * ORC_PHOTON_STRIDE,
                       ORC_PHOTON_STRIDE * sizeof(double));
                out->n++;
            }
            free(lists[t].p);
            s->visible_paths += vis[t];
            s->ctr.box_tests += ctrs[t].box_tests;
            s->ctr.tri_tests += ctrs[t].tri_tests;
        }
        free(lists);
        free(vis);
        free(ctrs);
        return;
    }
#endif
    if (!s->rng_serial) {
        Draws dr = {0, NULL, {{(uint32_t)s->seed, (uint32_t)(s->seed >> 32)}, 0, (uint32_t)iteration, STREAM_LIGHT}};
        for (long long p = 0; p < N; ++p) {
            dr.key.path = (uint32_t)p;
            Sampler smp;
            sampler_plain(&smp, &dr.key);
            if (s->guide_cdf) {
                uint32_t c0[4] = {0u, (uint32_t)p, (uint32_t)iteration, STREAM_LIGHT}, c1[4] = {1u, (uint32_t)p,
                    (uint32_t)iteration, STREAM_LIGHT}, w0[4], w1[4];
                orc_philox(c0, dr.key.key, w0);
                orc_philox(c1, dr.key.key, w1);
                s->guide_words[0] = w0[3];
                s->guide_words[1] = w1[0];
                s->guide_words[2] = w1[1];
                s->guide_words[3] = w1[2];
            }
            s->visible_paths += trace_light_path(s, &smp, &dr, (uint32_t)p, out);
        }
        return;
    }
    /* Reference light loop with the adaptive MLT branch (Vertex_Merging.cpp:220-282). */
    Draws dr = {1, &s->xs, {{0, 0}, 0, 0, 0}};
    Sampler smp;
    sampler_init_serial(&smp, s->max_path_length, &s->xs); /* MLT_Sampler(6, 1.0f) (:220) */
    int accepted = 1, mutated = 0;
    long long visible_uniform = 0;
    for (long long p = 0; p < N; ++p) {
        (void)random_float_0_1(&s->xs); /* progress print draw (:233) */
        sampler_init_uniform(&smp, &s->xs);
        smp.index = 0;
        int vis = trace_light_path(s, &smp, &dr, (uint32_t)p, out);
        if (vis) { ++visible_uniform; }
        else {
            sampler_reject(&smp);
            sampler_mutate(&smp, &s->xs);
            mutated += 1;
            smp.index = 0;
            vis = trace_light_path(s, &smp, &dr, (uint32_t)p, out);
            if (vis) accepted += 1; else sampler_reject(&smp);
            real R = (real)accepted / (real)mutated;
            smp.mutation_size += (R - K(0.234f)) / (real)mutated;
        }
    }
    s->visible_paths = visible_uniform;
    s->mlt_mutations = mutated;
    s->mlt_accepted = accepted - 1; /* the reference's count starts at 1 (:227) */
    real c = (real)visible_uniform / (real)N;
    for (long long i = 0; i < out->n; ++i)
        for (int k = 6; k < 9; ++k) out->p[(size_t)i * ORC_PHOTON_STRIDE + k] = (double)((real)out->p[(size_t)i * ORC_PHOTON_STRIDE + k] * c);
    s->serial_smp = smp;
}

/* ------------------------------------------------------------------ */
/* KD tree over photons (KD_Tree.cpp:24-280)                           */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t index; real split; } KDNode;
typedef struct {
    KDNode *nodes;
    size_t nnodes;
    const real *pos;
    uint32_t *tmp;
    Xorshift rng;
} KD;

static inline real kd_key(const KD *k, uint32_t i, int axis) { return k->pos[3 * (size_t)i + axis]; }

/* aproximate_biggest_axis (KD_Tree.cpp:24-95); the sampled branch draws its
 * 400 indices from a private xorshift stream instead of mt19937(0), which only
 * changes the tree layout.  The y/z slip of :84 is kept. */
static int kd_axis(KD *k, int start, int count) {
    V lo = vs(INF), hi = vs(-INF);
    int samples = count < 500 ? count : 400;
    for (int i = 0; i < samples; ++i) {
        int idx = count < 500 ? start + i : start + (int)(xorshift32(&k->rng) % (uint32_t)count);
        uint32_t pi = k->tmp[idx];
        V p = v3(kd_key(k, pi, 0), kd_key(k, pi, 1), kd_key(k, pi, 2));
        if (p.x < lo.x) lo.x = p.x;
        if (p.y < lo.y) lo.y = p.y;
        if (p.z < lo.z) lo.z = p.z;
        if (p.x > hi.x) hi.x = p.x;
        if (p.y > hi.y) hi.y = p.y;
        if (count < 500) { if (p.z > hi.z) hi.z = p.z; }
        else if (p.y > hi.z) hi.z = p.z;
    }
    real x = hi.x - lo.x, y = hi.y - lo.y, z = hi.z - lo.z;
    real big = (real)fmax(fmax(x, y), z);
    return big == x ? 0 : (big == y ? 1 : 2);
}

/* Selection of the k-th element (the role of std::nth_element). */
static void kd_select(KD *k, uint32_t *a, int n, int nth, int axis) {
    int lo = 0, hi = n - 1;
    while (hi > lo) {
        int mid = lo + (hi - lo) / 2;
        real x = kd_key(k, a[lo], axis), y = kd_key(k, a[mid], axis), z = kd_key(k, a[hi], axis);
        real pv = (x < y) ? ((y < z) ? y : ((x < z) ? z : x)) : ((x < z) ? x : ((y < z) ? z : y));
        int i = lo, j = hi;
        while (i <= j) {
            while (kd_key(k, a[i], axis) < pv) i++;
            while (pv < kd_key(k, a[j], axis)) j--;
            if (i <= j) { uint32_t t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
        }
        if (nth <= j) hi = j;
        else if (nth >= i) lo = i;
        else return;
    }
}

static void kd_recurse(KD *k, int start, int count, size_t element) {
    if (count < 2) { k->nodes[element].index = k->tmp[start]; return; }
    int axis = kd_axis(k, start, count);
    kd_select(k, k->tmp + start, count, count / 2, axis);
    uint32_t med = k->tmp[start + count / 2];
    real m = kd_key(k, med, axis);
    if (count & 1) {
        uint32_t mask = axis == 0 ? 0x90000000u : (axis == 1 ? 0xb0000000u : 0xf0000000u);
        k->nodes[element].index = med | mask;
        k->nodes[element].split = m;
        kd_recurse(k, start, count / 2, 2 * element);
        kd_recurse(k, start + count / 2 + 1, count / 2, 2 * element + 1);
    } else {
        uint32_t mask = axis == 0 ? 0x80000000u : (axis == 1 ? 0xa0000000u : 0xe0000000u);
        k->nodes[element].index = mask;
        k->nodes[element].split = m + (real)1.175494e-38f;
        kd_recurse(k, start, count / 2, 2 * element);
        kd_recurse(k, start + count / 2, count / 2, 2 * element + 1);
    }
}

static void kd_build(KD *k, const real *pos, long long P) {
    memset(k, 0, sizeof *k);
    k->pos = pos;
    k->rng.a = 0x2545F491u;
    if (P <= 0) return;
    int levels = (int)(log2((double)P) + 1);
    k->nnodes = (size_t)1 << levels;
    /* the recursion can address up to index 2^levels - 1; size one level more to stay safe */
    k->nnodes *= 2;
    k->nodes = (KDNode *)calloc(k->nnodes, sizeof(KDNode));
    k->tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)P);
    for (long long i = 0; i < P; ++i) k->tmp[i] = (uint32_t)i;
    kd_recurse(k, 0, (int)P, 1);
}
static void kd_free(KD *k) { free(k->nodes); free(k->tmp); }

typedef struct { int *idx; long long n, cap; } IList;
static void ilist_add(IList *l, int v) {
    if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 256; l->idx = realloc(l->idx, (size_t)l->cap * sizeof(int)); }
    l->idx[l->n++] = v;
}

static inline real dist2(const real *pos, uint32_t i, V q) {
    V d = vsub(q, v3(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]));
    return vlen2(d);
}

/* KD_Tree::find_closest_in_radius (KD_Tree.cpp:217-280) */
static void kd_query(const KD *k, V q, real r2, size_t element, IList *out) {
    const KDNode *n = &k->nodes[element];
    if (n->index & 0x80000000u) {
        real dp;
        uint32_t ax = n->index & 0x60000000u;
        if (ax == 0) dp = q.x - n->split;
        else if (ax == 0x20000000u) dp = q.y - n->split;
        else dp = q.z - n->split;
        size_t nearc = dp < 0 ? 2 * element : 2 * element + 1;
        size_t farc = dp < 0 ? 2 * element + 1 : 2 * element;
        kd_query(k, q, r2, nearc, out);
        if (dp * dp < r2) {
            if (n->index & 0x10000000u) {
                uint32_t i = n->index & 0xfffffffu;
                if (dist2(k->pos, i, q) < r2) ilist_add(out, (int)i);
            }
            kd_query(k, q, r2, farc, out);
        }
    } else {
        if (dist2(k->pos, n->index, q) < r2) ilist_add(out, (int)n->index);
    }
}

static void brute_query(const real *pos, long long P, V q, real r2, IList *out) {
    for (long long i = 0; i < P; ++i)
        if (dist2(pos, (uint32_t)i, q) < r2) ilist_add(out, (int)i);
}

static long long gather_sets(const float *pos, long long P, const float *q, long long Q, float r2,
                             long long *offsets, int *indices, long long cap, int use_kd) {
    real *rp = (real *)malloc(sizeof(real) * 3 * (size_t)(P > 0 ? P : 1));
    for (long long i = 0; i < 3 * P; ++i) rp[i] = pos[i];
    KD k;
    if (use_kd) kd_build(&k, rp, P);
    IList l = {0};
    long long total = 0;
    offsets[0] = 0;
    for (long long j = 0; j < Q; ++j) {
        l.n = 0;
        V qq = v3(q[3 * j], q[3 * j + 1], q[3 * j + 2]);
        if (P > 0) {
            if (use_kd) kd_query(&k, qq, (real)r2, 1, &l);
            else brute_query(rp, P, qq, (real)r2, &l);
        }
        for (long long i = 0; i < l.n; ++i) {
            if (total + i < cap) indices[total + i] = l.idx[i];
        }
        total += l.n;
        offsets[j + 1] = total;
    }
    if (use_kd) kd_free(&k);
    free(l.idx);
    free(rp);
    return total;
}

long long orc_gather_sets(const float *pos, long long P, const float *q, long long Q, float r2,
                          long long *offsets, int *indices, long long cap) {
    return gather_sets(pos, P, q, Q, r2, offsets, indices, cap, 0);
}
long long orc_gather_sets_kd(const float *pos, long long P, const float *q, long long Q, float r2,
                             long long *offsets, int *indices, long long cap) {
    return gather_sets(pos, P, q, Q, r2, offsets, indices, cap, 1);
}

/* ------------------------------------------------------------------ */
/* Camera pass (Vertex_Merging.cpp:27-46, 98-135, 309-363)             */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Stratified camera samples (camera_qmc), the engine's rc_vec.h restated:  */
/* sample i of a pixel = point i of the (0,2)-sequence (van der Corput,     */
/* Sobol's second dimension) for the jitter and the base-3 radical inverse  */
/* for the first bounce's first dimension, each shifted modulo 1 by the     */
/* pixel's Philox offset (stream 4, iteration-independent counter).         */
/* ------------------------------------------------------------------ */
enum { STREAM_QMC_SHIFT = 4 };
static uint32_t qmc_vdc(uint32_t i) {
    uint32_t r = 0;
    for (int b = 0; b < 32; ++b) r |= ((i >> b) & 1u) << (31 - b);
    return r;
}
static uint32_t qmc_sobol2(uint32_t i) {
    uint32_t r = 0, v = 1u << 31;
    for (; i; i >>= 1, v ^= v >> 1)
        if (i & 1u) r ^= v;
    return r;
}
static uint32_t qmc_ri3_24(uint32_t i) {
    uint64_t num = 0, den = 1;
    for (; i; i /= 3u) { num = num * 3u + i % 3u; den *= 3u; }
    return (uint32_t)((num << 24) / den);
}
static uint32_t qmc_word(uint32_t v24, uint32_t shift_word) { return ((v24 + (shift_word >> 8)) & 0xFFFFFFu) << 8; }
static void qmc_camera_words(const uint32_t key[2], uint32_t pixel, uint32_t iteration, uint32_t w[3]) {
    uint32_t ctr[4] = {0u, pixel, 0xFFFFFFFFu, STREAM_QMC_SHIFT}, sh[4];
    orc_philox(ctr, key, sh);
    w[0] = qmc_word(qmc_vdc(iteration) >> 8, sh[0]);
    w[1] = qmc_word(qmc_sobol2(iteration) >> 8, sh[1]);
    w[2] = qmc_word(qmc_ri3_24(iteration), sh[2]);
}
/* The three camera words of a pixel's sample `iteration` (seed = Philox key). */
void orc_qmc_words(unsigned long long seed, unsigned int pixel, unsigned int iteration, unsigned int *out3) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[3];
    qmc_camera_words(key, pixel, iteration, w);
    for (int k = 0; k < 3; ++k) out3[k] = w[k];
}

/* What a camera path's merges read: the photon list, its positions and KD
 * tree, the radius and the VM normalisation of the iteration. */
typedef struct {
    const double *ph;
    const real *pos;
    long long P;
    const KD *kd;
    real r2, norm;
} GatherCtx;

/* The camera path of pixel sample p (Vertex_Merging.cpp:309-363): its colour
 * and the film element it adds to (*at = -1: none). */
static void camera_path(Scene *s, int iteration, const GatherCtx *g, long long p, Draws *dr, Sampler *smp,
                        IList *found, V *color_out, long long *at_out) {
    const int W = s->W, H = s->H;
    *at_out = -1;
    real jx, jy;
    if (s->rng_serial) {
        (void)random_float_0_1(&s->xs); /* progress draw (:311) */
        jx = random_float_0_1(&s->xs);  /* argument order unspecified in C++; taken left to right */
        jy = random_float_0_1(&s->xs);
        sampler_init_uniform(smp, &s->xs);
    } else {
        dr->key.path = (uint32_t)p;
        jx = path_u(&dr->key, 0);
        jy = path_u(&dr->key, 1);
        sampler_plain(smp, &dr->key);
        if (s->camera_qmc) {
            uint32_t w[3];
            qmc_camera_words(dr->key.key, (uint32_t)p, (uint32_t)iteration, w);
            jx = u24(w[0]);
            jy = u24(w[1]);
            smp->value[0] = u24(w[2]); /* bounce 0, first dimension */
        }
    }
    smp->index = 0;
    real sx = (real)(int)(p % W) + K(0.0f) + K(1.0f) * jx;
    real sy = (real)(int)(p / W) + K(0.0f) + K(1.0f) * jy;
    /* Camera::get_ray (Camera.h:29-34) */
    V dir = vadd(vadd(s->cam_bl, vscale(s->cam_h, sx / (real)W)), vscale(s->cam_v, sy / (real)H));
    dir = vsub(dir, s->cam_o);
    V o = s->cam_o, d = vnormalize(dir), tp = vs(1), color = vs(0);
    real dvm = 0;
    for (int pl = 0;; ++pl) {
        Ray r = {vadd(o, vscale(d, K(0.001f))), d};
        Isect h;
        if (!scene_hit(s, &r, K(0.001f), &h)) break;
        Bsdf b;
        bsdf_init(&b, &s->mats[h.mat], h.sn, h.dir, 1);
        if (!b.valid) break;
        dvm /= (real)fabs(b.wo.z);
        if (!bsdf_is_delta(&b)) {
            /* Path_Vertices::get_contrib (Path_Vertices.h:46-68) */
            found->n = 0;
            if (g->P > 0) {
                if (s->gather_brute) brute_query(g->pos, g->P, h.pos, g->r2, found);
                else kd_query(g->kd, h.pos, g->r2, 1, found);
            }
            V c = vs(0);
            for (long long i = 0; i < found->n; ++i) {
                const double *pr = g->ph + (size_t)found->idx[i] * ORC_PHOTON_STRIDE;
                V pdir = v3((real)pr[3], (real)pr[4], (real)pr[5]);
                V ptp = v3((real)pr[6], (real)pr[7], (real)pr[8]);
                real ldvm = (real)pr[9];
                real fwd, rev;
                bsdf_pdf(&b, pdir, &fwd, &rev);
                real w = K(1.0f) / (K(1.0f) + dvm * rev + ldvm * fwd);
                V e = bsdf_eval(&b, pdir, h.gn);
                c = vadd(c, vscale(vmul(e, ptp), w));
            }
            color = vadd(color, vmul(vscale(tp, g->norm), c));
        }
        if (pl > s->max_path_length) break;
        real u0 = get1d(smp), u1 = get1d(smp), u2 = get1d(smp);
        Sample sm;
        V f = bsdf_sample(&b, h.gn, u0, u1, u2, &sm);
        if (f.x == 0 && f.y == 0 && f.z == 0) break;
        if (pl > 2) {
            if (draw_rr(dr, pl) > sm.cont) break;
            f = vdiv_each(f, sm.cont);
        }
        o = h.pos;
        d = sm.wi;
        tp = vmul(tp, f);
        if (sm.event & EV_SPECULAR) dvm *= (sm.cos_out / sm.fwd) * sm.rev;
        else dvm = (sm.cos_out / sm.fwd) * (K(1) + sm.rev * dvm);
    }
    *color_out = color;
    /* Image_Buffer::add_color (Image_Buffer.h:24-37) */
    if (sx < 0 || sx >= (real)W) return;
    if (sy < 0 || sy >= (real)H) return;
    int X = (int)sx;
    int Y = s->compat_rows ? H - (int)sy : H - 1 - (int)sy;
    long long at = (long long)X + (long long)Y * W;
    if (at > (long long)W * H - 1 || at < 0) return;
    *at_out = at;
}

static void camera_pass(Scene *s, int iteration, const double *ph, long long P, real *image) {
    double dr_, dr2, dnorm;
    orc_iteration_params(s, iteration, &dr_, &dr2, &dnorm);
    real *pos = (real *)malloc(sizeof(real) * 3 * (size_t)(P > 0 ? P : 1));
    for (long long i = 0; i < P; ++i)
        for (int c = 0; c < 3; ++c) pos[3 * i + c] = (real)ph[(size_t)i * ORC_PHOTON_STRIDE + c];
    KD kd;
    if (!s->gather_brute) kd_build(&kd, pos, P);
    const GatherCtx g = {ph, pos, P, &kd, (real)dr2, (real)dnorm};
    const long long pixels = (long long)s->W * s->H, stride = s->camera_stride;
#ifdef _OPENMP
    if (!s->rng_serial && s->threads > 1) {
        /* Camera paths are independent: sample k = p / stride is traced on any
         * thread (shallow scene copy for the counters), its colour kept, and
         * the colours are added to the film in sample order as the serial loop
         * adds them. */
        const int T = s->threads;
        const long long n = (pixels + stride - 1) / stride;
        V *col = (V *)malloc(sizeof(V) * (size_t)(n > 0 ? n : 1));
        long long *at = (long long *)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
        Counters *ctrs = (Counters *)calloc((size_t)T, sizeof(Counters));
#pragma omp parallel num_threads(T)
        {
            Scene local = *s;
            local.ctr.box_tests = local.ctr.tri_tests = 0;
            Draws dr = {0, NULL, {{(uint32_t)s->seed, (uint32_t)(s->seed >> 32)}, 0, (uint32_t)iteration, STREAM_CAMERA}};
            Sampler plain;
            IList found = {0};
#pragma omp for schedule(static)
            for (long long k = 0; k < n; ++k) camera_path(&local, iteration, &g, k * stride, &dr, &plain, &found, &col[k], &at[k]);
            free(found.idx);
            ctrs[omp_get_thread_num()] = local.ctr;
        }
        for (long long k = 0; k < n; ++k) {
            if (at[k] < 0) continue;
            image[3 * at[k]] += col[k].x;
            image[3 * at[k] + 1] += col[k].y;
            image[3 * at[k] + 2] += col[k].z;
        }
        for (int t = 0; t < T; ++t) {
            s->ctr.box_tests += ctrs[t].box_tests;
            s->ctr.tri_tests += ctrs[t].tri_tests;
        }
        free(col);
        free(at);
        free(ctrs);
        if (!s->gather_brute) kd_free(&kd);
        free(pos);
        return;
    }
#endif
    IList found = {0};
    Draws dr = {s->rng_serial, &s->xs, {{(uint32_t)s->seed, (uint32_t)(s->seed >> 32)}, 0, (uint32_t)iteration, STREAM_CAMERA}};
    Sampler *smp = &s->serial_smp, plain;
    if (!s->rng_serial) smp = &plain;
    for (long long p = 0; p < pixels; p += stride) {
        V color;
        long long at;
        camera_path(s, iteration, &g, p, &dr, smp, &found, &color, &at);
        if (at < 0) continue;
        image[3 * at] += color.x;
        image[3 * at + 1] += color.y;
        image[3 * at + 2] += color.z;
    }
    if (!s->gather_brute) kd_free(&kd);
    free(found.idx);
    free(pos);
}

long long orc_trace_light(void *sp, int iteration, double *photons, long long cap) {
    Scene *s = (Scene *)sp;
    s->ctr.box_tests = s->ctr.tri_tests = 0;
    PList l = {0};
    light_pass(s, iteration, &l);
    long long n = l.n < cap ? l.n : cap;
    if (n > 0) memcpy(photons, l.p, (size_t)n * ORC_PHOTON_STRIDE * sizeof(double));
    free(l.p);
    return l.n;
}

/* Debug AOVs: Vertex_Merging::run_debug_normals / _facing / _materials
 * (Debug_scene.cpp:21-119, 156-205): one unjittered camera ray per pixel,
 * Scene::hit with tmin 1e-4 (brute: the true closest hit instead of the
 * world-BVH traversal); adds into image (W*H*3). */
void orc_debug_aov(void *sp, int mode, double ior_d, int brute, double *image) {
    Scene *s = (Scene *)sp;
    const int W = s->W, H = s->H;
    const real ior = (real)(float)ior_d;
    for (long long p = 0; p < (long long)W * H; ++p) {
        real sx = (real)(int)(p % W), sy = (real)(int)(p / W);
        V dir = vadd(vadd(s->cam_bl, vscale(s->cam_h, sx / (real)W)), vscale(s->cam_v, sy / (real)H));
        dir = vsub(dir, s->cam_o);
        Ray r = {s->cam_o, vnormalize(dir)};
        Isect h;
        if (!(brute ? scene_hit_brute(s, &r, K(0.0001f), &h) : scene_hit(s, &r, K(0.0001f), &h))) continue;
        V c;
        if (mode == 2) {
            c = h.sn;                                    /* Debug_scene.cpp:59 */
        } else if (mode == 3) {
            int front = vdot(h.sn, r.d) < 0;             /* Triangle.cpp:75 */
            real factor = (real)fabs(vdot(r.d, h.sn));   /* Debug_scene.cpp:110 */
            c = vscale(front ? v3(0, 0, 1) : v3(1, 0, 0), factor);
        } else {
            Frame fr = frame_from_z(h.sn);               /* Debug_scene.cpp:185-199 */
            V wl = to_local(&fr, vscale(r.d, K(-1)));
            real eta = wl.z > 0 ? K(1) / ior : ior;
            real f = fresnel_dielectric_eta((real)fabs(wl.z), eta);
            c = v3(f, f, f);
        }
        int X = (int)sx;
        int Y = s->compat_rows ? H - (int)sy : H - 1 - (int)sy;
        long long at = (long long)X + (long long)Y * W;
        if (at > (long long)W * H - 1 || at < 0) continue;
        image[3 * at] += c.x;
        image[3 * at + 1] += c.y;
        image[3 * at + 2] += c.z;
    }
}

void orc_camera_pass(void *sp, int iteration, const double *photons, long long count, double *image) {
    Scene *s = (Scene *)sp;
    size_t n = (size_t)s->W * s->H * 3;
    real *img = (real *)malloc(sizeof(real) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) img[i] = (real)image[i];
    camera_pass(s, iteration, photons, count, img);
    for (size_t i = 0; i < n; ++i) image[i] = img[i];
    free(img);
}

void orc_render(void *sp, int iterations, double *image) {
    Scene *s = (Scene *)sp;
    size_t n = (size_t)s->W * s->H * 3;
    real *img = (real *)calloc(n ? n : 1, sizeof(real));
    for (int it = 0; it < iterations; ++it) {
        PList l = {0};
        light_pass(s, it, &l);
        if (s->rng_serial == 2) { /* visibility normalisation of the chain-parallel MLT pass */
            real c = (real)s->visible_paths / (real)light_paths_count(s);
            for (long long i = 0; i < l.n; ++i)
                for (int k = 6; k < 9; ++k)
                    l.p[(size_t)i * ORC_PHOTON_STRIDE + k] = (double)((real)l.p[(size_t)i * ORC_PHOTON_STRIDE + k] * c);
        }
        camera_pass(s, it, l.p, l.n, img);
        free(l.p);
    }
    /* Image_Buffer::scale (Image_Buffer.h:38-44) */
    real sc = K(1.0f) / (real)iterations;
    for (size_t i = 0; i < n; ++i) image[i] = (double)(img[i] * sc);
    free(img);
}

void orc_counters(void *sp, long long *out3) {
    Scene *s = (Scene *)sp;
    out3[0] = s->ctr.box_tests;
    out3[1] = s->ctr.tri_tests;
    out3[2] = s->visible_paths;
}

void orc_bsdf(const float *mat, int mode, const float *shade_n, const float *geo_n, const float *dir,
              const float *u3, const float *wi, int op, double *out) {
    unsigned char pm[28] = {0};
    const int rough_glass = (int)mat[0] == 'R'; /* test code: glass as the rough dielectric */
    pm[0] = rough_glass ? (unsigned char)'G' : (unsigned char)(int)mat[0];
    memcpy(pm + 4, &mat[1], 4);
    memcpy(pm + 8, &mat[2], 4);
    memcpy(pm + 12, &mat[3], 4);
    memcpy(pm + 16, &mat[4], 12);
    Material m = material_from_python(pm);
    m.rough_glass = rough_glass;
    Bsdf b;
    bsdf_init(&b, &m, v3(shade_n[0], shade_n[1], shade_n[2]), v3(dir[0], dir[1], dir[2]), mode == 0);
    V gn = v3(geo_n[0], geo_n[1], geo_n[2]);
    for (int i = 0; i < 13; ++i) out[i] = 0;
    out[11] = b.valid;
    out[12] = bsdf_is_delta(&b);
    if (op == 0) {
        Sample sm;
        memset(&sm, 0, sizeof sm);
        V f = bsdf_sample(&b, gn, u3[0], u3[1], u3[2], &sm);
        out[0] = f.x; out[1] = f.y; out[2] = f.z;
        out[3] = sm.wi.x; out[4] = sm.wi.y; out[5] = sm.wi.z;
        out[6] = sm.fwd; out[7] = sm.rev; out[8] = sm.cos_out; out[9] = sm.cont; out[10] = sm.event;
    } else if (op == 1) {
        V e = bsdf_eval(&b, v3(wi[0], wi[1], wi[2]), gn);
        out[0] = e.x; out[1] = e.y; out[2] = e.z;
    } else {
        real fwd, rev;
        bsdf_pdf(&b, v3(wi[0], wi[1], wi[2]), &fwd, &rev);
        out[6] = fwd; out[7] = rev;
    }
}
