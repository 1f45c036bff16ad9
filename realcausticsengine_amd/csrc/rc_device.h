// rc_device.h -- device functions of the vertex-merging path: BSDFs
// (materials.h, BXDF.cpp), light emission (Lights.cpp, Light_list.cpp),
// BVH closest hit with an LDS stack, and the photon-map radius query.
#pragma once
#include "rc_engine.h"
#include "rc_vec.h"

namespace rc {

constexpr float kPI = 3.14159265359f;        // main.cpp:44
constexpr float kPI2 = 6.28318530718f;       // main.cpp:47
constexpr float kInvPI = 0.31830988618f;     // main.cpp:48
constexpr float kPIOver2 = 1.57079632679f;   // main.cpp:49
constexpr float kPIOver4 = 0.78539816339f;   // main.cpp:50

__device__ __forceinline__ F3 ld3(const float* p) { return f3(p[0], p[1], p[2]); }

// Sampling.h:15-46 (Shirley-Chiu concentric disk, cosine hemisphere)
__device__ __forceinline__ F3 cos_hemisphere(float ex, float ey) {
    float ux = ex * 2.0f - 1.0f, uy = ey * 2.0f - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (!(ux == 0.0f && uy == 0.0f)) {
        float theta, r;
        if (fabsf(ux) > fabsf(uy)) {
            r = ux;
            theta = kPIOver4 * (uy / ux);
        } else {
            r = uy;
            theta = kPIOver2 - kPIOver4 * (ux / uy);
        }
        dx = own_cos(theta) * r;
        dy = own_sin(theta) * r;
    }
    float z = sqrtf(smax(0.0f, 1.0f - dx * dx - dy * dy));
    return f3(dx, dy, z);
}

// ------------------------------------------------------------------
// BSDFs
// ------------------------------------------------------------------
struct Bsdf {
    int kind;
    F3 R;
    float a2, s, ior, cont;
    F3 wo;
    Frame fr;
    bool radiance;
};

__device__ __forceinline__ bool bsdf_setup(Bsdf& b, const DevMaterial& m, F3 shade_n, F3 dir, bool radiance) {
    b.kind = m.kind;
    b.R = f3(m.R[0], m.R[1], m.R[2]);
    b.a2 = m.a2;
    b.s = m.s;
    b.ior = m.ior;
    b.cont = m.cont;
    b.radiance = radiance;
    b.fr = frame_from_z(shade_n);
    b.wo = to_local(b.fr, -dir);
    return !(fabsf(b.wo.z) < 1e-6f);
}
__device__ __forceinline__ bool bsdf_delta(const Bsdf& b) { return b.kind == BK_SPEC_REFL || b.kind == BK_FRESNEL_SPEC; }

__device__ __forceinline__ float fresnel_eta(float c, float eta) {
    float g = eta * eta - 1.0f + c * c;
    if (g > 0.0f) {
        g = sqrtf(g);
        float A = (g - c) / (g + c);
        float B = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
        return 0.5f * A * A * (1.0f + B * B);
    }
    return 1.0f;
}
__device__ __forceinline__ float fresnel_refl(float c, float refl) {
    if (refl >= 1.0f) return 1.0f;
    float n = (1.0f + sqrtf(refl)) / (1.0f - sqrtf(refl));
    float g = n * n - 1.0f + c * c;
    if (g > 0.0f) {
        g = sqrtf(g);
        float A = (g - c) / (g + c);
        float B = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
        return 0.5f * A * A * (1.0f + B * B);
    }
    return 1.0f;
}
__device__ __forceinline__ float ggx_D(float ndoth, float a2) {
    float den = ndoth * ndoth * (a2 - 1.0f) + 1.0f;
    return a2 / (kPI * den * den);
}
__device__ __forceinline__ float ggx_G2(float nw, float nl, float a2) {
    float A = nw * sqrtf(a2 + (1.0f - a2) * nl * nl);
    float B = nl * sqrtf(a2 + (1.0f - a2) * nw * nw);
    return 2.0f * nl * nw / (A + B);
}
__device__ __forceinline__ F3 ggx_half(float a2, float e1, float e2) {
    float theta = own_acos(sqrtf((1.0f - e1) / (e1 * (a2 - 1.0f) + 1.0f)));
    float phi = e2 * kPI2;
    float st = own_sin(theta);
    return f3(st * own_cos(phi), st * own_sin(phi), own_cos(theta));
}
__device__ __forceinline__ bool same_side(const Bsdf& b, F3 wi, F3 gn) {
    return !(dot(wi, gn) * dot(to_world(b.fr, b.wo), gn) <= 0.0f);
}

struct BsdfSample {
    F3 wi;
    float fwd, rev, cos_out, cont;
    int event;
};

// BxDF::sample (BXDF.cpp:25-53, 136-205, 302-337, 384-489, 568-632)
__device__ F3 bsdf_sample(const Bsdf& b, F3 gn, float u0, float u1, float u2, BsdfSample& o) {
    const F3 zero = f3s(0.0f);
    F3 wo = b.wo;
    if (b.kind == BK_SPEC_REFL) {
        o.event = EV_REFLECTION | EV_GLOSSY | EV_SPECULAR;
        F3 wil = f3(-wo.x, -wo.y, wo.z);
        o.cos_out = fabsf(wil.z);
        if (o.cos_out < 1e-6f) return zero;
        o.wi = to_world(b.fr, wil);
        if (!same_side(b, o.wi, gn)) return zero;
        o.fwd = 1.0f;
        o.rev = 1.0f;
        o.cont = b.cont;
        return b.R;
    }
    if (b.kind == BK_FRESNEL_SPEC) {
        float eta = b.ior;
        if (wo.z > 0.0f) eta = 1.0f / b.ior;
        float f = fresnel_eta(fabsf(wo.z), 1.0f / eta);
        if (u0 < f) {
            o.event = EV_GLOSSY | EV_SPECULAR | EV_REFLECTION;
            F3 wil = f3(-wo.x, -wo.y, wo.z);
            o.cos_out = fabsf(wil.z);
            if (o.cos_out < 1e-6f) return zero;
            o.wi = to_world(b.fr, wil);
            if (!same_side(b, o.wi, gn)) return zero;
            o.fwd = f;
            o.rev = f;
            o.cont = b.cont;
            return b.R;
        }
        o.event = EV_TRANSMISSION | EV_GLOSSY | EV_SPECULAR;
        F3 wil;
        if (!refract(wo, wo.z > 0.0f ? f3(0.0f, 0.0f, 1.0f) : f3(0.0f, 0.0f, -1.0f), eta, &wil)) return zero;
        o.cos_out = fabsf(wil.z);
        if (o.cos_out < 1e-6f) return zero;
        o.wi = to_world(b.fr, wil);
        if (dot(o.wi, gn) * dot(to_world(b.fr, wo), gn) >= 0.0f) return zero;
        o.fwd = 1.0f - f;
        o.rev = 1.0f - f;
        o.cont = b.cont;
        if (b.radiance) return (b.R * eta) * eta;
        return b.R;
    }
    if (b.kind == BK_MRG) {
        o.event = EV_GLOSSY | EV_REFLECTION;
        F3 h = ggx_half(b.a2, u1, u2);
        if (wo.z < 0.0f) h.z = -h.z;
        F3 wil = reflect(-wo, h);
        o.cos_out = fabsf(wil.z);
        if (o.cos_out < 1e-6f) return zero;
        o.wi = to_world(b.fr, wil);
        if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
        float chn = fabsf(h.z);
        float d = ggx_D(chn, b.a2);
        o.fwd = o.rev = d * chn * 0.25f / absdot(h, wo);
        o.cont = b.cont;
        float g2 = ggx_G2(fabsf(wo.z), fabsf(wil.z), b.a2);
        return div_recip((b.R * g2) * absdot(wo, h), fabsf(wo.z) * chn);
    }
    if (b.kind == BK_RGLASS) {
        // rough dielectric (Walter, Marschner, Li, Torrance 2007): a GGX
        // microfacet normal on wo's side; the Fresnel term of wo at the
        // microfacet picks reflection or refraction through it.  The weight
        // (f |cos wi| / pdf) is G2 (wo.m) / (|wo.n| |m.n|) for both, times
        // eta^2 for refraction in radiance mode -- Fresnel_Specular's weights
        // in the limit of zero roughness.
        F3 m = ggx_half(b.a2, u1, u2);
        if (wo.z < 0.0f) m.z = -m.z;
        const float wom = dot(wo, m);
        if (!(wom > 0.0f)) return zero;
        float eta = b.ior;
        if (wo.z > 0.0f) eta = 1.0f / b.ior;
        const float f = fresnel_eta(wom, 1.0f / eta);
        const float mz = fabsf(m.z);
        const float d = ggx_D(mz, b.a2);
        o.cont = b.cont;
        if (u0 < f) {
            o.event = EV_GLOSSY | EV_REFLECTION;
            F3 wil = reflect(-wo, m);
            o.cos_out = fabsf(wil.z);
            if (o.cos_out < 1e-6f) return zero;
            o.wi = to_world(b.fr, wil);
            if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
            o.fwd = o.rev = ((f * d) * mz) * 0.25f / wom;
            const float g2 = ggx_G2(fabsf(wo.z), fabsf(wil.z), b.a2);
            return (b.R * g2) * (wom / (fabsf(wo.z) * mz));
        }
        o.event = EV_GLOSSY | EV_TRANSMISSION;
        F3 wil;
        if (!refract(wo, m, eta, &wil)) return zero;
        o.cos_out = fabsf(wil.z);
        if (o.cos_out < 1e-6f || wil.z * wo.z >= 0.0f) return zero;
        o.wi = to_world(b.fr, wil);
        if (dot(o.wi, gn) * dot(to_world(b.fr, wo), gn) >= 0.0f) return zero;
        // pdfs with the Jacobian of the refracted direction, wi.m + eta wo.m
        // being the generalised half vector's length over the wi-side index
        const float wim = dot(wil, m);
        const float sq = wim + eta * wom;
        const float den = sq * sq;
        const float t = ((1.0f - f) * d) * mz;
        o.fwd = (t * fabsf(wim)) / den;
        o.rev = (t * ((eta * eta) * wom)) / den;
        const float g2 = ggx_G2(fabsf(wo.z), fabsf(wil.z), b.a2);
        F3 w = (b.R * g2) * (wom / (fabsf(wo.z) * mz));
        if (b.radiance) w = (w * eta) * eta;
        return w;
    }
    // plastics
    float f = fresnel_refl(fabsf(wo.z), b.s);
    if (b.kind == BK_FMG) {
        if (u0 < f) {
            o.event = EV_GLOSSY | EV_REFLECTION;
            F3 h = ggx_half(b.a2, u1, u2);
            if (wo.z < 0.0f) h.z = -h.z;
            F3 wil = reflect(-wo, h);
            o.cos_out = fabsf(wil.z);
            if (o.cos_out < 1e-6f) return zero;
            o.wi = to_world(b.fr, wil);
            if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
            float g2 = ggx_G2(fabsf(wo.z), fabsf(wil.z), b.a2);
            o.cont = smax(0.05f, f);
            float fd = fabsf(wil.z) * kInvPI, rd = fabsf(wo.z) * kInvPI;
            float chn = fabsf(h.z);
            float d = ggx_D(chn, b.a2);
            float sp = d * chn * 0.25f / absdot(wo, h);
            if (f > 1.0f) {
                o.fwd = o.rev = sp;
            } else {
                o.fwd = fd * (1.0f - f) + sp * f;
                o.rev = rd * (1.0f - f) + sp * f;
            }
            return f3s(g2 * absdot(wo, h) / (fabsf(wo.z) * chn));
        }
        o.event = EV_DIFFUSE;
        F3 wil = cos_hemisphere(u1, u2);
        o.cos_out = wil.z;
        if (o.cos_out < 1e-6f) return zero;
        if (wo.z < 0.0f) wil.z = -wil.z;
        o.wi = to_world(b.fr, wil);
        if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
        o.cont = smax(0.05f, maxc(b.R * (1.0f - f)));
        float fd = fabsf(wil.z) * kInvPI, rd = fabsf(wo.z) * kInvPI;
        F3 h = normalize(wo + wil);
        float chn = fabsf(h.z);
        float d = ggx_D(chn, b.a2);
        float sp = d * chn * 0.25f / absdot(wo, h);
        if (f > 1.0f) {
            o.fwd = o.rev = sp;
        } else {
            o.fwd = fd * (1.0f - f) + sp * f;
            o.rev = rd * (1.0f - f) + sp * f;
        }
        return b.R;
    }
    // BK_FSMG
    if (u0 < f) {
        o.event = EV_SPECULAR | EV_REFLECTION | EV_GLOSSY;
        F3 wil = f3(-wo.x, -wo.y, wo.z);
        o.cos_out = fabsf(wil.z);
        if (o.cos_out < 1e-6f) return zero;
        o.wi = to_world(b.fr, wil);
        if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
        o.fwd = f;
        o.rev = f;
        o.cont = smax(0.05f, f);
        return f3s(1.0f);
    }
    o.event = EV_DIFFUSE;
    F3 wil = cos_hemisphere(u1, u2);
    o.cos_out = wil.z;
    if (o.cos_out < 1e-6f) return zero;
    if (wo.z < 0.0f) wil.z = -wil.z;
    o.wi = to_world(b.fr, wil);
    if (!same_side(b, o.wi, gn) || wil.z * wo.z <= 0.0f) return zero;
    o.fwd = fabsf(wil.z) * kInvPI;
    o.rev = fabsf(wo.z) * kInvPI;
    o.cont = smax(0.05f, maxc(b.R * (1.0f - f)));
    return b.R;
}

// Per-query constants of the merge: the same operations the reference
// repeats for every gathered photon, evaluated once (bitwise identical).
struct MergeConst {
    float f;        // Fresnel reflectance of wo (plastics), fresnel_dielectric_reflectance
    float side_wo;  // dot(to_world(wo), gn) of the same-hemisphere test
    float rd;       // |wo.z| / pi, reverse diffuse pdf
    float awz;      // |wo.z|
    F3 diffuse;     // R / pi
    F3 gn;
};
__device__ __forceinline__ MergeConst merge_const(const Bsdf& b, F3 gn) {
    MergeConst m;
    m.awz = fabsf(b.wo.z);
    m.f = (b.kind == BK_FMG || b.kind == BK_FSMG) ? fresnel_refl(m.awz, b.s) : 0.0f;
    m.side_wo = dot(to_world(b.fr, b.wo), gn);
    m.rd = m.awz * kInvPI;
    m.diffuse = b.R * kInvPI;
    m.gn = gn;
    return m;
}

// Division and square root of the merge evaluation.  kFast uses the
// hardware reciprocal and square root (v_rcp_f32 / v_sqrt_f32, <= 1 ulp each,
// so a quotient is within 2 ulp); otherwise the correctly rounded forms that
// match the reference bit for bit.
template <bool kFast>
__device__ __forceinline__ float mdiv(float a, float b) {
    return kFast ? a * __builtin_amdgcn_rcpf(b) : a / b;
}
template <bool kFast>
__device__ __forceinline__ float msqrt(float x) {
    return kFast ? __builtin_amdgcn_sqrtf(x) : sqrtf(x);
}
template <bool kFast>
__device__ __forceinline__ F3 mnormalize(F3 a) {
    float s = mdiv<kFast>(1.0f, msqrt<kFast>(len2(a)));
    return a * s;
}
template <bool kFast>
__device__ __forceinline__ float mfresnel_eta(float c, float eta) {
    float g = eta * eta - 1.0f + c * c;
    if (g > 0.0f) {
        g = msqrt<kFast>(g);
        float A = mdiv<kFast>(g - c, g + c);
        float B = mdiv<kFast>(c * (g + c) - 1.0f, c * (g - c) + 1.0f);
        return 0.5f * A * A * (1.0f + B * B);
    }
    return 1.0f;
}
template <bool kFast>
__device__ __forceinline__ float mggx_D(float ndoth, float a2) {
    float den = ndoth * ndoth * (a2 - 1.0f) + 1.0f;
    return mdiv<kFast>(a2, kPI * den * den);
}
template <bool kFast>
__device__ __forceinline__ float mggx_G2(float nw, float nl, float a2) {
    float A = nw * msqrt<kFast>(a2 + (1.0f - a2) * nl * nl);
    float B = nl * msqrt<kFast>(a2 + (1.0f - a2) * nw * nw);
    return mdiv<kFast>(2.0f * nl * nw, A + B);
}

// BxDF::evaluate and BxDF::pdf together, for the merge (Path_Vertices.h:46-68):
// MRG BXDF.cpp:339-367, FMG :491-548, FSMG :634-667; delta lobes give zero.
template <bool kFast>
__device__ __forceinline__ F3 bsdf_eval_pdf(const Bsdf& b, const MergeConst& m, F3 wi, float& fwd, float& rev) {
    const F3 zero = f3s(0.0f);
    F3 wo = b.wo;
    F3 wil = to_local(b.fr, wi);
    bool side = !(dot(wi, m.gn) * m.side_wo <= 0.0f);
    if (b.kind == BK_MRG) {
        F3 h = mnormalize<kFast>(wo + wil);
        float chn = fabsf(h.z);
        float d = mggx_D<kFast>(chn, b.a2);
        fwd = rev = mdiv<kFast>(d * chn * 0.25f, absdot(h, wo));
        if (!side || wil.z * wo.z <= 0.0f) return zero;
        float g2 = mggx_G2<kFast>(m.awz, fabsf(wil.z), b.a2);
        return (((b.R * d) * g2) * 0.25f) * mdiv<kFast>(1.0f, m.awz);
    }
    if (kFast && b.kind == BK_FMG && m.f == 0.0f) {
        // specular 0: the lobe is R / pi with pdf |cos| / pi; the GGX terms are
        // multiplied by f = 0 (they only differ when wi = -wo, where they are NaN)
        fwd = fabsf(wil.z) * kInvPI;
        rev = m.rd;
        if (!side || wil.z * wo.z <= 0.0f) return zero;
        return m.diffuse;
    }
    if (b.kind == BK_FMG) {
        float fd = fabsf(wil.z) * kInvPI;
        // wo + wi and wi + wo round identically, so pdf and evaluate share h
        F3 h = mnormalize<kFast>(wo + wil);
        float chn = fabsf(h.z);
        float d = mggx_D<kFast>(chn, b.a2);
        float sp = mdiv<kFast>(d * chn * 0.25f, absdot(h, wo));
        float f = m.f;
        if (f > 1.0f) {
            fwd = rev = sp;
        } else {
            fwd = fd * (1.0f - f) + sp * f;
            rev = m.rd * (1.0f - f) + sp * f;
        }
        if (!side || wil.z * wo.z <= 0.0f) return zero;
        float g2 = mggx_G2<kFast>(m.awz, fabsf(wil.z), b.a2);
        float spec = mdiv<kFast>(f * d * g2 * 0.25f, m.awz);
        return f3(m.diffuse.x + spec, m.diffuse.y + spec, m.diffuse.z + spec);
    }
    if (b.kind == BK_RGLASS) {
        // the rough dielectric's evaluate (f |cos wi|) and pdfs for a photon
        // direction: a reflection about m = wo + wi, or a refraction through
        // m = -(wi + eta wo) (the generalised half vector), m turned to wo's side
        float eta = b.ior;
        if (wo.z > 0.0f) eta = 1.0f / b.ior;
        const bool refl = wil.z * wo.z > 0.0f;
        F3 mh = refl ? wo + wil : -(wil + wo * eta);
        mh = mnormalize<kFast>(mh);
        if (mh.z * wo.z < 0.0f) mh = -mh;
        const float wom = dot(wo, mh), wim = dot(wil, mh);
        const float f = mfresnel_eta<kFast>(wom, mdiv<kFast>(1.0f, eta));
        const float mz = fabsf(mh.z);
        const float d = mggx_D<kFast>(mz, b.a2);
        const float g2 = mggx_G2<kFast>(m.awz, fabsf(wil.z), b.a2);
        if (refl) {
            fwd = rev = mdiv<kFast>(((f * d) * mz) * 0.25f, wom);
            if (!side || !(wom > 0.0f)) return zero;
            return b.R * mdiv<kFast>(((f * d) * g2) * 0.25f, m.awz);
        }
        const float sq = wim + eta * wom;
        const float den = sq * sq;
        const float t = ((1.0f - f) * d) * mz;
        fwd = mdiv<kFast>(t * fabsf(wim), den);
        rev = mdiv<kFast>(t * ((eta * eta) * wom), den);
        if (!(dot(wi, m.gn) * m.side_wo < 0.0f) || !(wom > 0.0f) || !(wim < 0.0f)) return zero;
        // radiance mode (the merge's camera vertices): the eta^2 of radiance transport
        const float v = mdiv<kFast>(((((1.0f - f) * d) * g2) * (wom * fabsf(wim))) * (eta * eta), m.awz * den);
        return b.R * v;
    }
    if (b.kind == BK_FSMG) {
        float f = m.f;
        if (f > 1.0f) {
            fwd = rev = 0.0f;
        } else {
            fwd = fabsf(wil.z) * kInvPI * (1.0f - f);
            rev = m.rd * (1.0f - f);
        }
        if (!side || wil.z * wo.z <= 0.0f) return zero;
        return m.diffuse;
    }
    fwd = rev = 0.0f;
    return zero;
}

// ------------------------------------------------------------------
// Lights: Lights_list::sample_lights (Light_list.cpp:10-30) and
// Area_Light / Sun_Light / Point_Light::sample_light (Lights.cpp, Lights.h)
// ------------------------------------------------------------------
__device__ __forceinline__ int find_interval(const float* cdf, int size, float u) {
    int first = 0, len = size;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        if (cdf[middle] < u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    int r = first - 1;
    r = r < 0 ? 0 : r;
    r = r > size - 2 ? size - 2 : r;
    return r;
}

// Guided direction of light `light` (rc_guide.h): with probability alpha
// the plain words, otherwise a bin from the CDF of the position cell and a
// uniform point inside it (the bin's index goes into the top five bits of
// each word, the old words' top 19 bits below).  Returns q, the mixture
// density relative to plain sampling; `bin` is the sampled bin either way.
__device__ __forceinline__ float guide_direction(const GuideView& g, int light, float ax, float ay, const U4& w1,
                                                 uint32_t& wx, uint32_t& wy, int& bin) {
    bin = (int)(wx >> 27) | ((int)(wy >> 27) << 5);
    if (!g.cdf) return 1.0f;
    const int G = g.cells;
    int cx = (int)(ax * (float)G), cy = (int)(ay * (float)G);
    cx = cx > G - 1 ? G - 1 : cx;
    cy = cy > G - 1 ? G - 1 : cy;
    const float* cdf = g.cdf + ((size_t)(guide_slot(g, light) * G + cy) * G + cx) * kGuideCdfStride;
    if (u24(w1.b) >= g.alpha) {
        float uc = (float)((w1.c >> 8) + 1u) * 5.9604644775390625e-08f;  // (0, 1]
        int k = find_interval(cdf, kGuideBins + 1, uc);
        wx = ((uint32_t)(k & 31) << 27) | (wx >> 5);
        wy = ((uint32_t)(k >> 5) << 27) | (wy >> 5);
        bin = k;
    }
    float p = cdf[bin + 1] - cdf[bin];
    return g.alpha + (1.0f - g.alpha) * (float)kGuideBins * p;
}

// Returns false when the light emits nothing (the reference point light).
// bin / axis: the direction bin and emission axis of a guidable emission
// (area and sphere lights), bin = -1 otherwise.  The direction sample (block
// 1 draw, optional guided warp, cosine lobe) is one piece of code for both
// guidable kinds.
template <class Rng>
__device__ __forceinline__ bool sample_light(const SceneView& sc, const Rng& rng, const GuideView& gv, F3& o,
                                             F3& d, F3& tp, int& bin, F3& axis) {
    U4 w0 = rng.block(0);
    int idx = find_interval(sc.cdf, sc.nlights + 1, u24(w0.a));
    const DevLight& L = sc.lights[idx];
    float pick = sc.compat_pdf ? L.power / ((float)sc.nlights * sc.integral) : L.power / sc.integral;
    bin = -1;
    axis = f3s(0.0f);  // set on every path: a partly-set output keeps the compiler from promoting it to registers
    const float u1 = u24(w0.b), u2 = u24(w0.c);
    if (L.kind == L_SUN) {
        o = (ld3(L.bl) + ld3(L.h) * u1) + ld3(L.v) * u2;
        d = ld3(L.normal);
        tp = div_each(ld3(L.color), pick);
        return true;
    }
    const bool area = L.kind == L_AREA;
    if (!area && sc.compat_point) return false;
    F3 n = f3s(0.0f);
    if (area) {
        o = (ld3(L.bl) + ld3(L.h) * u1) + ld3(L.v) * u2;
    } else {
        float z = 1.0f - 2.0f * u1;
        float rr = sqrtf(smax(0.0f, 1.0f - z * z));
        float phi = kPI2 * u2;
        n = f3(rr * own_cos(phi), rr * own_sin(phi), z);
        o = L.radius > 0.0f ? ld3(L.pos) + n * L.radius : ld3(L.pos);
    }
    float q = 1.0f;
    if (area || L.radius > 0.0f) {
        U4 w1 = rng.block(1);
        uint32_t wx = w0.d, wy = w1.a;
        if (guide_slot(gv, idx) >= 0) q = guide_direction(gv, idx, u1, u2, w1, wx, wy, bin);
        F3 local = cos_hemisphere(u24(wx), u24(wy));
        Frame fr = area ? Frame{ld3(L.fx), ld3(L.fy), ld3(L.fz)} : frame_from_z(n);
        d = to_world(fr, local);
        axis = area ? ld3(L.fz) : n;
    } else {
        d = n;
    }
    F3 t = area ? ld3(L.color) * kPI : (ld3(L.color) * L.power) * kPI;
    tp = div_each(t, pick);
    if (q != 1.0f) tp = div_each(tp, q);
    return true;
}

// ------------------------------------------------------------------
// Closest hit: binary BVH with both child boxes in the parent node,
// near-child-first descent, stack in LDS (kLdsStack entries per lane,
// stride kBlock) with a private overflow.  Triangle test is
// Triangle::hit (Triangle.cpp:51-70) with strict tmin < t < t_best.
// ------------------------------------------------------------------
struct Hit {
    float t, u, v;
    int tri;
};

__device__ __forceinline__ bool tri_test(const float4* __restrict__ tri, int i, F3 o, F3 d, float tmin, float tbest,
                                         float& t_out, float& u_out, float& v_out) {
    float4 a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
    F3 v0 = f3(a.x, a.y, a.z), e1 = f3(b.x, b.y, b.z), e2 = f3(c.x, c.y, c.z);
    F3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (fabsf(det) < 0.0000001f) return false;
    float inv = 1.0f / det;
    F3 tvec = o - v0;
    float u = dot(tvec, pvec) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    F3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qvec) * inv;
    if (t > tmin && t < tbest) {
        t_out = t;
        u_out = u;
        v_out = v;
        return true;
    }
    return false;
}

template <bool kCount>
__device__ __forceinline__ bool trace_closest2(const SceneView& sc, F3 o, F3 d, float tmin, Hit& h, int2* lds_stack,
                                               uint32_t& n_nodes, uint32_t& n_tris) {
    int2 ext[kMaxDepth - kLdsStack];
    float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    float tbest = INFINITY;
    int best = -1;
    float bu = 0.0f, bv = 0.0f;
    int sp = 0;
    int ref = sc.root;
    const float4* __restrict__ nodes = sc.nodes;
    // Slab far distances are widened by 1 + 2^-19 so rounding in the slab
    // arithmetic never culls a box the ray touches (boxes are also padded
    // outwards when built); NaN slabs (0 * inf) drop out of fminf/fmaxf.
    const float slack = 1.0000019f;
    while (true) {
        if (ref >= 0) {
            if (kCount) ++n_nodes;
            const float4* n = nodes + 4 * ref;
            float4 n0 = n[0], n1 = n[1], n2 = n[2], n3 = n[3];
            float l0x = (n0.x - o.x) * ix, l1x = (n0.y - o.x) * ix;
            float l0y = (n0.z - o.y) * iy, l1y = (n0.w - o.y) * iy;
            float l0z = (n1.x - o.z) * iz, l1z = (n1.y - o.z) * iz;
            float r0x = (n1.z - o.x) * ix, r1x = (n1.w - o.x) * ix;
            float r0y = (n2.x - o.y) * iy, r1y = (n2.y - o.y) * iy;
            float r0z = (n2.z - o.z) * iz, r1z = (n2.w - o.z) * iz;
            float ln = fmaxf(fmaxf(fminf(l0x, l1x), fminf(l0y, l1y)), fmaxf(fminf(l0z, l1z), tmin));
            float lf = fminf(fminf(fmaxf(l0x, l1x), fmaxf(l0y, l1y)), fmaxf(l0z, l1z)) * slack;
            float rn = fmaxf(fmaxf(fminf(r0x, r1x), fminf(r0y, r1y)), fmaxf(fminf(r0z, r1z), tmin));
            float rf = fminf(fminf(fmaxf(r0x, r1x), fmaxf(r0y, r1y)), fmaxf(r0z, r1z)) * slack;
            lf = fminf(lf, tbest);
            rf = fminf(rf, tbest);
            bool hl = ln <= lf, hr = rn <= rf;
            int cl = __float_as_int(n3.x), cr = __float_as_int(n3.y);
            if (hl && hr) {
                int nearc = ln <= rn ? cl : cr, farc = ln <= rn ? cr : cl;
                const int2 e = make_int2(farc, __float_as_int(ln <= rn ? rn : ln));
                if (sp < kLdsStack) lds_stack[sp * kBlock] = e;
                else ext[sp - kLdsStack] = e;
                ++sp;
                ref = nearc;
                continue;
            }
            if (hl || hr) {
                ref = hl ? cl : cr;
                continue;
            }
        } else {
            int enc = ~ref;
            int first = enc >> 3, count = (enc & 7) + 1;
            if (kCount) n_tris += count;
            for (int i = first; i < first + count; ++i) {
                float t, u, v;
                if (tri_test(sc.tri, i, o, d, tmin, tbest, t, u, v)) {
                    tbest = t;
                    best = i;
                    bu = u;
                    bv = v;
                }
            }
        }
        // pop the nearest pending entry whose box can still hold a closer hit
        // (the test a box passes when it is pushed: entry distance <= tbest)
        bool more = false;
        while (sp > 0) {
            --sp;
            const int2 e = sp < kLdsStack ? lds_stack[sp * kBlock] : ext[sp - kLdsStack];
            if (__int_as_float(e.y) <= tbest) {
                ref = e.x;
                more = true;
                break;
            }
        }
        if (!more) break;
    }
    h.t = tbest;
    h.u = bu;
    h.v = bv;
    h.tri = best;
    return best >= 0;
}

// Closest hit through the 4-wide BVH: 8 x float4 per node = the x, y, z
// low / high planes of the four child boxes (one float per child) and the
// four child refs (kEmptyChild where a node has fewer children).  The four
// slab tests are those of the binary traversal; hit children are ordered by
// entry distance (a 5-comparator network), the nearest is taken next and the
// others are pushed far to near.
template <bool kCount>
__device__ __forceinline__ bool trace_closest4(const SceneView& sc, F3 o, F3 d, float tmin, Hit& h, int2* lds_stack,
                                               uint32_t& n_nodes, uint32_t& n_tris) {
    int2 ext[kMaxDepth - kLdsStack];
    float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    float tbest = INFINITY;
    int best = -1;
    float bu = 0.0f, bv = 0.0f;
    int sp = 0;
    int ref = sc.root4;
    const float4* __restrict__ nodes = sc.nodes4;
    const float slack = 1.0000019f;
    auto push = [&](int r, float dist) {
        const int2 e = make_int2(r, __float_as_int(dist));
        if (sp < kLdsStack) lds_stack[sp * kBlock] = e;
        else ext[sp - kLdsStack] = e;
        ++sp;
    };
    while (true) {
        if (ref >= 0) {
            if (kCount) ++n_nodes;
            const float4* n = nodes + 8 * ref;
            const float4 xl = n[0], xh = n[1], yl = n[2], yh = n[3], zl = n[4], zh = n[5];
            const int4 cr = *reinterpret_cast<const int4*>(n + 6);
            const float xlo[4] = {xl.x, xl.y, xl.z, xl.w}, xhi[4] = {xh.x, xh.y, xh.z, xh.w};
            const float ylo[4] = {yl.x, yl.y, yl.z, yl.w}, yhi[4] = {yh.x, yh.y, yh.z, yh.w};
            const float zlo[4] = {zl.x, zl.y, zl.z, zl.w}, zhi[4] = {zh.x, zh.y, zh.z, zh.w};
            const int refs[4] = {cr.x, cr.y, cr.z, cr.w};
            float key[4];
            int cref[4];
            bool hit[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float a0 = (xlo[c] - o.x) * ix, a1 = (xhi[c] - o.x) * ix;
                float b0 = (ylo[c] - o.y) * iy, b1 = (yhi[c] - o.y) * iy;
                float c0 = (zlo[c] - o.z) * iz, c1 = (zhi[c] - o.z) * iz;
                float nr = fmaxf(fmaxf(fminf(a0, a1), fminf(b0, b1)), fmaxf(fminf(c0, c1), tmin));
                float fr = fminf(fminf(fmaxf(a0, a1), fmaxf(b0, b1)), fmaxf(c0, c1)) * slack;
                fr = fminf(fr, tbest);
                hit[c] = nr <= fr && refs[c] != kEmptyChild;
                key[c] = hit[c] ? nr : INFINITY;
                cref[c] = refs[c];
            }
            auto cx = [&](int i, int j) {
                if (key[j] < key[i]) {
                    float tk = key[i];
                    key[i] = key[j];
                    key[j] = tk;
                    int tr = cref[i];
                    cref[i] = cref[j];
                    cref[j] = tr;
                    bool th = hit[i];
                    hit[i] = hit[j];
                    hit[j] = th;
                }
            };
            cx(0, 1);
            cx(2, 3);
            cx(0, 2);
            cx(1, 3);
            cx(1, 2);
            if (hit[0]) {
                if (hit[3]) push(cref[3], key[3]);
                if (hit[2]) push(cref[2], key[2]);
                if (hit[1]) push(cref[1], key[1]);
                ref = cref[0];
                continue;
            }
        } else {
            int enc = ~ref;
            int first = enc >> 3, count = (enc & 7) + 1;
            if (kCount) n_tris += count;
            for (int i = first; i < first + count; ++i) {
                float t, u, v;
                if (tri_test(sc.tri, i, o, d, tmin, tbest, t, u, v)) {
                    tbest = t;
                    best = i;
                    bu = u;
                    bv = v;
                }
            }
        }
        // pop the nearest pending entry whose box can still hold a closer hit
        // (the test a box passes when it is pushed: entry distance <= tbest)
        bool more = false;
        while (sp > 0) {
            --sp;
            const int2 e = sp < kLdsStack ? lds_stack[sp * kBlock] : ext[sp - kLdsStack];
            if (__int_as_float(e.y) <= tbest) {
                ref = e.x;
                more = true;
                break;
            }
        }
        if (!more) break;
    }
    h.t = tbest;
    h.u = bu;
    h.v = bv;
    h.tri = best;
    return best >= 0;
}

template <bool kCount>
__device__ __forceinline__ bool trace_closest(const SceneView& sc, F3 o, F3 d, float tmin, Hit& h, int2* lds_stack,
                                              uint32_t& n_nodes, uint32_t& n_tris) {
    return sc.nodes4 ? trace_closest4<kCount>(sc, o, d, tmin, h, lds_stack, n_nodes, n_tris)
                     : trace_closest2<kCount>(sc, o, d, tmin, h, lds_stack, n_nodes, n_tris);
}

// Hit record of the reference (Triangle.cpp:71-80): geometric normal,
// barycentric (unnormalised) shading normal, position, material.
struct Surface {
    F3 gn, sn, pos;
    int mat;
};
__device__ __forceinline__ Surface surface_at(const SceneView& sc, const Hit& h, F3 o, F3 d) {
    float4 a = sc.tri[3 * h.tri], b = sc.tri[3 * h.tri + 1], c = sc.tri[3 * h.tri + 2];
    F3 e1 = f3(b.x, b.y, b.z), e2 = f3(c.x, c.y, c.z);
    float4 s0 = sc.shade[3 * h.tri], s1 = sc.shade[3 * h.tri + 1], s2 = sc.shade[3 * h.tri + 2];
    Surface s;
    s.gn = normalize(cross(e1, e2));
    s.sn = (f3(s0.x, s0.y, s0.z) * (1.0f - h.u - h.v) + f3(s1.x, s1.y, s1.z) * h.u) + f3(s2.x, s2.y, s2.z) * h.v;
    s.pos = o + d * h.t;
    s.mat = __float_as_int(a.w);
    return s;
}

}  // namespace rc
