// atmo_consumer.hip — the table consumer on gfx950: skyRadiance, inScattering,
// sunRadiance and skyIrradiance of src/atmo/examples/atmo/atmosphereShader.glsl
// evaluated on the .raw-layout tables (RGB T and E, RGBA S) with GL LINEAR +
// CLAMP_TO_EDGE filtering.  One thread per ray / point / (r, muS) pair.
//
// Precision follows the preprocess kernels: ray geometry (camera radius, r mu,
// the top-sphere and ground discriminants, the texture4D coordinates) in FP64,
// filtering, phase functions and the radiance combination in FP32.
//
// Every shader constant comes from atmo_params instead of the shader's
// hard-coded Earth (atmosphereShader.glsl:72-125); lengths are in the
// parameters' unit, and the 600 m near-ground offset of inScattering (:472)
// is 600 / meters_per_unit.  Where the shader (built with OPTIMIZE) takes
// sqrt of a negative number for a ray that misses the atmosphere from
// outside, max(NaN, 0) evaluates to 0 on GPUs, so din = 0 here.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "../../include/atmo_consumer.h"
#include "atmo_device.h"

namespace atmo {
atmo_status set_error(atmo_status st, const char *msg);
}

namespace atmo {
namespace {

struct CDesc {
    double Rg, Rt, H, near;       // H = sqrt(Rt^2 - Rg^2); near = Rg + 600 m
    double cam[3], sun[3], shaft;
    float bR[3];
    float kR, kM, g2p1, m2g;      // phase-function constants
    float f_tMu;                  // tan(1.5) / 1.15
    float f_inv_dRt;              // 1 / (Rt - Rg)
    int TW, TH, SW, SH, RR, RMU, RMUS, RNU, W4;
    const float *T, *S, *E;
};

constexpr float kSunIntensity = 100.0f;  // atmosphereShader.glsl:66

typedef __attribute__((address_space(1))) const float gfloat;

// bilinear RGB fetch from a [H][W][3] table at texel coordinates (x, y)
__device__ __forceinline__ float3 rgb_bilinear(const float *tab, int W, int H, float x, float y) {
    int i0, i1, j0, j1;
    float a, b;
    gl_axis(x, W, i0, i1, a);
    gl_axis(y, H, j0, j1, b);
    const gfloat *t = (const gfloat *)tab;
    auto at = [&](int j, int i) {
        const gfloat *q = t + ((size_t)j * W + i) * 3;
        return make_float3(q[0], q[1], q[2]);
    };
    return lerp3(lerp3(at(j0, i0), at(j0, i1), a), lerp3(at(j1, i0), at(j1, i1), a), b);
}

// getTransmittanceUV + transmittance(r, mu) (atmosphereShader.glsl:141-151, 232-235)
__device__ __forceinline__ float3 T_at(const CDesc &d, double r, double mu) {
    const float uR = sqrtf((float)fmax(r - d.Rg, 0.0) * d.f_inv_dRt);
    const float uMu = atanf(((float)mu + 0.15f) * d.f_tMu) * (1.0f / 1.5f);
    return rgb_bilinear(d.T, d.TW, d.TH, uMu * d.TW - 0.5f, uR * d.TH - 0.5f);
}

// trilinear RGBA fetch from S [RR][RMU][W4][4] at texel coordinates
__device__ __forceinline__ float4 rgba_trilinear(const CDesc &d, double x, double y, double z) {
    int i0, i1, j0, j1, k0, k1;
    float a, b, g;
    gl_axis_d(x, d.W4, i0, i1, a);
    gl_axis_d(y, d.RMU, j0, j1, b);
    gl_axis_d(z, d.RR, k0, k1, g);
    const gfloat *t = (const gfloat *)d.S;
    auto at = [&](int k, int j, int i) {
        const gfloat *q = t + (((size_t)k * d.RMU + j) * d.W4 + i) * 4;
        return make_float4(q[0], q[1], q[2], q[3]);
    };
    auto l4 = [](float4 u, float4 v, float w) {
        return make_float4(fmaf(w, v.x - u.x, u.x), fmaf(w, v.y - u.y, u.y), fmaf(w, v.z - u.z, u.z),
                           fmaf(w, v.w - u.w, u.w));
    };
    const float4 c0 = l4(l4(at(k0, j0, i0), at(k0, j0, i1), a), l4(at(k0, j1, i0), at(k0, j1, i1), a), b);
    const float4 c1 = l4(l4(at(k1, j0, i0), at(k1, j0, i1), a), l4(at(k1, j1, i0), at(k1, j1, i1), a), b);
    return l4(c0, c1, g);
}

// texture4D (atmosphereShader.glsl:175-199), INSCATTER_NON_LINEAR; the ground
// branch of uMu in the cancellation-free form -rho / (sqrt(delta) - r mu)
__device__ float4 texture4D(const CDesc &d, double r, double mu, double muS, double nu) {
    const double rho2 = r * r - d.Rg * d.Rg;
    const double rho = sqrt(fmax(rho2, 0.0));
    const double rmu = r * mu;
    const double delta = rmu * rmu - rho2;
    double f, c;
    if (rmu < 0.0 && delta > 0.0) {
        f = -rho / (sqrt(delta) - rmu);
        c = 0.5 - 0.5 / d.RMU;
    } else {
        f = (sqrt(fmax(delta + d.H * d.H, 0.0)) - rmu) / (rho + d.H);
        c = 0.5 + 0.5 / d.RMU;
    }
    const double uMu = c + f * (0.5 - 1.0 / d.RMU);
    const double uR = 0.5 / d.RR + rho / d.H * (1.0 - 1.0 / d.RR);
    const double uMuS = 0.5 / d.RMUS + (atan(fmax(muS, -0.1975) * tan(1.26 * 1.1)) / 1.1 + (1.0 - 0.26)) *
                                           0.5 * (1.0 - 1.0 / d.RMUS);
    double lerp = (nu + 1.0) / 2.0 * (d.RNU - 1.0);
    const double uNu = floor(lerp);
    lerp -= uNu;
    const double y = uMu * d.RMU - 0.5, z = uR * d.RR - 0.5;
    const float4 a = rgba_trilinear(d, (uNu + uMuS) / d.RNU * d.W4 - 0.5, y, z);
    const float4 b = rgba_trilinear(d, (uNu + uMuS + 1.0) / d.RNU * d.W4 - 0.5, y, z);
    const float w = (float)lerp;
    return make_float4(fmaf(w, b.x - a.x, a.x), fmaf(w, b.y - a.y, a.y), fmaf(w, b.z - a.z, a.z),
                       fmaf(w, b.w - a.w, a.w));
}

__host__ __device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// rgb of inScatter * phaseR(nu) + getMie(inScatter) * phaseM(nu), times SUN_INTENSITY
// (:300-302 getMie, :290-297 phases)
__device__ __forceinline__ float4 combine(const CDesc &d, float4 is, double nu) {
    const float n = (float)nu;
    const float pr = d.kR * fmaf(n, n, 1.0f);
    const float y = rsqrtf(fmaf(d.m2g, n, d.g2p1));
    const float pm = d.kM * fmaf(n, n, 1.0f) * (y * y * y);
    const float k = is.w / fmaxf(is.x, 1e-4f);
    const float m0 = is.x * k, m1 = is.y * k * (d.bR[0] / d.bR[1]), m2 = is.z * k * (d.bR[0] / d.bR[2]);
    return make_float4(fmaf(is.x, pr, m0 * pm) * kSunIntensity, fmaf(is.y, pr, m1 * pm) * kSunIntensity,
                       fmaf(is.z, pr, m2 * pm) * kSunIntensity, 0.f);
}

// transmittance between points A and B of a ray, as the shaders write it:
// up ? T(A)/T(B) : T(B reversed)/T(A reversed), clamped to 1
__device__ __forceinline__ float3 T_ratio(const CDesc &d, double ra, double mua, double rb, double mub,
                                         bool up) {
    float3 a, b;
    if (up) {
        a = T_at(d, ra, mua);
        b = T_at(d, rb, mub);
    } else {
        a = T_at(d, rb, -mub);
        b = T_at(d, ra, -mua);
    }
    return make_float3(fminf(a.x / b.x, 1.f), fminf(a.y / b.y, 1.f), fminf(a.z / b.z, 1.f));
}

// skyRadiance, atmosphereShader.glsl:371-418
__global__ __launch_bounds__(256) void k_sky_radiance(CDesc d, const float4 *__restrict__ dirs, int n,
                                                      float4 *__restrict__ rad, float4 *__restrict__ ext) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v4 = dirs[i];
    const double v[3] = {v4.x, v4.y, v4.z};
    double cam[3];
    for (int c = 0; c < 3; ++c) cam[c] = d.cam[c] + v[c] * fmax(d.shaft, 0.0);
    double r = sqrt(dot3(cam, cam));
    double rMu = dot3(cam, v);
    double mu = rMu / r;
    const double r0 = r, mu0 = mu;
    const double q = rMu * rMu - r * r + d.Rt * d.Rt;
    const double din = q >= 0.0 ? fmax(-rMu - sqrt(q), 0.0) : 0.0;
    if (din > 0.0) {
        for (int c = 0; c < 3; ++c) cam[c] += din * v[c];
        rMu += din;
        mu = rMu / d.Rt;
        r = d.Rt;
    }
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f), ex = make_float4(1.f, 1.f, 1.f, 0.f);
    if (r <= d.Rt) {
        const double nu = dot3(v, d.sun);
        const double muS = dot3(cam, d.sun) / r;
        float4 is = texture4D(d, r, rMu / r, muS, nu);
        if (d.shaft > 0.0) {
            const float3 f = T_ratio(d, r0, mu0, r, mu, mu > 0.0);
            is = make_float4(is.x * f.x, is.y * f.y, is.z * f.z, is.w * f.x);
        }
        const float3 e = T_at(d, r, mu);
        ex = make_float4(e.x, e.y, e.z, 0.f);
        out = combine(d, is, nu);
    }
    rad[i] = out;
    if (ext) ext[i] = ex;
}

// inScattering, atmosphereShader.glsl:425-533 (HORIZON_HACK on)
__global__ __launch_bounds__(256) void k_in_scattering(CDesc d, const float4 *__restrict__ pts, int n,
                                                       float4 *__restrict__ rad,
                                                       float4 *__restrict__ ext) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p4 = pts[i];
    double cam[3] = {d.cam[0], d.cam[1], d.cam[2]};
    double pt[3] = {p4.x, p4.y, p4.z};
    double v[3] = {pt[0] - cam[0], pt[1] - cam[1], pt[2] - cam[2]};
    double dist = sqrt(dot3(v, v));
    for (int c = 0; c < 3; ++c) v[c] /= dist;
    double r = sqrt(dot3(cam, cam));
    if (r < 0.9 * d.Rg) {  // flat-terrain coordinates
        cam[2] += d.Rg;
        pt[2] += d.Rg;
        r = sqrt(dot3(cam, cam));
    }
    double rMu = dot3(cam, v);
    double mu = rMu / r;
    const double sw = fmin(fmax(d.shaft, 0.0), dist);
    for (int c = 0; c < 3; ++c) pt[c] -= v[c] * sw;
    const double q = rMu * rMu - r * r + d.Rt * d.Rt;
    const double din = q >= 0.0 ? fmax(-rMu - sqrt(q), 0.0) : 0.0;
    if (din > 0.0 && din < dist) {
        for (int c = 0; c < 3; ++c) cam[c] += din * v[c];
        rMu += din;
        mu = rMu / d.Rt;
        r = d.Rt;
        dist -= din;
    }
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f), ex = make_float4(1.f, 1.f, 1.f, 0.f);
    if (r <= d.Rt) {
        const double nu = dot3(v, d.sun);
        const double muS = dot3(cam, d.sun) / r;
        if (r < d.near) {
            const double f = d.near / r;
            r *= f;
            rMu *= f;
            for (int c = 0; c < 3; ++c) pt[c] *= f;
        }
        double r1 = sqrt(dot3(pt, pt));
        double mu1 = dot3(pt, v) / r1;
        const double muS1 = dot3(pt, d.sun) / r1;
        const float3 e = T_ratio(d, r, mu, r1, mu1, mu > 0.0);
        ex = make_float4(e.x, e.y, e.z, 0.f);
        auto diff = [&](double m, double rr1, double mm1) {
            const float4 s0 = texture4D(d, r, m, muS, nu);
            const float4 s1 = texture4D(d, rr1, mm1, muS1, nu);
            return make_float4(fmaxf(s0.x - s1.x * e.x, 0.f), fmaxf(s0.y - s1.y * e.y, 0.f),
                               fmaxf(s0.z - s1.z * e.z, 0.f), fmaxf(s0.w - s1.w * e.x, 0.f));
        };
        float4 is;
        const double EPS = 0.004;
        const double qg = d.Rg / r;
        const double lim = -sqrt(1.0 - qg * qg);
        if (fabs(mu - lim) < EPS) {
            const float w = (float)(((mu - lim) + EPS) / (2.0 * EPS));
            double m = lim - EPS;
            r1 = sqrt(r * r + dist * dist + 2.0 * r * dist * m);
            mu1 = (r * m + dist) / r1;
            const float4 A = diff(m, r1, mu1);
            m = lim + EPS;
            r1 = sqrt(r * r + dist * dist + 2.0 * r * dist * m);
            mu1 = (r * m + dist) / r1;
            const float4 B = diff(m, r1, mu1);
            is = make_float4(fmaf(w, B.x - A.x, A.x), fmaf(w, B.y - A.y, A.y), fmaf(w, B.z - A.z, A.z),
                             fmaf(w, B.w - A.w, A.w));
        } else {
            is = diff(mu, r1, mu1);
        }
        const float t = fminf(fmaxf((float)muS / 0.02f, 0.f), 1.f);  // smoothstep(0, 0.02, muS)
        is.w *= t * t * (3.f - 2.f * t);
        out = combine(d, is, nu);
    }
    rad[i] = out;
    if (ext) ext[i] = ex;
}

// sunRadiance (:342-350, transmittanceWithShadow :263-265) and skyIrradiance (:355-361)
__global__ __launch_bounds__(256) void k_sun_sky(CDesc d, const float2 *__restrict__ rmus, int n,
                                                 float4 *__restrict__ sun, float4 *__restrict__ sky) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = rmus[i].x, muS = rmus[i].y;
    const double qg = d.Rg / r;
    float3 s = make_float3(0.f, 0.f, 0.f);
    if (!(muS < -sqrt(1.0 - qg * qg))) s = T_at(d, r, muS);
    sun[i] = make_float4(s.x * kSunIntensity, s.y * kSunIntensity, s.z * kSunIntensity, 0.f);
    const float uR = (float)((r - d.Rg) / (d.Rt - d.Rg));
    const float uMuS = ((float)muS + 0.2f) / 1.2f;
    const float3 e = rgb_bilinear(d.E, d.SW, d.SH, uMuS * d.SW - 0.5f, uR * d.SH - 0.5f);
    sky[i] = make_float4(e.x * kSunIntensity, e.y * kSunIntensity, e.z * kSunIntensity, 0.f);
}

atmo_status cfail(atmo_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return atmo::set_error(st, buf);
}

bool al(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

atmo_status make_desc(const atmo_params *p, const atmo_view *v, CDesc &d) {
    if (!p) return cfail(ATMO_ERR_INVALID_ARGUMENT, "null parameters");
    if (!(p->Rg > 0.f) || !(p->Rt > p->Rg)) return cfail(ATMO_ERR_INVALID_ARGUMENT, "need 0 < Rg < Rt");
    if (p->TRANSMITTANCE_W < 2 || p->TRANSMITTANCE_H < 2 || p->SKY_W < 2 || p->SKY_H < 2 || p->RES_R < 2 ||
        p->RES_MU < 2 || p->RES_MU_S < 2 || p->RES_NU < 2)
        return cfail(ATMO_ERR_INVALID_ARGUMENT, "table sizes must be >= 2");
    if ((long long)p->RES_R * p->RES_MU * p->RES_MU_S * p->RES_NU > (1LL << 28))
        return cfail(ATMO_ERR_INVALID_ARGUMENT, "inscatter table too large");
    if (!(p->betaR.x > 0.f && p->betaR.y > 0.f && p->betaR.z > 0.f))
        return cfail(ATMO_ERR_INVALID_ARGUMENT, "getMie divides by betaR: need betaR > 0");
    if (!(p->mieG > -1.f && p->mieG < 1.f)) return cfail(ATMO_ERR_INVALID_ARGUMENT, "mieG must be in (-1, 1)");
    d.Rg = p->Rg;
    d.Rt = p->Rt;
    d.H = sqrt(d.Rt * d.Rt - d.Rg * d.Rg);
    d.bR[0] = p->betaR.x;
    d.bR[1] = p->betaR.y;
    d.bR[2] = p->betaR.z;
    const double g = p->mieG, pi = 3.141592657;  // M_PI of atmosphereShader.glsl:99
    d.kR = (float)(3.0 / (16.0 * pi));
    d.kM = (float)(1.5 / (4.0 * pi) * (1.0 - g * g) / (2.0 + g * g));
    d.g2p1 = (float)(1.0 + g * g);
    d.m2g = (float)(-2.0 * g);
    d.f_tMu = (float)(tan(1.5) / 1.15);
    d.f_inv_dRt = (float)(1.0 / (d.Rt - d.Rg));
    d.TW = p->TRANSMITTANCE_W;
    d.TH = p->TRANSMITTANCE_H;
    d.SW = p->SKY_W;
    d.SH = p->SKY_H;
    d.RR = p->RES_R;
    d.RMU = p->RES_MU;
    d.RMUS = p->RES_MU_S;
    d.RNU = p->RES_NU;
    d.W4 = p->RES_MU_S * p->RES_NU;
    d.near = d.Rg;
    d.shaft = 0.0;
    for (int c = 0; c < 3; ++c) d.cam[c] = d.sun[c] = 0.0;
    if (v) {
        if (!(v->meters_per_unit > 0.0)) return cfail(ATMO_ERR_INVALID_ARGUMENT, "meters_per_unit must be > 0");
        for (int c = 0; c < 3; ++c) {
            d.cam[c] = v->camera[c];
            d.sun[c] = v->sun[c];
        }
        if (!(dot3(d.cam, d.cam) > 0.0)) return cfail(ATMO_ERR_INVALID_ARGUMENT, "camera at the planet centre");
        d.shaft = v->shaft_width;
        d.near = d.Rg + 600.0 / v->meters_per_unit;
    }
    d.T = d.S = d.E = nullptr;
    return ATMO_OK;
}

}  // namespace
}  // namespace atmo

using atmo::CDesc;

extern "C" {

atmo_status atmo_sky_radiance(const atmo_params *p, const float *T, const float *S, const atmo_view *view,
                              const float *dirs, int n, float *radiance, float *extinction, void *stream) {
    CDesc d;
    atmo_status st = atmo::make_desc(p, view, d);
    if (st != ATMO_OK) return st;
    if (!view) return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "null view");
    if (n < 0) return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "negative count");
    if (n == 0) return ATMO_OK;
    if (!T || !S || !dirs || !radiance || !atmo::al(T, 4) || !atmo::al(S, 16) || !atmo::al(dirs, 16) ||
        !atmo::al(radiance, 16) || (extinction && !atmo::al(extinction, 16)))
        return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "bad table or array pointers");
    d.T = T;
    d.S = S;
    hipLaunchKernelGGL(atmo::k_sky_radiance, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d,
                       (const float4 *)dirs, n, (float4 *)radiance, (float4 *)extinction);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return atmo::cfail(ATMO_ERR_DEVICE, "k_sky_radiance: %s", hipGetErrorString(e));
    return ATMO_OK;
}

atmo_status atmo_in_scattering(const atmo_params *p, const float *T, const float *S, const atmo_view *view,
                               const float *points, int n, float *radiance, float *extinction, void *stream) {
    CDesc d;
    atmo_status st = atmo::make_desc(p, view, d);
    if (st != ATMO_OK) return st;
    if (!view) return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "null view");
    if (n < 0) return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "negative count");
    if (n == 0) return ATMO_OK;
    if (!T || !S || !points || !radiance || !atmo::al(T, 4) || !atmo::al(S, 16) || !atmo::al(points, 16) ||
        !atmo::al(radiance, 16) || (extinction && !atmo::al(extinction, 16)))
        return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "bad table or array pointers");
    d.T = T;
    d.S = S;
    hipLaunchKernelGGL(atmo::k_in_scattering, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d,
                       (const float4 *)points, n, (float4 *)radiance, (float4 *)extinction);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return atmo::cfail(ATMO_ERR_DEVICE, "k_in_scattering: %s", hipGetErrorString(e));
    return ATMO_OK;
}

atmo_status atmo_sun_sky_irradiance(const atmo_params *p, const float *T, const float *E, const float *rmus,
                                    int n, float *sun, float *sky, void *stream) {
    CDesc d;
    atmo_status st = atmo::make_desc(p, nullptr, d);
    if (st != ATMO_OK) return st;
    if (n < 0) return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "negative count");
    if (n == 0) return ATMO_OK;
    if (!T || !E || !rmus || !sun || !sky || !atmo::al(T, 4) || !atmo::al(E, 4) || !atmo::al(rmus, 8) ||
        !atmo::al(sun, 16) || !atmo::al(sky, 16))
        return atmo::cfail(ATMO_ERR_INVALID_ARGUMENT, "bad table or array pointers");
    d.T = T;
    d.E = E;
    hipLaunchKernelGGL(atmo::k_sun_sky, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, d,
                       (const float2 *)rmus, n, (float4 *)sun, (float4 *)sky);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return atmo::cfail(ATMO_ERR_DEVICE, "k_sun_sky: %s", hipGetErrorString(e));
    return ATMO_OK;
}

}  // extern "C"
