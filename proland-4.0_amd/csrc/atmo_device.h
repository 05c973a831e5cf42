// atmo_device.h — device-side data layout and math of the MI355X atmo precompute.
//
// Layout in HBM (one "set" = one atmosphere parameter set):
//   T4   float4[TH][TW]                 transmittance (rgb, w unused)
//   dE4  float4[SH][SW]                 irradiance delta of the current order
//   E4   float4[SH][SW]                 irradiance accumulator
//   dSR4, dSM4, dJ4  float4[RR][RMU][RNU*RMUS]   per-order deltas
//   S4   float4[RR][RMU][RNU*RMUS]      inscatter accumulator == the .raw layout
//   outT, outE  float[.][.][3]          RGB tables in the .raw layouts
// Every table is addressed through a per-set descriptor, so one launch covers
// any number of sets of mixed resolutions.
//
// Precision: the reference computes everything in FP32 GLSL.  Here the
// cancellation-prone geometry (r^2 - Rg^2, the ray/ground discriminant, ray
// lengths, sun-visibility tests, texel (mu, muS, nu)) is FP64 and the bulk
// (table filtering, phase functions, transcendentals, accumulation) is FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atmo {

// common.glsl:99 — the reference's value of pi.
constexpr double kPI = 3.141592657;

struct LayerGeom {
    double r;        // FP32 uniform r of the layer (PreprocessAtmo.cpp:255-271)
    double dhdH[4];  // FP32 uniform dhdH
};

struct SetDesc {
    // FP32 parameters promoted to double
    double Rg, Rt, RL, HR, HM, mieG, alpha;
    double H;            // sqrt(Rt^2 - Rg^2)
    double Rg2, Rt2, RL2;
    double tMu;          // tan(1.5) / 1.15  (getTransmittanceUV, FP64 lookups)
    float bR[3], bMS[3], bME[3];
    // FP32 constants
    float f_Rg, f_dRt, f_inv_dRt;   // Rg, Rt - Rg, 1/(Rt - Rg)
    float f_H, f_invH;
    float f_tMu;                    // tan(1.5) / 1.15  (getTransmittanceUV)
    float f_tMuS;                   // tan(1.26 * 1.1)  (texture4D uMuS)
    float kR;                       // 3 / (16 pi)
    float kM;                       // 1.5 / (4 pi) (1 - g^2) / (2 + g^2)
    float g2p1, m2g;                // 1 + g^2, -2 g
    int TW, TH, SW, SH, RR, RMU, RMUS, RNU;
    int W4;                         // RNU * RMUS
    int n4;                         // RR * RMU * W4
    // tables (device pointers)
    float4 *T4, *dE4, *E4, *dSR4, *dSM4, *dJ4, *S4;
    float4 *tau4;                   // optical depth -ln T of the T texels (rgb), or null
    // transmittances along the view ray of each (layer, mu row) at the inscatter
    // samples, [RR][RMU][nsamp + 1], and from the ground point to each layer
    // radius per theta row of the spherical integral, [RR][NS]; null: computed
    // inline from T (pass-level entry points)
    float4 *tview, *tground;
    int dJb;                        // bytes per dJ texel: 16 (float4) or 12 (packed rgb)
    int hs;                         // 1: table values rounded to FP16 on store (q16)
    int linT, lin4;                 // 1: the linear transmittance / inscatter parameterisation
    float *outT, *outE;
    const LayerGeom *layers;
};

// ---------------------------------------------------------------------------
// GL LINEAR + CLAMP_TO_EDGE along one axis (GL 3.3 core 3.8.11) for a texel
// coordinate x = u * N - 0.5.  Out-of-range coordinates clamp first, which
// yields the same edge texel as GL.
__device__ __forceinline__ void gl_axis(float x, int N, int &i0, int &i1, float &a) {
    x = fminf(fmaxf(x, -1.0f), (float)N);
    float f = floorf(x);
    a = x - f;
    int k = (int)f;
    i0 = min(max(k, 0), N - 1);
    i1 = min(max(k + 1, 0), N - 1);
}

__device__ __forceinline__ void gl_axis_d(double x, int N, int &i0, int &i1, float &a) {
    x = fmin(fmax(x, -1.0), (double)N);
    double f = floor(x);
    a = (float)(x - f);
    int k = (int)f;
    i0 = min(max(k, 0), N - 1);
    i1 = min(max(k + 1, 0), N - 1);
}

__device__ __forceinline__ void gl_axis_dd(double x, int N, int &i0, int &i1, double &a) {
    x = fmin(fmax(x, -1.0), (double)N);
    double f = floor(x);
    a = x - f;
    int k = (int)f;
    i0 = min(max(k, 0), N - 1);
    i1 = min(max(k + 1, 0), N - 1);
}

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }

// Nearest IEEE binary16 value of x, as a float: round to nearest even on the
// 10 fraction bits for |x| >= 2^-14, on the subnormal quantum 2^-24 below,
// +-inf from 65520 on (the FP16 overflow threshold).  Integer and exact
// power-of-two steps only, so it is bit-reproducible.
__device__ __forceinline__ float q16(float x) {
    const float ax = fabsf(x);
    if (!(ax < 65520.0f)) return ax == ax ? copysignf(__builtin_huge_valf(), x) : x;
    if (ax < 6.103515625e-05f) return rintf(x * 16777216.0f) * 5.9604644775390625e-08f;
    unsigned b = __float_as_uint(x);
    b += 0x0FFFu + ((b >> 13) & 1u);
    return __uint_as_float(b & 0xFFFFE000u);
}
__device__ __forceinline__ float4 q16(float4 v, bool on) {
    return on ? make_float4(q16(v.x), q16(v.y), q16(v.z), q16(v.w)) : v;
}

// Table pointers reach the kernels through the set descriptors, so the
// compiler sees generic (flat) pointers; this re-declares them as global
// memory so loads use global_load with a scalar base.
struct F4 {
    float x, y, z, w;
};
typedef __attribute__((address_space(1))) const F4 gfloat4;
__device__ __forceinline__ const gfloat4 *gptr(const float4 *p) { return (const gfloat4 *)p; }
__device__ __forceinline__ float3 ld3(const gfloat4 *p) { return make_float3(p->x, p->y, p->z); }

// Raw buffer loads of a table: the descriptor (base address, size in bytes) sits
// in scalar registers, and a load's byte offset is one VGPR plus one SGPR, so a
// gather needs no 64-bit address arithmetic per load.  The fourth descriptor word
// is the gfx9 one for untyped access (data format 32).  A load past `bytes`
// returns zeros, which no caller relies on.
typedef float buf_f3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float3 ld3b(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    const buf_f3 v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
    return make_float3(v.x, v.y, v.z);
}
__device__ __forceinline__ float3 lerp3(float3 a, float3 b, float t) {
    return make_float3(fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y), fmaf(t, b.z - a.z, a.z));
}
__device__ __forceinline__ float3 operator*(float3 a, float3 b) {
    return make_float3(a.x * b.x, a.y * b.y, a.z * b.z);
}
__device__ __forceinline__ float3 operator*(float3 a, float s) {
    return make_float3(a.x * s, a.y * s, a.z * s);
}
__device__ __forceinline__ float3 operator+(float3 a, float3 b) {
    return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}

// Bilinear fetch from a float4 [H][W] table at texel coordinates (x, y).
__device__ __forceinline__ float3 bilinear(const float4 *__restrict__ tab, int W, int H, float x,
                                          float y) {
    int i0, i1, j0, j1;
    float a, b;
    gl_axis(x, W, i0, i1, a);
    gl_axis(y, H, j0, j1, b);
    const gfloat4 *t = gptr(tab);
    float3 v00 = ld3(t + j0 * W + i0), v01 = ld3(t + j0 * W + i1);
    float3 v10 = ld3(t + j1 * W + i0), v11 = ld3(t + j1 * W + i1);
    return lerp3(lerp3(v00, v01, a), lerp3(v10, v11, a), b);
}

// getTransmittanceUV (common.glsl:114-124): uR from the altitude h = r - Rg
// (computed without cancellation by the caller; below the ground it clamps to
// row 0 in both parameterisations), and uMu.
__device__ __forceinline__ float T_ur(const SetDesc &s, float h) {
    return s.linT ? fmaxf(h, 0.0f) * s.f_inv_dRt : sqrtf(fmaxf(h, 0.0f) * s.f_inv_dRt);
}
template <bool LIN>   // the parameterisation fixed at compile time (inner loops)
__device__ __forceinline__ float T_umu_t(const SetDesc &s, float mu) {
    return LIN ? (mu + 0.15f) * (1.0f / 1.15f) : atanf((mu + 0.15f) * s.f_tMu) * (1.0f / 1.5f);
}
__device__ __forceinline__ float T_umu(const SetDesc &s, float mu) {
    return s.linT ? T_umu_t<true>(s, mu) : T_umu_t<false>(s, mu);
}

// The same T coordinates in FP64 (the sun's lookups of the single-scattering
// pass: next to the horizon T is steep, and an FP32 coordinate's ~3e-8 moves a
// grazing sun's T by ~1e-4).
__device__ __forceinline__ double T_ur_d(const SetDesc &s, double h) {
    const double hc = fmax(h, 0.0), dRt = s.Rt - s.Rg;
    return s.linT ? hc / dRt : sqrt(hc / dRt);
}
template <bool LIN>
__device__ __forceinline__ double T_umu_d(const SetDesc &s, double mu) {
    return LIN ? (mu + 0.15) / 1.15 : atan((mu + 0.15) * s.tMu) / 1.5;
}

// common.glsl:223-226: T(r, mu), bilinear.
__device__ __forceinline__ float3 T_lookup(const SetDesc &s, float h, float mu) {
    return bilinear(s.T4, s.TW, s.TH, T_umu(s, mu) * s.TW - 0.5f, T_ur(s, h) * s.TH - 0.5f);
}

// T(x) / T(y), min 1: the segment transmittance of common.glsl:272-282 for the
// bilinear lookups T(x) = T_lookup(s, hx, mux), T(y) = T_lookup(s, hy, muy).
// T_ratio_lin is the plain FP32 quotient (the single-pass entry points, which
// have no tables of ratios).
__device__ __forceinline__ float3 T_ratio_lin(const SetDesc &s, float hx, float mux, float hy, float muy) {
    const float3 a = T_lookup(s, hx, mux), b = T_lookup(s, hy, muy);
    return make_float3(fminf(a.x / b.x, 1.f), fminf(a.y / b.y, 1.f), fminf(a.z / b.z, 1.f));
}

// T_ratio_d is the form of the precompute's ratio tables (k_tview, k_tground:
// one ratio per layer, mu row and view-ray sample, and per layer and theta
// row, evaluated once per precompute).  Its texel coordinates and filter
// weights are computed in FP64: along a grazing view ray the ratio divides two
// lookups of nearly equal, large optical depth, and an FP32 error of 1e-7 in a
// lookup's mu coordinate moves it by ~1e-3 there.  With the optical-depth
// table the ratio is evaluated as
//   T(x) = e^-m sum_i w_i e^(m - tau_i),  m = the smallest depth among the taps
//   of nonzero weight,
// which equals the bilinear lookup in exact arithmetic but cannot underflow:
// two transmittances far below FP32's range (long grazing paths through dense
// or large atmospheres) still give their finite ratio.  With FP16 storage
// (which keeps the reference's stored T) it is the plain quotient.
__device__ __forceinline__ void T_taps_d(const SetDesc &s, double h, double mu, int idx[4],
                                         float w[4]) {
    const double hc = fmax(h, 0.0), dRt = s.Rt - s.Rg;
    const double ur = s.linT ? hc / dRt : sqrt(hc / dRt);
    const double um = s.linT ? (mu + 0.15) / 1.15 : atan((mu + 0.15) / 1.15 * tan(1.5)) / 1.5;
    int i0, i1, j0, j1;
    double a, b;
    gl_axis_dd(um * s.TW - 0.5, s.TW, i0, i1, a);
    gl_axis_dd(ur * s.TH - 0.5, s.TH, j0, j1, b);
    idx[0] = j0 * s.TW + i0; idx[1] = j0 * s.TW + i1;
    idx[2] = j1 * s.TW + i0; idx[3] = j1 * s.TW + i1;
    w[0] = (float)((1.0 - a) * (1.0 - b)); w[1] = (float)(a * (1.0 - b));
    w[2] = (float)((1.0 - a) * b); w[3] = (float)(a * b);
}
// ln T = -m + ln S of one channel c of a lookup with taps idx, weights w
__device__ __forceinline__ void T_log_channel(const float *tau, const int idx[4], const float w[4],
                                              int c, float &m, float &S) {
    float t[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) t[q] = tau[4 * idx[q] + c];
    m = __builtin_huge_valf();
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (w[q] > 0.0f) m = fminf(m, t[q]);
    S = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (w[q] > 0.0f) S = fmaf(w[q], expf(m - t[q]), S);
}
__device__ __forceinline__ float3 T_ratio_d(const SetDesc &s, double hx, double mux, double hy,
                                            double muy) {
    int ix[4], iy[4];
    float wx[4], wy[4], r[3];
    T_taps_d(s, hx, mux, ix, wx);
    T_taps_d(s, hy, muy, iy, wy);
    if (!s.tau4 || s.hs) {   // the plain quotient of the stored T (FP16 storage)
        const float *T = (const float *)s.T4;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float a = 0.0f, b = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a = fmaf(wx[q], T[4 * ix[q] + c], a);
                b = fmaf(wy[q], T[4 * iy[q] + c], b);
            }
            r[c] = fminf(a / b, 1.0f);
        }
    } else {
        const float *tau = (const float *)s.tau4;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float mx, Sx, my, Sy;
            T_log_channel(tau, ix, wx, c, mx, Sx);
            T_log_channel(tau, iy, wy, c, my, Sy);
            const float lr = (my - mx) + logf(Sx / Sy);
            r[c] = lr >= 0.0f ? 1.0f : expf(lr);
        }
    }
    return make_float3(r[0], r[1], r[2]);
}

// Phase functions, common.glsl:290-297.
__device__ __forceinline__ float phaseR(const SetDesc &s, float mu) {
    return s.kR * fmaf(mu, mu, 1.0f);
}
__device__ __forceinline__ float phaseM(const SetDesc &s, float mu) {
    float x = fmaf(s.m2g, mu, s.g2p1);
    float y = rsqrtf(x);
    return s.kM * fmaf(mu, mu, 1.0f) * (y * y * y);
}

// common.glsl:209-219 limit(r, mu), FP64.
__device__ __forceinline__ double limit_d(const SetDesc &s, double r, double mu) {
    double rr1 = r * r * (mu * mu - 1.0);
    double dout = -r * mu + sqrt(rr1 + s.RL2);
    double delta2 = rr1 + s.Rg2;
    if (delta2 >= 0.0) {
        double din = -r * mu - sqrt(delta2);
        if (din >= 0.0) dout = fmin(dout, din);
    }
    return dout;
}

// common.glsl:175-201 getMuMuSNu for texel (x, y) of a layer, FP64.
__device__ __forceinline__ double texel_mus(const SetDesc &s, int xm);
__device__ __forceinline__ void mu_mus_nu(const SetDesc &s, const LayerGeom &L, int x, int y,
                                          double &mu, double &muS, double &nu) {
    double r = L.r;
    double half = s.RMU * 0.5;
    if (s.lin4) {   // linear parameterisation (common.glsl:194-199)
        mu = -1.0 + 2.0 * y / (s.RMU - 1.0);
        muS = texel_mus(s, x % s.RMUS);
        nu = -1.0 + (double)(x / s.RMUS) / (s.RNU - 1.0) * 2.0;
        return;
    }
    if (y < half) {
        double d = 1.0 - y / (half - 1.0);
        d = fmin(fmax(L.dhdH[2], d * L.dhdH[3]), L.dhdH[3] * 0.999);
        mu = (s.Rg2 - r * r - d * d) / (2.0 * r * d);
        double q = s.Rg / r;
        mu = fmin(mu, -sqrt(1.0 - q * q) - 0.001);
    } else {
        double d = (y - half) / (half - 1.0);
        d = fmin(fmax(L.dhdH[0], d * L.dhdH[1]), L.dhdH[1] * 0.999);
        mu = (s.Rt2 - r * r - d * d) / (2.0 * r * d);
    }
    int xm = x % s.RMUS;
    double m = xm / (s.RMUS - 1.0);
    muS = tan((2.0 * m - 1.0 + 0.26) * 1.1) / tan(1.26 * 1.1);
    nu = -1.0 + (double)(x / s.RMUS) / (s.RNU - 1.0) * 2.0;
}

// muS of a texel column (depends on x only).
__device__ __forceinline__ double texel_mus(const SetDesc &s, int xm) {
    double m = xm / (s.RMUS - 1.0);
    // linear: -0.2 + m 1.2 rounded twice, never fused (a column with muS = 0
    // in exact arithmetic keeps the sign of the two-step rounding)
    if (s.lin4) return __dadd_rn(-0.2, __dmul_rn(m, 1.2));
    return tan((2.0 * m - 1.0 + 0.26) * 1.1) / tan(1.26 * 1.1);
}

// texture4D coordinates (common.glsl:149-173) as texel coordinates:
// tz (r axis), ty (mu axis), txs (mu_s axis inside a nu slice).  FP64 inputs.
// rho2 = r^2 - Rg^2, rmu = r mu, delta = rmu^2 - rho2 (ray/ground discriminant).
__device__ __forceinline__ void tex4d_rmu(const SetDesc &s, double rho2, double rmu,
                                          double delta, double mu, double &tz, double &ty) {
    double rho = sqrt(fmax(rho2, 0.0));
    double f, c;
    if (s.lin4) {   // common.glsl:164: uMu = 0.5/RES_MU + (mu + 1)/2 (1 - 1/RES_MU)
        ty = (0.5 / s.RMU + (mu + 1.0) / 2.0 * (1.0 - 1.0 / s.RMU)) * s.RMU - 0.5;
        tz = (0.5 / s.RR + rho / s.H * (1.0 - 1.0 / s.RR)) * s.RR - 0.5;
        return;
    }
    if (rmu < 0.0 && delta > 0.0) {
        // (rmu + sqrt(delta)) / rho == -rho / (sqrt(delta) - rmu), finite at rho = 0
        f = -rho / (sqrt(delta) - rmu);
        c = 0.5 - 0.5 / s.RMU;
    } else {
        f = (sqrt(fmax(delta + s.H * s.H, 0.0)) - rmu) / (rho + s.H);
        c = 0.5 + 0.5 / s.RMU;
    }
    double uMu = c + f * (0.5 - 1.0 / s.RMU);
    ty = uMu * s.RMU - 0.5;
    double uR = 0.5 / s.RR + rho / s.H * (1.0 - 1.0 / s.RR);
    tz = uR * s.RR - 0.5;
}

__device__ __forceinline__ double tex4d_umus(const SetDesc &s, double muS) {
    if (s.lin4) return 0.5 / s.RMUS + fmax(muS + 0.2, 0.0) / 1.2 * (1.0 - 1.0 / s.RMUS);   // :165
    return 0.5 / s.RMUS +
           (atan(fmax(muS, -0.1975) * tan(1.26 * 1.1)) / 1.1 + (1.0 - 0.26)) * 0.5 *
               (1.0 - 1.0 / s.RMUS);
}

// FP32 uMuS texel offset inside a slice: uMuS * RMUS - 0.5.
template <bool LIN>   // the parameterisation fixed at compile time (inner loops)
__device__ __forceinline__ float tex4d_txs_f(const SetDesc &s, float muS) {
    if (LIN) return fmaxf(muS + 0.2f, 0.0f) * (1.0f / 1.2f) * (float)(s.RMUS - 1);
    float a = atanf(fmaxf(muS, -0.1975f) * s.f_tMuS) * (1.0f / 1.1f) + 0.74f;
    return a * 0.5f * (float)(s.RMUS - 1);
}

}  // namespace atmo
