// atmo_kernels.hip — gfx950 kernels of the atmo precompute, one per reference pass.
//
// Every kernel is launched once per pass for a whole batch of parameter sets:
// blockIdx.x indexes a tile table of (set, tile) pairs built by the host plan,
// so sets of different resolutions share one launch.
#include "atmo_kernels.h"

namespace atmo {

// ---------------------------------------------------------------------------
// Pass T — transmittance.glsl:78-98.  FP64 trapezoid of exp(-h/H) for both
// scale heights at once; 16 lanes share one texel's samples.
// tile = 16 consecutive texels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_transmittance(const SetDesc *__restrict__ sets,
                                                       const uint2 *__restrict__ tiles,
                                                       int nsamp) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int lane = threadIdx.x & 15;
    const int t = (int)tl.y * 16 + (threadIdx.x >> 4);
    const bool valid = t < s.TW * s.TH;
    const int tt = valid ? t : 0;
    const int i = tt % s.TW, j = tt / s.TW;
    // getTransmittanceRMu (common.glsl:126-136), gl_FragCoord = (i + .5, j + .5)
    double rr = (j + 0.5) / s.TH;
    double m = (i + 0.5) / s.TW;
    double r = s.Rg + rr * rr * (s.Rt - s.Rg);
    double mu = -0.15 + tan(1.5 * m) / tan(1.5) * (1.0 + 0.15);
    double dx = limit_d(s, r, mu) / nsamp;
    double sR = 0.0, sM = 0.0;
    const double iHR = 1.0 / s.HR, iHM = 1.0 / s.HM;
    for (int k = lane; k <= nsamp; k += 16) {
        double x = k * dx;
        double h = sqrt(r * r + x * x + 2.0 * x * r * mu) - s.Rg;
        double w = (k == 0 || k == nsamp) ? 0.5 : 1.0;
        sR += w * exp(-h * iHR);
        sM += w * exp(-h * iHM);
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
        sR += __shfl_xor(sR, off, 16);
        sM += __shfl_xor(sM, off, 16);
    }
    if (lane == 0 && valid) {
        double q = s.Rg / r;
        bool below = mu < -sqrt(1.0 - q * q);
        double dR = below ? 1e9 : sR * dx;
        double dM = below ? 1e9 : sM * dx;
        float4 o;
        o.x = (float)exp(-((double)s.bR[0] * dR + (double)s.bME[0] * dM));
        o.y = (float)exp(-((double)s.bR[1] * dR + (double)s.bME[1] * dM));
        o.z = (float)exp(-((double)s.bR[2] * dR + (double)s.bME[2] * dM));
        o.w = 0.0f;
        s.T4[t] = o;
        if (s.outT) {
            s.outT[3 * t + 0] = o.x;
            s.outT[3 * t + 1] = o.y;
            s.outT[3 * t + 2] = o.z;
        }
    }
}

// ---------------------------------------------------------------------------
// Pass dE1 — irradiance1.glsl:78-82, fused with copyIrradiance k = 0
// (PreprocessAtmo.cpp:298-303: E = 0).  tile = 256 texels.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_irradiance1(const SetDesc *__restrict__ sets,
                                                     const uint2 *__restrict__ tiles) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y * 256 + threadIdx.x;
    if (t >= s.SW * s.SH) return;
    const int i = t % s.SW, j = t / s.SW;
    // getIrradianceRMuS (common.glsl:144-147): r - Rg = j/(SH-1) (Rt - Rg)
    double h = (double)j / (s.SH - 1.0) * (s.Rt - s.Rg);
    double muS = -0.2 + (double)i / (s.SW - 1.0) * (1.0 + 0.2);
    float3 T = T_lookup(s, (float)h, (float)muS);
    float c = (float)fmax(muS, 0.0);
    if (s.dE4) s.dE4[t] = make_float4(T.x * c, T.y * c, T.z * c, 0.0f);
    if (s.E4) s.E4[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (s.outE) {
        s.outE[3 * t + 0] = 0.0f;
        s.outE[3 * t + 1] = 0.0f;
        s.outE[3 * t + 2] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Pass dS1 — inscatter1.glsl:104-150 fused with copyInscatter1.glsl:104-109.
// One thread per 4D texel; tile = 256 texels of one set.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_inscatter1(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles,
                                                    int nsamp) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y * 256 + threadIdx.x;
    if (t >= s.n4) return;
    const int x = t % s.W4;
    const int y = (t / s.W4) % s.RMU;
    const int l = t / (s.W4 * s.RMU);
    const LayerGeom L = s.layers[l];
    double mu, muS, nu;
    mu_mus_nu(s, L, x, y, mu, muS, nu);
    const double r = L.r;
    const double rmu = r * mu;
    const double rho2 = r * r - s.Rg2;
    const double dx = limit_d(s, r, mu) / nsamp;
    const bool up = mu > 0.0;
    const float3 Tfix = T_lookup(s, (float)(r - s.Rg), (float)(up ? mu : -mu));
    const double muSr = muS * r;
    const float iHR = (float)(1.0 / s.HR), iHM = (float)(1.0 / s.HM);
    float3 aR = make_float3(0.f, 0.f, 0.f), aM = make_float3(0.f, 0.f, 0.f);
    for (int i = 0; i <= nsamp; ++i) {
        const double ti = i * dx;
        const double q = rho2 + ti * (ti + 2.0 * rmu);  // ri^2 - Rg^2
        const double num = fma(nu, ti, muSr);          // ri * muSi
        // muSi >= -sqrt(1 - Rg^2 / max(Rg, ri)^2)  <=>  num >= 0 or num^2 <= ri^2 - Rg^2
        if (!(num >= 0.0 || num * num <= q)) continue;
        const float ri = sqrtf((float)(q + s.Rg2));
        const float h1 = (float)q / (ri + s.f_Rg);      // ri - Rg (may be < 0 at a ground hit)
        const float mu1 = (float)(rmu + ti) / ri;
        float3 Tseg;
        if (up) {
            float3 b = T_lookup(s, h1, mu1);
            Tseg = make_float3(fminf(Tfix.x / b.x, 1.f), fminf(Tfix.y / b.y, 1.f),
                               fminf(Tfix.z / b.z, 1.f));
        } else {
            float3 a = T_lookup(s, h1, -mu1);
            Tseg = make_float3(fminf(a.x / Tfix.x, 1.f), fminf(a.y / Tfix.y, 1.f),
                               fminf(a.z / Tfix.z, 1.f));
        }
        const float hc = fmaxf(h1, 0.0f);              // ri = max(Rg, ri)
        const float muSi = (float)num / ri;
        const float3 Ts = T_lookup(s, hc, muSi);
        const float w = (i == 0 || i == nsamp) ? 0.5f : 1.0f;
        const float eR = w * __expf(-hc * iHR), eM = w * __expf(-hc * iHM);
        const float3 tt = Tseg * Ts;
        aR = aR + tt * eR;
        aM = aM + tt * eM;
    }
    const float fdx = (float)dx;
    float4 oR = make_float4(aR.x * fdx * s.bR[0], aR.y * fdx * s.bR[1], aR.z * fdx * s.bR[2], 0.f);
    float4 oM = make_float4(aM.x * fdx * s.bMS[0], aM.y * fdx * s.bMS[1], aM.z * fdx * s.bMS[2], 0.f);
    s.dSR4[t] = oR;
    s.dSM4[t] = oM;
    if (s.S4) s.S4[t] = make_float4(oR.x, oR.y, oR.z, oM.x);
}

// ---------------------------------------------------------------------------
// Shared staging of the spherical integrals (inscatterS, irradianceN).
//
// For a fixed r and mu_s, texture4D(table, r, cos(theta), mu_s, nu) is a
// linear interpolation in nu between RNU + 2 "slice values" (slices -1 and
// RNU are the CLAMP_TO_EDGE columns), each slice value being the trilinear
// filter of the 8 texels around (r, cos theta, mu_s).  Trilinear filtering is
// separable, so the 4D lookup of every direction (theta, phi) reduces to a
// 2-entry lerp of LDS data once those slice values are staged.
//
// Staged pair layout: P[it][kk] = {A = V[kk-1], D = V[kk] - V[kk-1]} for
// kk = 0..RNU, so lookup(nu) = A + f D with pos = (nu+1)/2 (RNU-1) + 1,
// kk = floor(pos), f = pos - kk.
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void stage_slice(const SetDesc &s, const float4 *const *tabs,
                                            double tz, double ty, double txs, int k,
                                            float4 *out) {
    int z0, z1, y0, y1, x0, x1;
    float az, ay, ax;
    gl_axis_d(tz, s.RR, z0, z1, az);
    gl_axis_d(ty, s.RMU, y0, y1, ay);
    gl_axis_d((double)k * s.RMUS + txs, s.W4, x0, x1, ax);
    const size_t W = s.W4, HW = (size_t)s.RMU * s.W4;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const float4 *tb = tabs[n];
        float3 c000 = f3(tb[z0 * HW + y0 * W + x0]), c001 = f3(tb[z0 * HW + y0 * W + x1]);
        float3 c010 = f3(tb[z0 * HW + y1 * W + x0]), c011 = f3(tb[z0 * HW + y1 * W + x1]);
        float3 c100 = f3(tb[z1 * HW + y0 * W + x0]), c101 = f3(tb[z1 * HW + y0 * W + x1]);
        float3 c110 = f3(tb[z1 * HW + y1 * W + x0]), c111 = f3(tb[z1 * HW + y1 * W + x1]);
        float3 v = lerp3(lerp3(lerp3(c000, c001, ax), lerp3(c010, c011, ax), ay),
                         lerp3(lerp3(c100, c101, ax), lerp3(c110, c111, ax), ay), az);
        out[n] = make_float4(v.x, v.y, v.z, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Pass dJn — inscatterS.glsl:111-189.  One workgroup per (set, r layer,
// mu_s column); its RMU x RNU texels share r and mu_s, hence the staged
// slice values.  FIRST = order 2 (Rayleigh and Mie stored separately).
// ---------------------------------------------------------------------------
struct DirJ {      // per-theta constants of the 16x32 (NSx2NS) grid
    float ct, st, dw, pad;
};

template <bool FIRST>
__global__ __launch_bounds__(256) void k_inscatterS(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles, int NS,
                                                    const DirJ *__restrict__ dirT,
                                                    const float2 *__restrict__ dirW) {
    constexpr int NT = FIRST ? 2 : 1;
    extern __shared__ float4 lds[];
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int l = (int)tl.y / s.RMUS;
    const int m = (int)tl.y % s.RMUS;
    const int NU = s.RNU;
    const int NP = NU + 1;                 // pairs per theta
    const int NV = NU + 2;                 // slice values per theta
    float4 *pairs = lds;                                   // [NS][NP][NT][2]
    float4 *vtmp = pairs + (size_t)NS * NP * NT * 2;       // [NS][NV][NT]
    float4 *epair = vtmp + (size_t)NS * NV * NT;           // [SW][2]
    float4 *gdat = epair + (size_t)s.SW * 2;               // [NS][2]

    const LayerGeom L = s.layers[l];
    const double r = fmin(fmax(L.r, s.Rg), s.Rt);          // inscatterS.glsl:112
    double muS = fmin(fmax(texel_mus(s, m), -1.0), 1.0);   // :114
    const double rho2 = r * r - s.Rg2;
    const double txs = tex4d_umus(s, muS) * s.RMUS - 0.5;
    const float4 *tabs[2] = {s.dSR4, s.dSM4};

    // (1) slice values V[it][kk], kk = slice + 1
    for (int e = threadIdx.x; e < NS * NV; e += blockDim.x) {
        const int it = e / NV, kk = e % NV;
        const double ct = cos((it + 0.5) * (kPI / NS));
        const double rmu = r * ct;
        double tz, ty;
        tex4d_rmu(s, rho2, rmu, rmu * rmu - rho2, tz, ty);
        float4 v[NT];
        stage_slice<NT>(s, tabs, tz, ty, txs, kk - 1, v);
        if (FIRST) {  // fold the phase-function constants in
            v[0].x *= s.kR; v[0].y *= s.kR; v[0].z *= s.kR;
            v[NT - 1].x *= s.kM; v[NT - 1].y *= s.kM; v[NT - 1].z *= s.kM;
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) vtmp[(size_t)e * NT + n] = v[n];
    }
    // (2) row 0 of dE (irradiance at r = Rg reads only row 0 under CLAMP_TO_EDGE)
    for (int k = threadIdx.x; k < s.SW; k += blockDim.x) {
        float4 a = s.dE4[k];
        float4 b = s.dE4[min(k + 1, s.SW - 1)];
        epair[2 * k] = a;
        epair[2 * k + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
    }
    // (3) ground visibility and ground-to-x transmittance per theta (:129-138)
    for (int it = threadIdx.x; it < NS; it += blockDim.x) {
        const double ct = cos((it + 0.5) * (kPI / NS));
        const double q = s.Rg / r;
        const double cthetamin = -sqrt(1.0 - q * q);
        float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f), g1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ct < cthetamin) {
            const double dg = -r * ct - sqrt(r * r * (ct * ct - 1.0) + s.Rg2);
            const double mug = -(r * ct + dg) / s.Rg;
            // transmittance(Rg, mug, dg) (common.glsl:272-282); its end point is x itself
            const double r1sq = s.Rg2 + dg * dg + 2.0 * s.Rg * mug * dg;
            const double r1 = sqrt(r1sq);
            const double mu1 = (s.Rg * mug + dg) / r1;
            const float h1 = (float)((r1sq - s.Rg2) / (r1 + s.Rg));
            float3 tr;
            if (mug > 0.0) {
                float3 a = T_lookup(s, 0.0f, (float)mug), b = T_lookup(s, h1, (float)mu1);
                tr = make_float3(fminf(a.x / b.x, 1.f), fminf(a.y / b.y, 1.f), fminf(a.z / b.z, 1.f));
            } else {
                float3 a = T_lookup(s, h1, (float)-mu1), b = T_lookup(s, 0.0f, (float)-mug);
                tr = make_float3(fminf(a.x / b.x, 1.f), fminf(a.y / b.y, 1.f), fminf(a.z / b.z, 1.f));
            }
            const float gr = (float)(s.alpha / kPI);
            g0 = make_float4(gr * tr.x, gr * tr.y, gr * tr.z, 1.0f);
            // E texel coordinate: ((r muS + dg nu1)/Rg + 0.2)/1.2 * SW - 0.5 = c1 nu1 + c0
            const double sc = s.SW / 1.2;
            g1 = make_float4((float)(dg / s.Rg * sc), (float)((r * muS / s.Rg + 0.2) * sc - 0.5),
                             0.f, 0.f);
        }
        gdat[2 * it] = g0;
        gdat[2 * it + 1] = g1;
    }
    __syncthreads();
    // (4) pairs {V[kk-1], V[kk] - V[kk-1]}
    for (int e = threadIdx.x; e < NS * NP * NT; e += blockDim.x) {
        const int n = e % NT, ek = e / NT;
        const int it = ek / NP, kk = ek % NP;
        float4 a = vtmp[((size_t)it * NV + kk) * NT + n];
        float4 b = vtmp[((size_t)it * NV + kk + 1) * NT + n];
        pairs[((size_t)(it * NP + kk) * NT + n) * 2] = a;
        pairs[((size_t)(it * NP + kk) * NT + n) * 2 + 1] =
            make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
    }
    __syncthreads();

    // per-layer scattering coefficients (phase constants folded in)
    const double hr = r - s.Rg;
    const float eR = (float)exp(-hr / s.HR), eM = (float)exp(-hr / s.HM);
    const float A0 = s.bR[0] * eR * s.kR, A1 = s.bR[1] * eR * s.kR, A2 = s.bR[2] * eR * s.kR;
    const float B0 = s.bMS[0] * eM * s.kM, B1 = s.bMS[1] * eM * s.kM, B2 = s.bMS[2] * eM * s.kM;
    const float posA = 0.5f * (NU - 1), posB = 0.5f * (NU - 1) + 1.0f;
    const float g2p1 = s.g2p1, m2g = s.m2g;
    const float SWm1 = (float)(s.SW - 1);
    const double muSmu_base = muS;

    const int ntex = s.RMU * NU;
    for (int tt = threadIdx.x; tt < ntex; tt += blockDim.x) {
        const int y = tt % s.RMU, k = tt / s.RMU;
        const int x = k * s.RMUS + m;
        double mu, muS_t, nu;
        mu_mus_nu(s, L, x, y, mu, muS_t, nu);
        // inscatterS.glsl:113-122 (FP64)
        mu = fmin(fmax(mu, -1.0), 1.0);
        const double mS = muSmu_base;
        const double var = sqrt(1.0 - mu * mu) * sqrt(1.0 - mS * mS);
        nu = fmin(fmax(nu, mS * mu - var), mS * mu + var);
        const double vxd = sqrt(1.0 - mu * mu);
        const double sxd = vxd == 0.0 ? 0.0 : (nu - mS * mu) / vxd;
        const double syd = sqrt(fmax(0.0, 1.0 - sxd * sxd - mS * mS));
        const float vx = (float)vxd, vz = (float)mu;
        const float sx = (float)sxd, sy = (float)syd, sz = (float)mS;

        float r0 = 0.f, r1 = 0.f, r2 = 0.f;
        for (int it = 0; it < NS; ++it) {
            const DirJ dt = dirT[it];
            const float szwz = sz * dt.ct, vzwz = vz * dt.ct;
            const float4 g0 = gdat[2 * it], g1 = gdat[2 * it + 1];
            const bool ground = g0.w != 0.0f;
            const float4 *pt = pairs + (size_t)it * NP * NT * 2;
            const float2 *wrow = dirW + (size_t)it * 2 * NS;
            float aR0 = 0.f, aR1 = 0.f, aR2 = 0.f, aM0 = 0.f, aM1 = 0.f, aM2 = 0.f;
            for (int ip = 0; ip < 2 * NS; ++ip) {
                const float2 w = wrow[ip];
                const float nu1 = fmaf(sx, w.x, fmaf(sy, w.y, szwz));
                const float nu2 = fmaf(vx, w.x, vzwz);
                float pos = fmaf(nu1, posA, posB);
                float fl = floorf(pos);
                float f = pos - fl;
                int kk = min(max((int)fl, 0), NU);
                const float4 *pp = pt + (size_t)kk * NT * 2;
                float4 a = pp[0], d = pp[1];
                float l0 = fmaf(f, d.x, a.x), l1 = fmaf(f, d.y, a.y), l2 = fmaf(f, d.z, a.z);
                if (FIRST) {
                    float4 am = pp[2], dm = pp[3];
                    float m0 = fmaf(f, dm.x, am.x), m1 = fmaf(f, dm.y, am.y), m2 = fmaf(f, dm.z, am.z);
                    const float q1 = fmaf(nu1, nu1, 1.0f);
                    const float y1 = rsqrtf(fmaf(m2g, nu1, g2p1));
                    const float t1 = q1 * (y1 * y1 * y1);
                    l0 = fmaf(l0, q1, m0 * t1);
                    l1 = fmaf(l1, q1, m1 * t1);
                    l2 = fmaf(l2, q1, m2 * t1);
                }
                if (ground) {
                    float xg = fmaf(nu1, g1.x, g1.y);
                    xg = fminf(fmaxf(xg, 0.0f), SWm1);
                    float fg = floorf(xg);
                    float gf = xg - fg;
                    int kg = (int)fg;
                    float4 ea = epair[2 * kg], ed = epair[2 * kg + 1];
                    l0 = fmaf(g0.x, fmaf(gf, ed.x, ea.x), l0);
                    l1 = fmaf(g0.y, fmaf(gf, ed.y, ea.y), l1);
                    l2 = fmaf(g0.z, fmaf(gf, ed.z, ea.z), l2);
                }
                const float q2 = fmaf(nu2, nu2, 1.0f);
                const float y2 = rsqrtf(fmaf(m2g, nu2, g2p1));
                const float t2 = q2 * (y2 * y2 * y2);
                aR0 = fmaf(l0, q2, aR0); aR1 = fmaf(l1, q2, aR1); aR2 = fmaf(l2, q2, aR2);
                aM0 = fmaf(l0, t2, aM0); aM1 = fmaf(l1, t2, aM1); aM2 = fmaf(l2, t2, aM2);
            }
            r0 = fmaf(dt.dw, fmaf(A0, aR0, B0 * aM0), r0);
            r1 = fmaf(dt.dw, fmaf(A1, aR1, B1 * aM1), r1);
            r2 = fmaf(dt.dw, fmaf(A2, aR2, B2 * aM2), r2);
        }
        const size_t o = ((size_t)l * s.RMU + y) * s.W4 + x;
        s.dJ4[o] = make_float4(r0, r1, r2, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Pass dEn — irradianceN.glsl:85-114, fused with copyIrradiance k = 1 and its
// ADD blend (PreprocessAtmo.cpp:335-341).  One workgroup per E texel.
// ---------------------------------------------------------------------------
struct DirE {      // per-direction constants of the hemisphere grid
    float wx, wz, wzdw, pad;
};

template <bool FIRST>
__global__ __launch_bounds__(256) void k_irradianceN(const SetDesc *__restrict__ sets,
                                                     const uint2 *__restrict__ tiles, int NI,
                                                     const DirE *__restrict__ dirs) {
    constexpr int NT = FIRST ? 2 : 1;
    extern __shared__ float4 lds[];
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y;
    const int i = t % s.SW, j = t / s.SW;
    const int NU = s.RNU, NP = NU + 1, NV = NU + 2;
    const int NTH = NI / 2;
    float4 *pairs = lds;                                   // [NTH][NP][NT][2]
    float4 *vtmp = pairs + (size_t)NTH * NP * NT * 2;      // [NTH][NV][NT]
    float *red = (float *)(vtmp + (size_t)NTH * NV * NT);  // [3][blockDim/64]

    const double h = (double)j / (s.SH - 1.0) * (s.Rt - s.Rg);
    const double r = s.Rg + h;
    const double muS = -0.2 + (double)i / (s.SW - 1.0) * (1.0 + 0.2);
    const double rho2 = h * (2.0 * s.Rg + h);
    const double txs = tex4d_umus(s, muS) * s.RMUS - 0.5;
    const float4 *tabs[2] = {s.dSR4, s.dSM4};
    for (int e = threadIdx.x; e < NTH * NV; e += blockDim.x) {
        const int it = e / NV, kk = e % NV;
        const double ct = cos((it + 0.5) * (kPI / NI));
        const double rmu = r * ct;
        double tz, ty;
        tex4d_rmu(s, rho2, rmu, rmu * rmu - rho2, tz, ty);
        float4 v[NT];
        stage_slice<NT>(s, tabs, tz, ty, txs, kk - 1, v);
        if (FIRST) {
            v[0].x *= s.kR; v[0].y *= s.kR; v[0].z *= s.kR;
            v[NT - 1].x *= s.kM; v[NT - 1].y *= s.kM; v[NT - 1].z *= s.kM;
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) vtmp[(size_t)e * NT + n] = v[n];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NTH * NP * NT; e += blockDim.x) {
        const int n = e % NT, ek = e / NT;
        const int it = ek / NP, kk = ek % NP;
        float4 a = vtmp[((size_t)it * NV + kk) * NT + n];
        float4 b = vtmp[((size_t)it * NV + kk + 1) * NT + n];
        pairs[((size_t)(it * NP + kk) * NT + n) * 2] = a;
        pairs[((size_t)(it * NP + kk) * NT + n) * 2 + 1] =
            make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
    }
    __syncthreads();

    const float sx = (float)sqrt(fmax(1.0 - muS * muS, 0.0)), sz = (float)muS;
    const float posA = 0.5f * (NU - 1), posB = 0.5f * (NU - 1) + 1.0f;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    const int ndir = 2 * NI * NTH;
    for (int d = threadIdx.x; d < ndir; d += blockDim.x) {
        const int it = d % NTH;
        const DirE w = dirs[d];
        const float nu = fmaf(sx, w.wx, sz * w.wz);
        float pos = fmaf(nu, posA, posB);
        float fl = floorf(pos);
        float f = pos - fl;
        int kk = min(max((int)fl, 0), NU);
        const float4 *pp = pairs + ((size_t)it * NP + kk) * NT * 2;
        float4 a = pp[0], dd = pp[1];
        float l0 = fmaf(f, dd.x, a.x), l1 = fmaf(f, dd.y, a.y), l2 = fmaf(f, dd.z, a.z);
        if (FIRST) {
            float4 am = pp[2], dm = pp[3];
            float m0 = fmaf(f, dm.x, am.x), m1 = fmaf(f, dm.y, am.y), m2 = fmaf(f, dm.z, am.z);
            const float q1 = fmaf(nu, nu, 1.0f);
            const float y1 = rsqrtf(fmaf(s.m2g, nu, s.g2p1));
            const float t1 = q1 * (y1 * y1 * y1);
            l0 = fmaf(l0, q1, m0 * t1);
            l1 = fmaf(l1, q1, m1 * t1);
            l2 = fmaf(l2, q1, m2 * t1);
        }
        a0 = fmaf(l0, w.wzdw, a0);
        a1 = fmaf(l1, w.wzdw, a1);
        a2 = fmaf(l2, w.wzdw, a2);
    }
    // block reduction
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
    }
    const int nw = blockDim.x >> 6, wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv] = a0;
        red[nw + wv] = a1;
        red[2 * nw + wv] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s0 = 0.f, s1 = 0.f, s2 = 0.f;
        for (int w = 0; w < nw; ++w) {
            s0 += red[w];
            s1 += red[nw + w];
            s2 += red[2 * nw + w];
        }
        s.dE4[t] = make_float4(s0, s1, s2, 1.0f);
        if (s.E4) {
            float4 e = s.E4[t];
            e.x += s0; e.y += s1; e.z += s2;
            s.E4[t] = e;
            if (s.outE) {
                s.outE[3 * t + 0] = e.x;
                s.outE[3 * t + 1] = e.y;
                s.outE[3 * t + 2] = e.z;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pass dSn — inscatterN.glsl:105-131 fused with copyInscatterN.glsl:105-110 +
// ADD blend (S.rgb += dS / phaseR(nu)).  One thread per 4D texel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float3 trilinear(const float4 *__restrict__ tb, const SetDesc &s,
                                            float tx, float ty, float tz) {
    int z0, z1, y0, y1, x0, x1;
    float az, ay, ax;
    gl_axis(tx, s.W4, x0, x1, ax);
    gl_axis(ty, s.RMU, y0, y1, ay);
    gl_axis(tz, s.RR, z0, z1, az);
    const int W = s.W4, HW = s.RMU * s.W4;
    float3 c000 = f3(tb[z0 * HW + y0 * W + x0]), c001 = f3(tb[z0 * HW + y0 * W + x1]);
    float3 c010 = f3(tb[z0 * HW + y1 * W + x0]), c011 = f3(tb[z0 * HW + y1 * W + x1]);
    float3 c100 = f3(tb[z1 * HW + y0 * W + x0]), c101 = f3(tb[z1 * HW + y0 * W + x1]);
    float3 c110 = f3(tb[z1 * HW + y1 * W + x0]), c111 = f3(tb[z1 * HW + y1 * W + x1]);
    return lerp3(lerp3(lerp3(c000, c001, ax), lerp3(c010, c011, ax), ay),
                 lerp3(lerp3(c100, c101, ax), lerp3(c110, c111, ax), ay), az);
}

__global__ __launch_bounds__(256) void k_inscatterN(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles,
                                                    int nsamp) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y * 256 + threadIdx.x;
    if (t >= s.n4) return;
    const int x = t % s.W4;
    const int y = (t / s.W4) % s.RMU;
    const int l = t / (s.W4 * s.RMU);
    const LayerGeom L = s.layers[l];
    double mu, muS, nu;
    mu_mus_nu(s, L, x, y, mu, muS, nu);
    const double r = L.r;
    const double rmu = r * mu;
    const double rho2 = r * r - s.Rg2;
    const double delta = rmu * rmu - rho2;                 // constant along the ray
    const bool gnd_ray = delta > 0.0;
    const double sqd = gnd_ray ? sqrt(delta) : 0.0;
    const double sqsky = sqrt(fmax(delta + s.H * s.H, 0.0));
    const double dx = limit_d(s, r, mu) / nsamp;
    const bool up = mu > 0.0;
    const float3 Tfix = T_lookup(s, (float)(r - s.Rg), (float)(up ? mu : -mu));
    const double muSr = muS * r;
    // nu lies on a slice of the table: the nu lerp of texture4D has weight 0 or 1
    double pos = (nu + 1.0) / 2.0 * (s.RNU - 1.0);
    double kr = rint(pos);
    if (fabs(pos - kr) < 1e-9) pos = kr;
    const double kfl = floor(pos);
    const float fk = (float)(pos - kfl);
    const int k0 = (int)kfl;
    const float fH = s.f_H, fRR1 = (float)(s.RR - 1), fMU = (float)s.RMU;
    const float cG = 0.5f - 0.5f / s.RMU, cS = 0.5f + 0.5f / s.RMU, cM = 0.5f - 1.0f / s.RMU;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int i = 0; i <= nsamp; ++i) {
        const double ti = i * dx;
        const double q = rho2 + ti * (ti + 2.0 * rmu);  // ri^2 - Rg^2
        const double rmui = rmu + ti;                    // ri * mui
        const float ri = sqrtf((float)(q + s.Rg2));
        const float rho = sqrtf(fmaxf((float)q, 0.0f));
        // texture4D(dJ, ri, mui, muSi, nu) coordinates
        float uMu;
        if (rmui < 0.0 && gnd_ray) {
            uMu = fmaf(-rho / (float)(sqd - rmui), cM, cG);
        } else {
            uMu = fmaf((float)(sqsky - rmui) / (rho + fH), cM, cS);
        }
        const float ty = uMu * fMU - 0.5f;
        const float tz = rho / fH * fRR1;
        const float muSi = (float)fma(nu, ti, muSr) / ri;
        const float txs = tex4d_txs_f(s, muSi);
        float3 J = trilinear(s.dJ4, s, (float)k0 * s.RMUS + txs, ty, tz);
        if (fk != 0.0f) {
            float3 J1 = trilinear(s.dJ4, s, (float)(k0 + 1) * s.RMUS + txs, ty, tz);
            J = lerp3(J, J1, fk);
        }
        // transmittance(r, mu, ti) (common.glsl:272-282)
        const float h1 = (float)q / (ri + s.f_Rg);
        const float mu1 = (float)rmui / ri;
        float3 Tseg;
        if (up) {
            float3 b = T_lookup(s, h1, mu1);
            Tseg = make_float3(fminf(Tfix.x / b.x, 1.f), fminf(Tfix.y / b.y, 1.f),
                               fminf(Tfix.z / b.z, 1.f));
        } else {
            float3 a = T_lookup(s, h1, -mu1);
            Tseg = make_float3(fminf(a.x / Tfix.x, 1.f), fminf(a.y / Tfix.y, 1.f),
                               fminf(a.z / Tfix.z, 1.f));
        }
        const float w = (i == 0 || i == nsamp) ? 0.5f : 1.0f;
        a0 = fmaf(w * J.x, Tseg.x, a0);
        a1 = fmaf(w * J.y, Tseg.y, a1);
        a2 = fmaf(w * J.z, Tseg.z, a2);
    }
    const float fdx = (float)dx;
    const float4 o = make_float4(a0 * fdx, a1 * fdx, a2 * fdx, 0.f);
    s.dSR4[t] = o;
    if (s.S4) {
        const float pr = (float)((3.0 / (16.0 * kPI)) * (1.0 + nu * nu));
        float4 S = s.S4[t];
        S.x += o.x / pr;
        S.y += o.y / pr;
        S.z += o.z / pr;
        s.S4[t] = S;
    }
}

// ---------------------------------------------------------------------------
// Numeric validation of finished tables: counts non-finite values, T outside
// [0, 1] and negative E / S entries into flags[set].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_check(const SetDesc *__restrict__ sets,
                                               const uint2 *__restrict__ tiles,
                                               unsigned int *__restrict__ flags) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y * 256 + threadIdx.x;
    bool bad = false;
    if (t < s.n4) {
        float4 v = s.S4[t];
        bad |= !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w));
        bad |= v.x < 0.f || v.y < 0.f || v.z < 0.f || v.w < 0.f;
    }
    if (t < s.TW * s.TH) {
        float4 v = s.T4[t];
        bad |= !(v.x >= 0.f && v.x <= 1.f && v.y >= 0.f && v.y <= 1.f && v.z >= 0.f && v.z <= 1.f);
    }
    if (t < s.SW * s.SH && s.E4) {
        float4 v = s.E4[t];
        bad |= !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z));
        bad |= v.x < 0.f || v.y < 0.f || v.z < 0.f;
    }
    if (bad) atomicAdd(&flags[tl.x], 1u);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
size_t lds_inscatterS(int NS, int RNU, int SW, bool first) {
    const int NT = first ? 2 : 1;
    return sizeof(float4) * ((size_t)NS * (RNU + 1) * NT * 2 + (size_t)NS * (RNU + 2) * NT +
                             (size_t)SW * 2 + (size_t)NS * 2);
}
size_t lds_irradianceN(int NI, int RNU, bool first) {
    const int NT = first ? 2 : 1;
    return sizeof(float4) * ((size_t)(NI / 2) * (RNU + 1) * NT * 2 + (size_t)(NI / 2) * (RNU + 2) * NT) +
           sizeof(float) * 3 * 16;
}

hipError_t launch_transmittance(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                                hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_transmittance, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_irradiance1(const SetDesc *sets, const uint2 *tiles, int ntiles,
                              hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_irradiance1, dim3(ntiles), dim3(256), 0, st, sets, tiles);
    return hipGetLastError();
}
hipError_t launch_inscatter1(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inscatter1, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_inscatterS(const SetDesc *sets, const uint2 *tiles, int ntiles, int NS,
                             const void *dirT, const void *dirW, bool first, size_t lds,
                             hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (first)
        hipLaunchKernelGGL(k_inscatterS<true>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NS,
                           (const DirJ *)dirT, (const float2 *)dirW);
    else
        hipLaunchKernelGGL(k_inscatterS<false>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NS,
                           (const DirJ *)dirT, (const float2 *)dirW);
    return hipGetLastError();
}
hipError_t launch_irradianceN(const SetDesc *sets, const uint2 *tiles, int ntiles, int NI,
                              const void *dirs, bool first, size_t lds, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (first)
        hipLaunchKernelGGL(k_irradianceN<true>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NI,
                           (const DirE *)dirs);
    else
        hipLaunchKernelGGL(k_irradianceN<false>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NI,
                           (const DirE *)dirs);
    return hipGetLastError();
}
hipError_t launch_inscatterN(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inscatterN, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_check(const SetDesc *sets, const uint2 *tiles, int ntiles, unsigned int *flags,
                        hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check, dim3(ntiles), dim3(256), 0, st, sets, tiles, flags);
    return hipGetLastError();
}

}  // namespace atmo
