// atmo_kernels.hip — gfx950 kernels of the atmo precompute, one per reference pass.
//
// Every kernel is launched once per pass for a whole batch of parameter sets:
// blockIdx.x indexes a tile table of (set, tile) pairs built by the host plan,
// so sets of different resolutions share one launch.
#include "atmo_kernels.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

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
    double r, mu;
    if (s.linT) {
        r = s.Rg + rr * (s.Rt - s.Rg);
        mu = -0.15 + m * (1.0 + 0.15);
    } else {
        r = s.Rg + rr * rr * (s.Rt - s.Rg);
        mu = -0.15 + tan(1.5 * m) / tan(1.5) * (1.0 + 0.15);
    }
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
        o = q16(o, s.hs);
        s.T4[t] = o;
        if (s.tau4)
            s.tau4[t] = make_float4((float)((double)s.bR[0] * dR + (double)s.bME[0] * dM),
                                    (float)((double)s.bR[1] * dR + (double)s.bME[1] * dM),
                                    (float)((double)s.bR[2] * dR + (double)s.bME[2] * dM), 0.0f);
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
    if (s.dE4) s.dE4[t] = q16(make_float4(T.x * c, T.y * c, T.z * c, 0.0f), s.hs);
    if (s.E4) s.E4[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (s.outE) {
        s.outE[3 * t + 0] = 0.0f;
        s.outE[3 * t + 1] = 0.0f;
        s.outE[3 * t + 2] = 0.0f;
    }
}

constexpr int kNChunk = 64;   // samples staged per chunk
constexpr int kNRows = 4;     // max mu rows per workgroup

// rows per workgroup: as many whole rows as fit 256 threads, at most rpw_cap (<= 4)
__device__ __forceinline__ int sn_rows_per_wg(int W4, int rpw_cap) {
    return max(1, min(min(rpw_cap, kNRows), 256 / W4));
}

// ---------------------------------------------------------------------------
// Pass dS1 — inscatter1.glsl:104-150 fused with copyInscatter1.glsl:104-109.
// A texel's view ray depends only on its r layer and mu row, so everything of
// an integration sample that does not involve the sun — its distance and
// radius, the view-ray transmittance T(r, mu, t), the scale-height factors and
// the r-row of the T(r_i, mu_s_i) lookup — is shared by the RNU x RMU_S texels
// of the row.  A workgroup covers up to 4 rows (as k_inscatterN), stages those
// terms (FP64 geometry) in LDS, and each thread then needs only the sun
// visibility test, mu_s_i and the mu-axis half of the T lookup per sample.
// ---------------------------------------------------------------------------
struct S1Entry {
    double ti;       // sample distance
    double q;        // r_i^2 - Rg^2
    double iri;      // 1 / r_i
    float T[3];      // T(r, mu, t_i), trapezoid weight folded in
    float eR, eM;    // exp(-h_i/HR), exp(-h_i/HM)
    float b;         // T-table row weight of r_i
    unsigned j0, j1; // T-table row offsets (bytes) of r_i
};

constexpr int kS1Chunk = 64;
constexpr int kXkMax = 2048;   // T-table columns whose centres are staged for the sun lookups

template <bool LINT>   // linear transmittance parameterisation (the per-sample T lookup)
__global__ __launch_bounds__(256) void k_inscatter1(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles,
                                                    int nsamp) {
    __shared__ S1Entry ent[kNRows][kS1Chunk];
    __shared__ float rowdx[kNRows];
    __shared__ double xk[LINT ? 1 : kXkMax];
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    // The sun's T lookups at mu_s_i next to the horizon, where T is steep:
    // an FP32 texel coordinate (atanf, then * TW) is good to ~3e-5 of a texel
    // at TW = 512, which moves a grazing sun's T by ~1e-4.  So the filter
    // fraction is taken from the centre of the column the FP32 coordinate
    // falls in, u - u_k = (atan x - atan x_k) / 1.5 with
    // atan x - atan x_k = atan((x - x_k) / (1 + x x_k)) (less than 1.5 / TW,
    // FP32 is exact enough there), x = (mu_s_i + 0.15) tan(1.5) / 1.15 in FP64
    // and x_k = tan(1.5 (k + 0.5) / TW) staged once per workgroup (tables of at
    // most kXkMax columns; wider ones keep the FP32 coordinate).
    const bool tab = !LINT && s.TW <= kXkMax;
    const float sc = (float)s.TW / 1.5f;
    if (tab)
        for (int k = threadIdx.x; k < s.TW; k += blockDim.x) xk[k] = tan(1.5 * (k + 0.5) / s.TW);
    const int RPW = sn_rows_per_wg(s.W4, kNRows);
    const int ngy = (s.RMU + RPW - 1) / RPW;
    const int l = (int)tl.y / ngy;
    const int yb = ((int)tl.y % ngy) * RPW;
    const int nrows = min(RPW, s.RMU - yb);
    const LayerGeom L = s.layers[l];
    const double r = L.r;
    const double rho2 = r * r - s.Rg2;
    const int per = nrows * s.W4;
    const int nmine = (per + (int)blockDim.x - 1) / (int)blockDim.x;
    constexpr int MAXT = 4;   // W4 <= 1024 (validated on the host)
    float aR[MAXT][3], aM[MAXT][3];
    double nud[MAXT], muSr[MAXT];
    int rowi[MAXT];
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        aR[j][0] = aR[j][1] = aR[j][2] = aM[j][0] = aM[j][1] = aM[j][2] = 0.f;
        nud[j] = muSr[j] = 0.0;
        rowi[j] = 0;
        const int e = threadIdx.x + j * (int)blockDim.x;
        if (j < nmine && e < per) {
            const int row = e / s.W4, x = e % s.W4;
            double mu, muS, nu;
            mu_mus_nu(s, L, x, yb + row, mu, muS, nu);
            nud[j] = nu;
            muSr[j] = muS * r;
            rowi[j] = row;
        }
    }
    const float iHR = (float)(1.0 / s.HR), iHM = (float)(1.0 / s.HM);
    const __amdgpu_buffer_rsrc_t Tr = table_rsrc(s.T4, (unsigned)((size_t)s.TW * s.TH * sizeof(float4)));
    for (int c0 = 0; c0 <= nsamp; c0 += kS1Chunk) {
        const int nc = min(kS1Chunk, nsamp + 1 - c0);
        __syncthreads();
        for (int e = threadIdx.x; e < nrows * kS1Chunk; e += blockDim.x) {
            const int row = e / kS1Chunk, js = e % kS1Chunk;
            if (js >= nc) continue;
            const int i = c0 + js;
            double mu, muS_u, nu_u;
            mu_mus_nu(s, L, 0, yb + row, mu, muS_u, nu_u);
            const double rmu = r * mu;
            const double dx = limit_d(s, r, mu) / nsamp;
            if (i == 0) rowdx[row] = (float)dx;
            const bool up = mu > 0.0;
            S1Entry en;
            en.ti = i * dx;
            en.q = rho2 + en.ti * (en.ti + 2.0 * rmu);                 // ri^2 - Rg^2
            const float ri = sqrtf((float)(en.q + s.Rg2));
            const float h1 = (float)en.q / (ri + s.f_Rg);              // ri - Rg (< 0 at a ground hit)
            const float mu1 = (float)(rmu + en.ti) / ri;
            // transmittance(r, mu, t_i) (common.glsl:272-282)
            const float3 Tseg = s.tview ? ld3(gptr(s.tview) + ((size_t)l * s.RMU + yb + row) * (nsamp + 1) + i)
                              : up      ? T_ratio_lin(s, (float)(r - s.Rg), (float)mu, h1, mu1)
                                        : T_ratio_lin(s, h1, -mu1, (float)(r - s.Rg), (float)-mu);
            const float w = (i == 0 || i == nsamp) ? 0.5f : 1.0f;
            en.T[0] = Tseg.x * w; en.T[1] = Tseg.y * w; en.T[2] = Tseg.z * w;
            const double rid = sqrt(en.q + s.Rg2);
            en.iri = 1.0 / rid;
            const float hc = fmaxf(h1, 0.0f);                          // ri = max(Rg, ri)
            en.eR = __expf(-hc * iHR);
            en.eM = __expf(-hc * iHM);
            // r axis of T(max(Rg, ri), muSi): common.glsl:114-124 (FP64 coordinate)
            int j0, j1;
            double bd;
            gl_axis_dd(T_ur_d(s, en.q / (rid + s.Rg)) * s.TH - 0.5, s.TH, j0, j1, bd);
            en.b = (float)bd;
            en.j0 = (unsigned)(j0 * s.TW) * 16u;
            en.j1 = (unsigned)(j1 * s.TW) * 16u;
            ent[row][js] = en;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < MAXT; ++j) {
            if (j >= nmine) break;
            const int e = threadIdx.x + j * (int)blockDim.x;
            if (e >= per) break;
            const S1Entry *er = ent[rowi[j]];
            float r0 = aR[j][0], r1 = aR[j][1], r2 = aR[j][2];
            float m0 = aM[j][0], m1 = aM[j][1], m2 = aM[j][2];
            for (int js = 0; js < nc; ++js) {
                const S1Entry &en = er[js];
                const double num = fma(nud[j], en.ti, muSr[j]);        // ri * muSi
                // muSi >= -sqrt(1 - Rg^2 / max(Rg, ri)^2)  <=>  num >= 0 or num^2 <= ri^2 - Rg^2
                if (!(num >= 0.0 || num * num <= en.q)) continue;
                const double mus = num * en.iri;   // mu_s_i
                int i0, i1;
                float a;
                if (tab) {
                    gl_axis(T_umu_t<false>(s, (float)mus) * s.TW - 0.5f, s.TW, i0, i1, a);
                    if (i1 == i0 + 1) {   // between two columns (not clamped to an edge)
                        const double x = (mus + 0.15) * s.tMu, x0 = xk[i0];
                        const float d = (float)(x - x0) * __builtin_amdgcn_rcpf(fmaf((float)x, (float)x0, 1.0f));
                        // the FP64 coordinate can sit a rounding outside the FP32
                        // cell: keep the filter weight in [0, 1] (a convex lerp)
                        a = __builtin_amdgcn_fmed3f(atanf(d) * sc, 0.0f, 1.0f);
                    }
                } else if (LINT) {   // (mu_s_i + 0.15) / 1.15: FP64 is cheap here
                    double ad;
                    gl_axis_dd(T_umu_d<true>(s, mus) * s.TW - 0.5, s.TW, i0, i1, ad);
                    a = (float)ad;
                } else {   // more than kXkMax columns: the FP32 coordinate
                    gl_axis(T_umu_t<false>(s, (float)mus) * s.TW - 0.5f, s.TW, i0, i1, a);
                }
                // buffer loads at byte offsets (i0, i1 >= 0: clamped to the table)
                const unsigned o0 = (unsigned)i0 << 4, o1 = (unsigned)i1 << 4;
                const float3 t00 = ld3b(Tr, en.j0 + o0, 0u), t01 = ld3b(Tr, en.j0 + o1, 0u);
                const float3 t10 = ld3b(Tr, en.j1 + o0, 0u), t11 = ld3b(Tr, en.j1 + o1, 0u);
                const float3 Ts = lerp3(lerp3(t00, t01, a), lerp3(t10, t11, a), en.b);
                const float tx = en.T[0] * Ts.x, ty = en.T[1] * Ts.y, tz = en.T[2] * Ts.z;
                r0 = fmaf(tx, en.eR, r0); r1 = fmaf(ty, en.eR, r1); r2 = fmaf(tz, en.eR, r2);
                m0 = fmaf(tx, en.eM, m0); m1 = fmaf(ty, en.eM, m1); m2 = fmaf(tz, en.eM, m2);
            }
            aR[j][0] = r0; aR[j][1] = r1; aR[j][2] = r2;
            aM[j][0] = m0; aM[j][1] = m1; aM[j][2] = m2;
        }
    }
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        if (j >= nmine) break;
        const int e = threadIdx.x + j * (int)blockDim.x;
        if (e >= per) break;
        const int row = e / s.W4, x = e % s.W4;
        const size_t t = ((size_t)l * s.RMU + yb + row) * s.W4 + x;
        const float fdx = rowdx[row];
        const float4 oR = q16(make_float4(aR[j][0] * fdx * s.bR[0], aR[j][1] * fdx * s.bR[1], aR[j][2] * fdx * s.bR[2], 0.f), s.hs);
        const float4 oM = q16(make_float4(aM[j][0] * fdx * s.bMS[0], aM[j][1] * fdx * s.bMS[1], aM[j][2] * fdx * s.bMS[2], 0.f), s.hs);
        s.dSR4[t] = oR;
        s.dSM4[t] = oM;
        if (s.S4) s.S4[t] = make_float4(oR.x, oR.y, oR.z, oM.x);
    }
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
// Staged pairs per theta row, kk = 0..RNU, with pos = (nu+1)/2 (RNU-1) + 1,
// kk = floor(pos), f = pos - kk, A = V[kk-1], D = V[kk] - V[kk-1]:
// irradianceN stores {A, D} and looks up A + f D; inscatterS stores
// {A - kk D, D} (rounded once from FP64) and looks up fma(pos, D, A - kk D),
// the same line without the fraction.
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
        const gfloat4 *tb = gptr(tabs[n]);
        float3 c000 = ld3(tb + z0 * HW + y0 * W + x0), c001 = ld3(tb + z0 * HW + y0 * W + x1);
        float3 c010 = ld3(tb + z0 * HW + y1 * W + x0), c011 = ld3(tb + z0 * HW + y1 * W + x1);
        float3 c100 = ld3(tb + z1 * HW + y0 * W + x0), c101 = ld3(tb + z1 * HW + y0 * W + x1);
        float3 c110 = ld3(tb + z1 * HW + y1 * W + x0), c111 = ld3(tb + z1 * HW + y1 * W + x1);
        float3 v = lerp3(lerp3(lerp3(c000, c001, ax), lerp3(c010, c011, ax), ay),
                         lerp3(lerp3(c100, c101, ax), lerp3(c110, c111, ax), ay), az);
        out[n] = make_float4(v.x, v.y, v.z, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Pass dJn — inscatterS.glsl:111-189.  One workgroup per (set, r layer,
// mu_s column); its RMU x RNU texels share r and mu_s, hence the staged
// slice values.  FIRST = order 2 (Rayleigh and Mie stored separately).
//
// Each thread integrates G texels of one mu row (G nu slices): they share
// the view vector v, so the phase factors of nu2 = v.w are evaluated once per
// direction for all G, and since v has no y component the directions phi and
// 2 pi - phi share them too; the thread walks the grid in such mirrored pairs
// and folds both phases and the scattering coefficients into one rgb weight
// per pair, as the reference's sum does per direction.  Per direction and
// texel the lookup position is pos = (nu1 + 1)/2 (RNU - 1) + 1 =
// posA s'.w + posB, with s' = posA s folded in per texel and the mu_s term
// folded in per theta.
// ---------------------------------------------------------------------------
// Entries of padding in front of the staged ground-irradiance row: texel
// coordinates down to -(0.8/1.2) SW - 0.5 (mu_s = -1) plus rounding.
__host__ __device__ __forceinline__ int e_pad(int SW) { return (2 * SW) / 3 + 3; }
// The lookup positions pos = (nu1 + 1)/2 (RNU - 1) + 1 of a direction lie in
// [1, RNU] up to FP32 rounding (|nu1| <= 1 + a few ulp); the merged ground rows
// cover [kPosLo, RNU + kPosHi].
constexpr float kPosLo = 1.0f - 1.0f / 256.0f, kPosHi = 1.0f / 256.0f;

// The value at 0 of the line through (k, a) with slope d, a - k d, from FP64
// and rounded once; a lookup at pos in [k, k + 1) is then fma(pos, d, .).
__device__ __forceinline__ float at_zero(float a, float d, int k) {
    return (float)((double)a - (double)k * (double)d);
}

struct DirJ {      // per-theta constants of the NS x 2NS grid
    float ct, st, dw, pad;
};

// CHUNK: the theta rows are staged CH at a time (when the staging of all NS
// rows exceeds the LDS budget); between chunks a thread keeps its running sums
// in the dJ table itself and resumes them, so the sum over theta runs in the
// same order, and gives the same bits, as with all rows staged at once
// (!CHUNK, CH = NS: one round, compiled without the chunk bookkeeping).
// At most 128 VGPRs (four waves per SIMD, as the LDS of a typical staging
// allows): G = 8 then keeps its direction loops free of spills (its per-group
// set-up spills a few registers) and runs 2 % faster than G = 4 at 32 x 64
// directions; at 166 VGPRs (three waves) it ran 14 % slower.
template <bool FIRST, int G, bool CHUNK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_inscatterS(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles, int NS, int CH,
                                                    int JX, const DirJ *__restrict__ dirT,
                                                    const float2 *__restrict__ dirW) {
    constexpr int NT = FIRST ? 2 : 1;
    constexpr int PW = FIRST ? 3 : 2;      // float4 per staged pair entry
    extern __shared__ float4 lds[];
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int l = (int)tl.y / s.RMUS;
    const int m = (int)tl.y % s.RMUS;
    const int NU = s.RNU;
    const int NP = NU + 1;                 // pairs per theta
    const int NV = NU + 2;                 // slice values per theta
    const int EPAD = e_pad(s.SW);
    // orders >= 3: a ground row holds NP + (its E cells - 1) merged entries
    // (sky + ground terms); JX bounds the extra entries of a chunk's rows
    const int JXr = FIRST ? 0 : JX;
    const size_t npairs = (size_t)CH * NP + (size_t)JXr;
    __shared__ int s_jxmax;
    // order 2, unchunked: the entries' three float4 as planes kJ1Plane apart
    // (16-byte stride in each: a lookup's reads spread over the LDS banks, and
    // one address serves all three with immediate offsets); otherwise whole
    constexpr bool PLANES = FIRST && !CHUNK;
    constexpr int ES = PLANES ? 1 : PW;                    // entry stride (float4)
    constexpr int CS = PLANES ? kJ1Plane : 1;              // component stride (float4)
    if (PLANES && npairs > (size_t)kJ1Plane) return;       // the host never launches this
    float4 *pairs = lds;                                   // rows of NP (+ JX) entries
    // [CH][NV][NT] slice values (the planes compute each entry's two directly)
    float4 *vtmp = pairs + (PLANES ? 2 * kJ1Plane + npairs : npairs * PW);
    float4 *epair = vtmp + (PLANES ? 0 : (size_t)CH * NV * NT);   // [EPAD + SW + 1][2] (order 2)
    float4 *gdat = epair + (FIRST ? (size_t)(EPAD + s.SW + 1) * 2 : 0);   // [CH][3]

    const LayerGeom L = s.layers[l];
    const double r = fmin(fmax(L.r, s.Rg), s.Rt);          // inscatterS.glsl:112
    const double muS = fmin(fmax(texel_mus(s, m), -1.0), 1.0);   // :114
    const double rho2 = r * r - s.Rg2;
    const double txs = tex4d_umus(s, muS) * s.RMUS - 0.5;
    const float4 *tabs[2] = {s.dSR4, s.dSM4};
    const double posAd = 0.5 * (NU - 1), posBd = posAd + 1.0;

    // (2) order 2: row 0 of dE (irradiance at r = Rg reads only row 0 under
    // CLAMP_TO_EDGE), with EPAD copies of texel 0 in front and one of texel
    // SW-1 behind (zero differences): every texel coordinate the ground term can
    // produce, (mu_s + 0.2)/1.2 SW - 0.5 for |mu_s| <= 1, then lands on an entry
    // that reproduces the clamp, and the lookup needs no clamp instruction
    for (int k = threadIdx.x; FIRST && k < EPAD + s.SW + 1; k += blockDim.x) {
        const int k0 = min(max(k - EPAD, 0), s.SW - 1), k1 = min(max(k - EPAD + 1, 0), s.SW - 1);
        const float4 a = s.dE4[k0];
        const float4 b = s.dE4[k1];
        // {a - k d, d}: the lookup at texel coordinate xg is fma(xg, d, a - k d)
        const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
        epair[2 * k] = make_float4(at_zero(a.x, dx, k), at_zero(a.y, dy, k), at_zero(a.z, dz, k), dx);
        epair[2 * k + 1] = make_float4(dy, dz, 0.f, 0.f);
    }

    // per-layer scattering coefficients (phase constants folded in)
    const double hr = r - s.Rg;
    const float eR = (float)exp(-hr / s.HR), eM = (float)exp(-hr / s.HM);
    const float A0 = s.bR[0] * eR * s.kR, A1 = s.bR[1] * eR * s.kR, A2 = s.bR[2] * eR * s.kR;
    const float B0 = s.bMS[0] * eM * s.kM, B1 = s.bMS[1] * eM * s.kM, B2 = s.bMS[2] * eM * s.kM;
    const float iposA = (float)(1.0 / posAd), nposBi = (float)(-posBd / posAd);
    const float g2p1 = s.g2p1, m2g = s.m2g;
    const int ngk = (NU + G - 1) / G;
    const int ngroups = s.RMU * ngk;

    for (int c0 = 0; c0 < NS; c0 += CH) {
        const int nth = CHUNK ? min(CH, NS - c0) : NS;
        const bool last = !CHUNK || c0 + nth == NS;
        if (CHUNK && c0 > 0) __syncthreads();   // every thread is done with the previous chunk's rows
        // (1) slice values V[it][kk], kk = slice + 1
        auto slice_value = [&](int it, int kk, float4 v[NT]) {
            const double ct = cos((it + 0.5) * (kPI / NS));
            const double rmu = r * ct;
            double tz, ty;
            tex4d_rmu(s, rho2, rmu, rmu * rmu - rho2, ct, tz, ty);
            stage_slice<NT>(s, tabs, tz, ty, txs, kk - 1, v);
            if (FIRST) {  // fold the phase-function constants in
                v[0].x *= s.kR; v[0].y *= s.kR; v[0].z *= s.kR;
                v[NT - 1].x *= s.kM; v[NT - 1].y *= s.kM; v[NT - 1].z *= s.kM;
            }
        };
        for (int e = threadIdx.x; !PLANES && e < nth * NV; e += blockDim.x) {
            float4 v[NT];
            slice_value(c0 + e / NV, e % NV, v);
    #pragma unroll
            for (int n = 0; n < NT; ++n) vtmp[(size_t)e * NT + n] = v[n];
        }
        // (3) per theta: ground visibility, ground-to-x transmittance (:129-138), and
        // the theta-uniform parts of the lookup positions
        for (int itl = threadIdx.x; itl < nth; itl += blockDim.x) {
            const int it = c0 + itl;
            const double ct = cos((it + 0.5) * (kPI / NS));
            const double q = s.Rg / r;
            const double cthetamin = -sqrt(1.0 - q * q);
            float4 g0 = make_float4(0.f, 0.f, 0.f, 0.f);
            double c1 = 0.0, ce0 = 0.0;
            int jlo = 0, jhi = 0;
            if (ct < cthetamin) {
                const double dg = -r * ct - sqrt(r * r * (ct * ct - 1.0) + s.Rg2);
                const double mug = -(r * ct + dg) / s.Rg;
                // transmittance(Rg, mug, dg) (common.glsl:272-282); its end point is x itself
                const double r1sq = s.Rg2 + dg * dg + 2.0 * s.Rg * mug * dg;
                const double r1 = sqrt(r1sq);
                const double mu1 = (s.Rg * mug + dg) / r1;
                const float h1 = (float)((r1sq - s.Rg2) / (r1 + s.Rg));
                const float3 tr = s.tground ? ld3(gptr(s.tground) + (size_t)l * NS + it)
                                  : mug > 0.0 ? T_ratio_lin(s, 0.0f, (float)mug, h1, (float)mu1)
                                              : T_ratio_lin(s, h1, (float)-mu1, 0.0f, (float)-mug);
                const float gr = (float)(s.alpha / kPI);
                g0 = make_float4(gr * tr.x, gr * tr.y, gr * tr.z, 1.0f);
                // E texel coordinate ((r muS + dg nu1)/Rg + 0.2)/1.2 SW - 0.5 as c1 pos + ce0
                const double sc = s.SW / 1.2;
                const double e1 = dg / s.Rg * sc, e0 = (r * muS / s.Rg + 0.2) * sc - 0.5 + EPAD;
                c1 = e1 / posAd;
                ce0 = e0 - e1 * posBd / posAd;
                // E cells the lookups of this row can reach: positions [1, NU] up to rounding
                const float fc1 = (float)c1, fce0 = (float)ce0;
                jlo = (int)__float2uint_rz(fmaf(kPosLo, fc1, fce0));
                jhi = (int)__float2uint_rz(fmaf((float)NU + kPosHi, fc1, fce0));
                // the lookups move xg by at most c1 (RNU - 1 + 2/256) <= 2.0158 e1
                // (RNU >= 2): at most that many + 1 cells; the host sizes the rows
                // with ceil(2.05 e1) + 2 per ground row (merged_extra)
                jhi = min(jhi, jlo + (int)ceil(2.05 * e1) + 1);
            }
            gdat[3 * itl] = g0;
            gdat[3 * itl + 1] = make_float4((float)c1, (float)ce0, (float)(posAd * muS * ct + posBd),
                                            (float)((kPI / NS) * (kPI / NS) * sin((it + 0.5) * (kPI / NS))));
            gdat[3 * itl + 2] = make_float4(__int_as_float(jlo), __int_as_float(jhi),
                                            __int_as_float(FIRST ? NP : NP + (jhi - jlo)), 0.f);
        }
        __syncthreads();
        // first entry of every row: the sum of the entries of the rows before it
        // (a ground row needs NP + jhi - jlo), and the widest ground row
        if (threadIdx.x == 0) s_jxmax = 0;
        for (int itl = threadIdx.x; itl < nth; itl += blockDim.x) {
            int off = 0;
            for (int i = 0; i < itl; ++i) off += __float_as_int(gdat[3 * i + 2].z);
            gdat[3 * itl + 2].w = __int_as_float(off);
        }
        __syncthreads();
        if (!FIRST)
            for (int itl = threadIdx.x; itl < nth; itl += blockDim.x)
                atomicMax(&s_jxmax, __float_as_int(gdat[3 * itl + 2].z) - NP);
        // ground rows are a suffix of the chunk (ct decreases with theta): the
        // first one, found by bisection on the flags
        int ig0 = nth;
        if (!FIRST) {
            int lo = 0, hi = nth;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (gdat[3 * mid].w != 0.0f) hi = mid; else lo = mid + 1;
            }
            ig0 = lo;
        }
        auto rowoff = [&](int itl) { return __float_as_int(gdat[3 * itl + 2].w); };
        __syncthreads();   // s_jxmax
        // (4) pairs {V[kk-1], V[kk] - V[kk-1]}
        for (int e = threadIdx.x; e < nth * NP; e += blockDim.x) {
            const int it = e / NP, kk = e % NP;
            if (!FIRST && it >= ig0) continue;   // merged ground row (4b)
            float4 va[NT], vb[NT];   // V[kk-1], V[kk]
            if (PLANES) {            // computed here: no slice-value staging (LDS)
                slice_value(c0 + it, kk, va);
                slice_value(c0 + it, kk + 1, vb);
            } else {
    #pragma unroll
                for (int n = 0; n < NT; ++n) {
                    va[n] = vtmp[((size_t)it * NV + kk) * NT + n];
                    vb[n] = vtmp[((size_t)it * NV + kk + 1) * NT + n];
                }
            }
            const float4 a = va[0], b = vb[0];
            float4 *q = pairs + ((size_t)rowoff(it) + kk) * ES;
            const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            const float px = at_zero(a.x, dx, kk), py = at_zero(a.y, dy, kk), pz = at_zero(a.z, dz, kk);
            if (!FIRST) {
                // (P.rgb, D.r), (D.g, D.b): one 16-byte and one 8-byte LDS read
                q[0] = make_float4(px, py, pz, dx);
                q[1] = make_float4(dy, dz, 0.f, 0.f);
            } else {
                // Rayleigh and Mie pairs in 48 bytes: three 16-byte LDS reads
                const float4 am = va[NT - 1], bm = vb[NT - 1];
                const float ex = bm.x - am.x, ey = bm.y - am.y, ez = bm.z - am.z;
                q[0] = make_float4(px, py, pz, dx);
                q[CS] = make_float4(dy, dz, at_zero(am.x, ex, kk), at_zero(am.y, ey, kk));
                q[2 * CS] = make_float4(at_zero(am.z, ez, kk), ex, ey, ez);
            }
        }
        // (4b) orders >= 3, ground rows: the lookup of a ground direction is
        // S(pos) + g0 E(xg(pos)) with xg = c1 pos + ce0, two piecewise-linear
        // functions of the same position, with knots at the integers (slices) and
        // where xg crosses an integer (E texels).  The merged row holds one entry
        // per segment between consecutive knots of either kind: the slope Q and
        // P - p0 Q, where P is the value at the segment's start p0, so that a
        // lookup is fma(pos, Q, P - p0 Q).  The segment of (kk, jg) = (floor(pos), floor(xg)) is
        // entry kk + jg - jlo (each knot raises one of the two by one); segments
        // are found by intersecting every slice cell with every E cell of the row
        // (FP64), so entries of zero-length segments at coinciding knots exist too.
        if (!FIRST && ig0 < nth) {
            const int nj = s_jxmax + 1, per_row = NP * nj;
            for (int e = threadIdx.x; e < (nth - ig0) * per_row; e += blockDim.x) {
                const int itl = ig0 + e / per_row, kk = (e % per_row) / nj, jj = e % nj;
                // the whole merged row must lie inside the staging (guaranteed by
                // the host's bound; a row that would not fit is left unwritten)
                if (rowoff(itl) + __float_as_int(gdat[3 * itl + 2].z) > (int)npairs) continue;
                const float4 g0 = gdat[3 * itl], g1 = gdat[3 * itl + 1], g2 = gdat[3 * itl + 2];
                const int jlo = __float_as_int(g2.x), jhi = __float_as_int(g2.y), jg = jlo + jj;
                if (jg > jhi) continue;
                const double c1 = g1.x, ce0 = g1.y;
                double lo = kk == 0 ? (double)kPosLo : (double)kk;
                double hi = kk == NU ? (double)NU + kPosHi : kk + 1.0;
                if (c1 > 0.0) {
                    lo = fmax(lo, (jg - ce0) / c1);
                    hi = fmin(hi, (jg + 1 - ce0) / c1);
                }
                if (lo > hi) continue;
                const float4 a = vtmp[((size_t)itl * NV + kk) * NT];
                const float4 b = vtmp[((size_t)itl * NV + kk + 1) * NT];
                const float4 ea = s.dE4[min(max(jg - EPAD, 0), s.SW - 1)];
                const float4 eb = s.dE4[min(max(jg + 1 - EPAD, 0), s.SW - 1)];
                const double u = lo - kk, v = lo * c1 + ce0 - jg;   // fractions at p0
                const double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
                const double fx = (double)eb.x - ea.x, fy = (double)eb.y - ea.y, fz = (double)eb.z - ea.z;
                const double Px = a.x + u * dx + g0.x * (ea.x + v * fx);
                const double Py = a.y + u * dy + g0.y * (ea.y + v * fy);
                const double Pz = a.z + u * dz + g0.z * (ea.z + v * fz);
                const float Qx = (float)(dx + g0.x * c1 * fx), Qy = (float)(dy + g0.y * c1 * fy),
                            Qz = (float)(dz + g0.z * c1 * fz);
                // the FP32 slope's line through (p0, P), at pos = 0
                float4 *q = pairs + ((size_t)rowoff(itl) + kk + jj) * PW;
                q[0] = make_float4((float)(Px - lo * Qx), (float)(Py - lo * Qy), (float)(Pz - lo * Qz), Qx);
                q[1] = make_float4(Qy, Qz, 0.f, 0.f);
            }
        }
        __syncthreads();

        for (int grp = threadIdx.x; grp < ngroups; grp += blockDim.x) {
            const int y = grp % s.RMU, kb = (grp / s.RMU) * G;
            float vx = 0.f, vz = 0.f;
            float sxp[G], syp[G];
    #pragma unroll
            for (int g = 0; g < G; ++g) {
                const int k = min(kb + g, NU - 1);
                double mu, muS_t, nu;
                mu_mus_nu(s, L, k * s.RMUS + m, y, mu, muS_t, nu);
                // inscatterS.glsl:113-122 (FP64)
                mu = fmin(fmax(mu, -1.0), 1.0);
                const double var = sqrt(1.0 - mu * mu) * sqrt(1.0 - muS * muS);
                nu = fmin(fmax(nu, muS * mu - var), muS * mu + var);
                const double vxd = sqrt(1.0 - mu * mu);
                const double sxd = vxd == 0.0 ? 0.0 : (nu - muS * mu) / vxd;
                const double syd = sqrt(fmax(0.0, 1.0 - sxd * sxd - muS * muS));
                vx = (float)vxd;
                vz = (float)mu;
                sxp[g] = (float)(posAd * sxd);
                syp[g] = (float)(posAd * syd);
            }
            const float cx1 = m2g * vx;
            float acc[G][3];
    #pragma unroll
            for (int g = 0; g < G; ++g)
    #pragma unroll
                for (int c = 0; c < 3; ++c) acc[g][c] = 0.f;
            if (CHUNK && c0 > 0) {   // resume the sums of the previous chunks (written by this thread)
    #pragma unroll
                for (int g = 0; g < G; ++g) {
                    const int k = kb + g;
                    if (k < NU) {
                        const size_t o = ((size_t)l * s.RMU + y) * s.W4 + (size_t)k * s.RMUS + m;
                        if (s.dJb == 12) {
                            const float *q = (const float *)s.dJ4 + 3 * o;
                            acc[g][0] = q[0]; acc[g][1] = q[1]; acc[g][2] = q[2];
                        } else {
                            const float4 q = s.dJ4[o];
                            acc[g][0] = q.x; acc[g][1] = q.y; acc[g][2] = q.z;
                        }
                    }
                }
            }

            // 4D lookup of one direction from the staged pairs: the clamp of the
            // nu slice is implicit, |s.w| <= 1 keeps pos in [1, RNU] (+ rounding)
            // and kk = trunc(pos) in [0, RNU] = the staged pair range.
            // (order 2: the Rayleigh and Mie values weighted by the phase functions of
            // nu1 = s.w, as q1 (R + z1 M) with q1 = 1 + nu1^2, z1 = (1+g^2-2g nu1)^-3/2;
            // the caller applies q1, whose per-set constants are folded into the staging)
            auto lookup = [&](const float4 *pt, float pos, float out[3], float &q1) {
                const unsigned kk = __float2uint_rz(pos);
                // ES = 3: a 24-bit multiply (full rate; kk <= RNU) instead of the
                // 32-bit one the compiler picks; 2 and 1: a shift
                const float4 *pp = pt + (ES == 3 ? __umul24(kk, (unsigned)ES) : kk * ES);
                const float4 a = pp[0];
                if (!FIRST) {
                    const float2 d = *(const float2 *)(pp + 1);
                    out[0] = fmaf(pos, a.w, a.x);
                    out[1] = fmaf(pos, d.x, a.y);
                    out[2] = fmaf(pos, d.y, a.z);
                    q1 = 1.0f;
                } else {
                    const float4 b = pp[CS], c2 = pp[2 * CS];
                    const float nu1 = fmaf(pos, iposA, nposBi);
                    q1 = fmaf(nu1, nu1, 1.0f);
                    const float y1 = __builtin_amdgcn_rsqf(fmaf(m2g, nu1, g2p1));
                    const float z1 = y1 * y1 * y1;
                    const float m0 = fmaf(pos, c2.y, b.z), m1 = fmaf(pos, c2.z, b.w), m2 = fmaf(pos, c2.w, c2.x);
                    out[0] = fmaf(m0, z1, fmaf(pos, a.w, a.x));
                    out[1] = fmaf(m1, z1, fmaf(pos, b.x, a.y));
                    out[2] = fmaf(m2, z1, fmaf(pos, b.y, a.z));
                }
            };
            // orders >= 3, ground direction: the merged sky + ground entry of the
            // segment holding pos (4b)
            auto glookup = [&](unsigned gb, float pos, float cg1, float cg0, float out[3]) {
                const float xg = fmaf(pos, cg1, cg0);   // >= 0: EPAD is folded into cg0
                const unsigned e = __float2uint_rz(pos) + __float2uint_rz(xg);
                const float4 *pp = pairs + (gb + e * PW);
                const float4 a = pp[0];
                const float2 b = *(const float2 *)(pp + 1);
                out[0] = fmaf(pos, a.w, a.x);
                out[1] = fmaf(pos, b.x, a.y);
                out[2] = fmaf(pos, b.y, a.z);
            };
            // irradiance at the ground point seen in direction w, row 0 of dE
            auto ground_e = [&](float pos, float cg1, float cg0, float out[3]) {
                const float xg = fmaf(pos, cg1, cg0);   // >= 0: EPAD is folded into cg0
                const unsigned kg = __float2uint_rz(xg);
                const float4 ea = epair[2 * kg];
                const float2 ed = *(const float2 *)(epair + 2 * kg + 1);
                out[0] = fmaf(xg, ea.w, ea.x);
                out[1] = fmaf(xg, ed.x, ea.y);
                out[2] = fmaf(xg, ed.y, ea.z);
            };

            for (int itl = 0; itl < nth; ++itl) {
                const int it = c0 + itl;
                const float ct = dirT[it].ct;
                const float4 g0 = gdat[3 * itl], g1 = gdat[3 * itl + 1];
                const bool ground = g0.w != 0.0f;
                const float vzwz = vz * ct;
                const float cx0 = fmaf(m2g, vzwz, g2p1);
                const float dw = g1.w, cp = g1.z;
                const float4 *pt = pairs + (size_t)rowoff(itl) * ES;
                // merged ground row: entry (floor(pos) + floor(xg)) - jlo, as an
                // index that wraps modulo 2^32 (jlo may exceed the row offset)
                const unsigned gbase = (unsigned)(rowoff(itl) - __float_as_int(gdat[3 * itl + 2].x)) * PW;
                const float2 *wrow = dirW + (size_t)it * 2 * NS;
                // the directions of one theta row; GR: a merged ground row of orders
                // >= 3 (a compile-time branch, so the lookups of all G texels stay
                // one straight sequence)
                auto phi_row = [&](auto gr) {
                    constexpr bool GR = decltype(gr)::value;
                    float2 wn = wrow[0];
                    // directions phi and 2 pi - phi together: same w.x, opposite w.y,
                    // hence the same nu2 = v.w (v has no y component) and weights
                    for (int ip = 0; ip < NS; ++ip) {
                        const float2 w = wn;
                        if (ip + 1 < NS) wn = wrow[ip + 1];  // scalar prefetch of the next direction
                        const float nu2 = fmaf(vx, w.x, vzwz);
                        const float q2 = fmaf(nu2 * dw, nu2, dw);
                        const float y2 = __builtin_amdgcn_rsqf(fmaf(cx1, w.x, cx0));
                        const float t2 = q2 * (y2 * y2 * y2);
                        // betaR e^(-h/HR) pR(nu2) dw + betaMSca e^(-h/HM) pM(nu2) dw (:182-185)
                        const float W0 = fmaf(A0, q2, B0 * t2), W1 = fmaf(A1, q2, B1 * t2),
                                    W2 = fmaf(A2, q2, B2 * t2);
                        float pa[G], pb[G], lv[G][3];
    #pragma unroll
                        for (int g = 0; g < G; ++g) {
                            const float base = fmaf(sxp[g], w.x, cp);
                            pa[g] = fmaf(syp[g], w.y, base);
                            pb[g] = fmaf(-syp[g], w.y, base);
                            float la[3], lb[3], qa, qb;
                            if constexpr (!FIRST && GR) {
                                glookup(gbase, pa[g], g1.x, g1.y, la);
                                glookup(gbase, pb[g], g1.x, g1.y, lb);
                            } else {
                                lookup(pt, pa[g], la, qa);
                                lookup(pt, pb[g], lb, qb);
                            }
                            if (FIRST) {
    #pragma unroll
                                for (int c = 0; c < 3; ++c) lv[g][c] = fmaf(qa, la[c], qb * lb[c]);
                            } else {
    #pragma unroll
                                for (int c = 0; c < 3; ++c) lv[g][c] = la[c] + lb[c];
                            }
                        }
                        if (FIRST && ground) {   // order 2: the separate ground term (uniform branch)
    #pragma unroll
                            for (int g = 0; g < G; ++g) {
                                float ea[3], eb[3];
                                ground_e(pa[g], g1.x, g1.y, ea);
                                ground_e(pb[g], g1.x, g1.y, eb);
                                lv[g][0] = fmaf(g0.x, ea[0] + eb[0], lv[g][0]);
                                lv[g][1] = fmaf(g0.y, ea[1] + eb[1], lv[g][1]);
                                lv[g][2] = fmaf(g0.z, ea[2] + eb[2], lv[g][2]);
                            }
                        }
    #pragma unroll
                        for (int g = 0; g < G; ++g) {
                            acc[g][0] = fmaf(lv[g][0], W0, acc[g][0]);
                            acc[g][1] = fmaf(lv[g][1], W1, acc[g][1]);
                            acc[g][2] = fmaf(lv[g][2], W2, acc[g][2]);
                        }
                    }
                };
                if (!FIRST && ground) phi_row(std::true_type{});
                else phi_row(std::false_type{});
            }
    #pragma unroll
            for (int g = 0; g < G; ++g) {
                const int k = kb + g;
                if (k < NU) {
                    const size_t o = ((size_t)l * s.RMU + y) * s.W4 + (size_t)k * s.RMUS + m;
                    if (!FIRST && last) {
                        // dJ is a sum of non-negative terms; a merged ground entry
                        // (FP32 P and Q) can undershoot a zero value by rounding
                        acc[g][0] = fmaxf(acc[g][0], 0.f); acc[g][1] = fmaxf(acc[g][1], 0.f);
                        acc[g][2] = fmaxf(acc[g][2], 0.f);
                    }
                    if (s.hs && last) {
                        acc[g][0] = q16(acc[g][0]); acc[g][1] = q16(acc[g][1]); acc[g][2] = q16(acc[g][2]);
                    }
                    if (s.dJb == 12) {
                        float *q = (float *)s.dJ4 + 3 * o;
                        q[0] = acc[g][0]; q[1] = acc[g][1]; q[2] = acc[g][2];
                    } else {
                        s.dJ4[o] = make_float4(acc[g][0], acc[g][1], acc[g][2], 0.f);
                    }
                }
            }
        }
        if (!CHUNK) break;
    }
}

// ---------------------------------------------------------------------------
// Pass dEn — irradianceN.glsl:85-114, fused with copyIrradiance k = 1 and its
// ADD blend (PreprocessAtmo.cpp:335-341).  One workgroup per E texel.
// ---------------------------------------------------------------------------
struct DirE {      // per-direction constants of the hemisphere grid
    float wx, wz, wzdw, pad;
};

// The NI/2 theta rows are staged CH at a time (CH = NI/2 unless they exceed
// the LDS budget).
template <bool FIRST>
__global__ __launch_bounds__(256) void k_irradianceN(const SetDesc *__restrict__ sets,
                                                     const uint2 *__restrict__ tiles, int NI, int CH,
                                                     const DirE *__restrict__ dirs) {
    constexpr int NT = FIRST ? 2 : 1;
    extern __shared__ float4 lds[];
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int t = (int)tl.y;
    const int i = t % s.SW, j = t / s.SW;
    const int NU = s.RNU, NP = NU + 1, NV = NU + 2;
    const int NTH = NI / 2;
    float4 *pairs = lds;                                   // [CH][NP][NT][2]
    float4 *vtmp = pairs + (size_t)CH * NP * NT * 2;       // [CH][NV][NT]
    float *red = (float *)(vtmp + (size_t)CH * NV * NT);   // [3][blockDim/64]

    const double h = (double)j / (s.SH - 1.0) * (s.Rt - s.Rg);
    const double r = s.Rg + h;
    const double muS = -0.2 + (double)i / (s.SW - 1.0) * (1.0 + 0.2);
    const double rho2 = h * (2.0 * s.Rg + h);
    const double txs = tex4d_umus(s, muS) * s.RMUS - 0.5;
    const float4 *tabs[2] = {s.dSR4, s.dSM4};
    const float sx = (float)sqrt(fmax(1.0 - muS * muS, 0.0)), sz = (float)muS;
    const float posA = 0.5f * (NU - 1), posB = 0.5f * (NU - 1) + 1.0f;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int c0 = 0; c0 < NTH; c0 += CH) {
        const int nth = min(CH, NTH - c0);
        if (c0 > 0) __syncthreads();   // every thread is done with the previous chunk's rows
        for (int e = threadIdx.x; e < nth * NV; e += blockDim.x) {
            const int it = c0 + e / NV, kk = e % NV;
            const double ct = cos((it + 0.5) * (kPI / NI));
            const double rmu = r * ct;
            double tz, ty;
            tex4d_rmu(s, rho2, rmu, rmu * rmu - rho2, ct, tz, ty);
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
        for (int e = threadIdx.x; e < nth * NP * NT; e += blockDim.x) {
            const int n = e % NT, ek = e / NT;
            const int it = ek / NP, kk = ek % NP;
            float4 a = vtmp[((size_t)it * NV + kk) * NT + n];
            float4 b = vtmp[((size_t)it * NV + kk + 1) * NT + n];
            pairs[((size_t)(it * NP + kk) * NT + n) * 2] = a;
            pairs[((size_t)(it * NP + kk) * NT + n) * 2 + 1] =
                make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.f);
        }
        __syncthreads();

        // directions d = phi * NTH + theta of this chunk's theta rows
        const int ndir = 2 * NI * nth;
        for (int dl = threadIdx.x; dl < ndir; dl += blockDim.x) {
            const int it = dl % nth;   // staged row
            const DirE w = dirs[(dl / nth) * NTH + c0 + it];
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
        if (s.hs) {   // dE stored as FP16, then the FP16 ADD blend E = FP16(E + dE)
            s0 = q16(s0); s1 = q16(s1); s2 = q16(s2);
        }
        s.dE4[t] = make_float4(s0, s1, s2, 1.0f);
        if (s.E4) {
            float4 e = s.E4[t];
            e.x += s0; e.y += s1; e.z += s2;
            e = q16(e, s.hs);
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
// ADD blend (S.rgb += dS / phaseR(nu)).
//
// The view ray (r, mu) of a texel depends only on its layer and mu row, so
// every quantity of an integration sample that does not involve the sun
// (sample radius and zenith cosine, the texture4D r and mu coordinates and
// filter weights, the transmittance T(r, mu, t)) is shared by the
// RNU x RMU_S texels of the row.  A workgroup covers up to 4 rows: it stages
// those per-sample terms in LDS (FP64 geometry, in chunks of 64 samples) and
// each thread then integrates its texels with only the mu_s-dependent work.
// ---------------------------------------------------------------------------
struct SnEntry {
    float4 w;     // bilinear (r, mu) weights of the 4 table rows
    int4 base;    // offsets of the 4 rows (z0y0, z0y1, z1y0, z1y1)
    float4 tw;    // T(r, mu, t) * trapezoid weight (rgb), t
    float4 ir;    // 1 / r_i, r_i - r
};

// The (r, mu) blend of one table column x: sum_q w_q dJ[row_q][x], in one fixed
// FP32 order.  Both forms of k_inscatterN (direct gathers, LDS-staged rows)
// blend with this, then lerp two columns in x as fma(ax, b1 - b0, b0), so
// they give the same bits.
__device__ __forceinline__ float3 sn_blend(const float4 &w, float3 t0, float3 t1, float3 t2, float3 t3) {
    return make_float3(fmaf(w.w, t3.x, fmaf(w.z, t2.x, fmaf(w.y, t1.x, w.x * t0.x))),
                       fmaf(w.w, t3.y, fmaf(w.z, t2.y, fmaf(w.y, t1.y, w.x * t0.y))),
                       fmaf(w.w, t3.z, fmaf(w.z, t2.z, fmaf(w.y, t1.z, w.x * t0.z))));
}
__device__ __forceinline__ float3 sn_lerp(float ax, float3 b0, float3 b1) {
    return make_float3(fmaf(ax, b1.x - b0.x, b0.x), fmaf(ax, b1.y - b0.y, b0.y), fmaf(ax, b1.z - b0.z, b0.z));
}

// Texel offset of the sample's mu_s_i = (nu t + mu_s r) / r_i inside a nu
// slice (texture4D's uMuS * RMUS - 0.5), evaluated relative to the texel's own
// mu_s.  Near the start of a long, nearly horizontal ray mu_s_i moves away from
// mu_s by 1e-5 and less per sample, and where a thin layer's dJ rises steeply
// out of the shadow between two mu_s columns the lookup is proportional to
// that move: an FP32 mu_s_i (~2e-8 absolute) put through the coordinate
// formula (whose + 0.74, resp. + 0.2, cancels at the low end) would give it
// only to a few 1e-3.  So:
//   linear (:165), mus = (mu_s + 0.2) r per texel: offset proportional to
//     mu_s_i + 0.2 = (nu t + (mu_s + 0.2) r + 0.2 (r_i - r)) / r_i;
//   non-linear (:161), mus = mu_s and txb = the texel's own offset (FP64): with
//     d = max(mu_s_i, -0.1975) - mu_b, mu_b = max(mu_s, -0.1975), and
//     mu_s_i - mu_s = (nu t - mu_s (r_i - r)) / r_i,
//     offset = txb + C atan(k d / (1 + k mu_b k (mu_b + d))), k = tan(1.386),
//     C = 0.5 (RMUS - 1) / 1.1 (atan a - atan b = atan((a - b) / (1 + a b)),
//     used while |k d| < 0.5, where 1 + a b > 0; larger moves take the
//     reference's form, which has no cancellation there).
template <bool LIN4>
__device__ __forceinline__ float sn_txs(const SetDesc &s, float nu, float mus, float txb,
                                        const SnEntry &en) {
    const float half = 0.5f * (float)(s.RMUS - 1);
    if (LIN4)
        return fmaxf(fmaf(nu, en.tw.w, fmaf(0.2f, en.ir.y, mus)) * en.ir.x, 0.0f) * (1.0f / 1.2f) *
               (2.0f * half);
    // Every sum of products is an explicit fmaf: left to contraction, the
    // compiler fuses kb + kd as fma(k, d, kb) in one inlined copy and as
    // fma(k, mub, kd) in another, and the paths of k_inscatterN must agree
    // bit for bit.
    const float mub = fmaxf(mus, -0.1975f);
    // d = mu_s_i - mu_s + (mu_s - mu_b), dmu = (nu t - mu_s (r_i - r)) / r_i
    const float d = fmaxf(fmaf(fmaf(nu, en.tw.w, -mus * en.ir.y), en.ir.x, mus - mub), -0.1975f - mub);
    const float k = s.f_tMuS, kb = k * mub, kd = k * d;
    const float kbd = fmaf(k, d, kb);   // kb + kd
    const bool near = fabsf(kd) < 0.5f;
    const float arg = near ? kd * __builtin_amdgcn_rcpf(fmaf(kb, kbd, 1.0f)) : kbd;
    return fmaf(atanf(arg), half * (1.0f / 1.1f), near ? txb : 0.74f * half);
}

// STAGED = false: each texel gathers its 8 taps per sample from dJ directly
// (global loads).  STAGED = true: per sample the workgroup first blends the
// 4 (r, mu) table rows of every row it covers into one row
// B_i[x] = sum_q w_q dJ[row_q][x] in LDS (coalesced loads, 4 per texel
// instead of 8, shared by all texels that hit x), then each texel reads two
// neighbours of B_i and lerps in x.  The direct form blends the two columns
// it needs in the same order (sn_blend, sn_lerp), so both give the same bits.
// `sb` samples are blended per LDS round (dynamic LDS: sb * rows * RNU*RMU_S
// float4).  A workgroup covers rows of RNU*RMU_S <= 256 texels with one texel
// per thread, or one row of more with 2-4 texels per thread: the staged form
// has a path for each.  In the second, a sample whose 4 rows are those of the
// sample before it in the same LDS round blends the loaded rows again (the
// row starts are workgroup-uniform there, so the test is a scalar branch).
// Three waves per SIMD (up to 168 VGPRs): at four (128 VGPRs) the staged form
// spills 116 bytes per lane, whose traffic shows in the kernel's HBM counters,
// and both builds run at the same speed (2.31 and 2.33 ms per launch at
// 64x256x64x16), the loads of a thread's texels being in flight together.
template <bool STAGED, bool LIN4>   // LIN4: linear inscatter parameterisation (per-sample mu_s)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_inscatterN(const SetDesc *__restrict__ sets,
                                                    const uint2 *__restrict__ tiles,
                                                    int nsamp, int rpw_cap, int sb) {
    __shared__ SnEntry ent[kNRows][kNChunk];   // exactly 16 KiB: two workgroups of 64 KiB
                                               // dynamic staging fit one CU's 160 KiB
    extern __shared__ float4 blend[];   // [sb][nrows * W4] (STAGED)
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int RPW = sn_rows_per_wg(s.W4, rpw_cap);
    const int ngy = (s.RMU + RPW - 1) / RPW;
    const int l = (int)tl.y / ngy;
    const int yb = ((int)tl.y % ngy) * RPW;
    const int nrows = min(RPW, s.RMU - yb);
    const LayerGeom L = s.layers[l];
    const double r = L.r;
    const double rho2 = r * r - s.Rg2;
    const int HW = s.RMU * s.W4;

    // texels of this thread: rows yb.. (row-major, x fastest)
    const int per = nrows * s.W4;
    const int nmine = (per + (int)blockDim.x - 1) / (int)blockDim.x;
    constexpr int MAXT = 4;   // W4 <= 1024 (validated on the host)
    float acc[MAXT][3];
    float nuf[MAXT], muSr[MAXT], txb[MAXT], fk[MAXT];
    int kq[MAXT];   // the texel's nu slice k0 (< RNU) | its row in the workgroup << 16
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        acc[j][0] = acc[j][1] = acc[j][2] = 0.f;
        nuf[j] = muSr[j] = txb[j] = fk[j] = 0.f;
        kq[j] = 0;
        const int e = threadIdx.x + j * (int)blockDim.x;
        if (j < nmine && e < per) {
            const int row = e / s.W4, x = e % s.W4;
            double mu, muS, nu;
            mu_mus_nu(s, L, x, yb + row, mu, muS, nu);
            // nu lies on a slice of the table: the nu lerp of texture4D has weight 0 or 1
            double pos = (nu + 1.0) / 2.0 * (s.RNU - 1.0);
            const double kr = rint(pos);
            if (fabs(pos - kr) < 1e-9) pos = kr;
            const double kfl = floor(pos);
            fk[j] = (float)(pos - kfl);
            kq[j] = (int)kfl | (row << 16);
            nuf[j] = (float)nu;
            // the texel's terms of its samples' mu_s offsets (sn_txs)
            muSr[j] = (float)(LIN4 ? (muS + 0.2) * r : muS);
            txb[j] = LIN4 ? 0.0f : (float)(tex4d_umus(s, muS) * s.RMUS - 0.5);
        }
    }

    // the dJ table as a buffer (< 2^31 bytes: at most 2^27 texels of 16 bytes)
    const __amdgpu_buffer_rsrc_t Jr = table_rsrc(s.dJ4, (unsigned)((size_t)s.RR * HW * s.dJb));
    // A lookup's texel column is kbase + txs: kbase the first column of its nu
    // slice, txs its offset, kept apart so that the fraction ax has full FP32
    // precision.  txs runs past RMU_S - 1 into the next slice when mu_s_i > 1
    // (as GL's coordinate does); only the row's last column clamps (CLAMP_TO_EDGE),
    // and i0 = kbase + is <= W4 - 2 keeps both taps in the row.
    const unsigned W4m2 = (unsigned)(s.W4 - 2);
    const unsigned jb = (unsigned)s.dJb;   // bytes per dJ texel
    for (int c0 = 0; c0 <= nsamp; c0 += kNChunk) {
        const int nc = min(kNChunk, nsamp + 1 - c0);
        __syncthreads();
        // stage the row-uniform terms of samples c0 .. c0+nc-1 (FP64 geometry)
        for (int e = threadIdx.x; e < nrows * kNChunk; e += blockDim.x) {
            const int row = e / kNChunk, js = e % kNChunk;
            if (js >= nc) continue;
            const int i = c0 + js;
            double mu, muS_u, nu_u;
            mu_mus_nu(s, L, 0, yb + row, mu, muS_u, nu_u);
            const double rmu = r * mu;
            const double delta = rmu * rmu - rho2;            // constant along the ray
            const double dx = limit_d(s, r, mu) / nsamp;
            const double ti = i * dx;
            const double q = rho2 + ti * (ti + 2.0 * rmu);   // ri^2 - Rg^2
            const double rmui = rmu + ti;                     // ri * mui
            const double ri = sqrt(q + s.Rg2);
            const double rho = sqrt(fmax(q, 0.0));
            double f, c;
            if (rmui < 0.0 && delta > 0.0) {
                f = -rho / (sqrt(delta) - rmui);
                c = 0.5 - 0.5 / s.RMU;
            } else {
                f = (sqrt(fmax(delta + s.H * s.H, 0.0)) - rmui) / (rho + s.H);
                c = 0.5 + 0.5 / s.RMU;
            }
            const double ty = s.lin4 ? (0.5 / s.RMU + (rmui / ri + 1.0) / 2.0 * (1.0 - 1.0 / s.RMU)) * s.RMU - 0.5
                                     : (c + f * (0.5 - 1.0 / s.RMU)) * s.RMU - 0.5;
            const double tz = rho / s.H * (s.RR - 1.0);
            int z0, z1, y0, y1;
            float az, ay;
            gl_axis_d(tz, s.RR, z0, z1, az);
            gl_axis_d(ty, s.RMU, y0, y1, ay);
            // transmittance(r, mu, ti) (common.glsl:272-282)
            const float h1 = (float)(q / (ri + s.Rg));
            const float mu1 = (float)(rmui / ri);
            const float3 Tseg = s.tview ? ld3(gptr(s.tview) + ((size_t)l * s.RMU + yb + row) * (nsamp + 1) + i)
                              : mu > 0.0 ? T_ratio_lin(s, (float)(r - s.Rg), (float)mu, h1, mu1)
                                         : T_ratio_lin(s, h1, -mu1, (float)(r - s.Rg), (float)-mu);
            const float wt = (i == 0 || i == nsamp) ? 0.5f : 1.0f;
            SnEntry en;
            en.w = make_float4((1.f - az) * (1.f - ay), (1.f - az) * ay, az * (1.f - ay), az * ay);
            en.base = make_int4((z0 * HW + y0 * s.W4) * s.dJb, (z0 * HW + y1 * s.W4) * s.dJb,
                                (z1 * HW + y0 * s.W4) * s.dJb, (z1 * HW + y1 * s.W4) * s.dJb);
            en.tw = make_float4(Tseg.x * wt, Tseg.y * wt, Tseg.z * wt, (float)ti);
            en.ir = make_float4((float)(1.0 / ri), (float)(ri - r), 0.f, 0.f);
            ent[row][js] = en;
        }
        __syncthreads();
        if (!STAGED) {
#pragma unroll
            for (int j = 0; j < MAXT; ++j) {
                if (j >= nmine) break;
                const int e = threadIdx.x + j * (int)blockDim.x;
                if (e >= per) break;
                const SnEntry *er = ent[kq[j] >> 16];
                const unsigned kbase = (unsigned)((kq[j] & 0xffff) * s.RMUS);
                const float txmax = (float)(W4m2 + 1 - kbase);
                const unsigned ismax = W4m2 - kbase;
                float a0 = acc[j][0], a1 = acc[j][1], a2 = acc[j][2];
                for (int js = 0; js < nc; ++js) {
                    const SnEntry en = er[js];
                    const float txs = __builtin_amdgcn_fmed3f(sn_txs<LIN4>(s, nuf[j], muSr[j], txb[j], en), 0.0f, txmax);
                    const unsigned is = min(__float2uint_rz(txs), ismax);
                    const float ax = txs - (float)is;
                    const unsigned i0 = kbase + is;
                    // the blended columns i0 and i0 + 1 (as the staged form's rows hold them)
                    auto column = [&](unsigned x) {
                        const unsigned o = x * jb;
                        return sn_blend(en.w, ld3b(Jr, (unsigned)en.base.x + o, 0u), ld3b(Jr, (unsigned)en.base.y + o, 0u),
                                        ld3b(Jr, (unsigned)en.base.z + o, 0u), ld3b(Jr, (unsigned)en.base.w + o, 0u));
                    };
                    float3 v = sn_lerp(ax, column(i0), column(i0 + 1));
                    if (fk[j] != 0.0f) {  // off-grid nu: blend with the next slice (k0 + 1 <= RNU - 1)
                        unsigned j0 = i0 + (unsigned)s.RMUS;
                        float bx = ax;
                        if (j0 > W4m2) { j0 = W4m2; bx = 1.0f; }   // past the row's last column
                        v = lerp3(v, sn_lerp(bx, column(j0), column(j0 + 1)), fk[j]);
                    }
                    a0 = fmaf(v.x, en.tw.x, a0);
                    a1 = fmaf(v.y, en.tw.y, a1);
                    a2 = fmaf(v.z, en.tw.z, a2);
                }
                acc[j][0] = a0; acc[j][1] = a1; acc[j][2] = a2;
            }
        } else if (nmine >= 2) {
            // Rows of more than 256 texels (one row per workgroup, 2-4 texels
            // per thread): every slot reads the same sample entries, so (a)
            // issues all 4 * NM table loads of a thread before blending any,
            // and (b) interleaves the NM independent lookups of a sample.  The
            // arithmetic of each texel-sample is that of the general path.
            auto one_row = [&](auto nmc) {
                constexpr int NM = decltype(nmc)::value;
                for (int s0 = 0; s0 < nc; s0 += sb) {
                    const int ns = min(sb, nc - s0);
                    // the table rows of the last sample loaded in this round: a later
                    // sample on the same 4 rows (same first and last row start, which fix
                    // the two layers and the two mu rows) blends them again without
                    // loading (kept within a round: carried over the lookups, t would
                    // spill at three waves per SIMD)
                    float3 t[NM][4];
                    unsigned P0 = 0u, P3 = 0u;
                    for (int k = 0; k < ns; ++k) {
                        const float4 w = ent[0][s0 + k].w;
                        const int4 base = ent[0][s0 + k].base;
                        // the 4 row starts are the same for the whole workgroup: in
                        // scalar registers, the loads take them as their base and
                        // the thread's own column offset (fixed over the samples)
                        const unsigned R0 = (unsigned)__builtin_amdgcn_readfirstlane(base.x);
                        const unsigned R1 = (unsigned)__builtin_amdgcn_readfirstlane(base.y);
                        const unsigned R2 = (unsigned)__builtin_amdgcn_readfirstlane(base.z);
                        const unsigned R3 = (unsigned)__builtin_amdgcn_readfirstlane(base.w);
                        if (k == 0 || R0 != P0 || R3 != P3) {
                            P0 = R0;
                            P3 = R3;
#pragma unroll
                            for (int j = 0; j < NM; ++j) {
                                // past the row's end: a valid address, the blend is not stored
                                const unsigned xo = (unsigned)min((int)threadIdx.x + j * (int)blockDim.x, per - 1) * jb;
                                t[j][0] = ld3b(Jr, xo, R0);
                                t[j][1] = ld3b(Jr, xo, R1);
                                t[j][2] = ld3b(Jr, xo, R2);
                                t[j][3] = ld3b(Jr, xo, R3);
                            }
                        }
#pragma unroll
                        for (int j = 0; j < NM; ++j) {
                            const int e = threadIdx.x + j * (int)blockDim.x;
                            const float3 b = sn_blend(w, t[j][0], t[j][1], t[j][2], t[j][3]);
                            if (e < per) blend[(size_t)k * per + e] = make_float4(b.x, b.y, b.z, 0.f);
                        }
                    }
                    __syncthreads();
                    for (int k = 0; k < ns; ++k) {
                        const SnEntry en = ent[0][s0 + k];
                        const float4 *B = blend + (size_t)k * per;
                        float3 v[NM];
                        unsigned i0[NM];
                        float ax[NM];
#pragma unroll
                        for (int j = 0; j < NM; ++j) {
                            // slots past the row's end (kq 0) read in-row values and are never stored
                            const unsigned kbase = (unsigned)((kq[j] & 0xffff) * s.RMUS);
                            const float txs = __builtin_amdgcn_fmed3f(sn_txs<LIN4>(s, nuf[j], muSr[j], txb[j], en), 0.0f,
                                                                      (float)(W4m2 + 1 - kbase));
                            const unsigned is = min(__float2uint_rz(txs), W4m2 - kbase);
                            ax[j] = txs - (float)is;
                            i0[j] = kbase + is;
                            v[j] = sn_lerp(ax[j], f3(B[i0[j]]), f3(B[i0[j] + 1]));
                        }
#pragma unroll
                        for (int j = 0; j < NM; ++j) {
                            if (fk[j] != 0.0f) {  // off-grid nu: blend with the next slice (k0 + 1 <= RNU - 1)
                                unsigned j0 = i0[j] + (unsigned)s.RMUS;
                                float bx = ax[j];
                                if (j0 > W4m2) { j0 = W4m2; bx = 1.0f; }   // past the row's last column
                                v[j] = lerp3(v[j], sn_lerp(bx, f3(B[j0]), f3(B[j0 + 1])), fk[j]);
                            }
                            acc[j][0] = fmaf(v[j].x, en.tw.x, acc[j][0]);
                            acc[j][1] = fmaf(v[j].y, en.tw.y, acc[j][1]);
                            acc[j][2] = fmaf(v[j].z, en.tw.z, acc[j][2]);
                        }
                    }
                    __syncthreads();
                }
            };
            if (nmine == 2) one_row(std::integral_constant<int, 2>{});
            else if (nmine == 3) one_row(std::integral_constant<int, 3>{});
            else one_row(std::integral_constant<int, MAXT>{});
        } else if (nmine == 1) {
            // Rows of at most 256 texels (one texel per thread): samples are
            // blended and looked up in pairs, so a thread has 8 table loads and
            // two independent lookups in flight; the accumulation keeps sample
            // order, and each texel-sample's arithmetic is the general path's.
            const int e = threadIdx.x;
            const bool live = e < per;
            const int rowj = kq[0] >> 16;   // 0 past the rows' end: valid addresses, nothing stored
            const SnEntry *er = ent[rowj];
            const unsigned xo = (unsigned)(live ? e - rowj * s.W4 : 0) * jb;
            const unsigned kbase = (unsigned)((kq[0] & 0xffff) * s.RMUS);
            const float txmax = (float)(W4m2 + 1 - kbase);
            const unsigned ismax = W4m2 - kbase;
            const float4 *Brow = blend + rowj * s.W4;
            auto group = [&](int s0, int k, auto pc) {   // samples s0+k .. s0+k+P-1
                constexpr int P = decltype(pc)::value;
                float3 t[P][4];
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    // the row starts differ between the rows of a workgroup: in the VGPR offset
                    const int4 base = er[s0 + k + q].base;
                    t[q][0] = ld3b(Jr, (unsigned)base.x + xo, 0u);
                    t[q][1] = ld3b(Jr, (unsigned)base.y + xo, 0u);
                    t[q][2] = ld3b(Jr, (unsigned)base.z + xo, 0u);
                    t[q][3] = ld3b(Jr, (unsigned)base.w + xo, 0u);
                }
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float3 b = sn_blend(er[s0 + k + q].w, t[q][0], t[q][1], t[q][2], t[q][3]);
                    if (live) blend[(size_t)(k + q) * per + e] = make_float4(b.x, b.y, b.z, 0.f);
                }
            };
            auto lookup = [&](int s0, int k, auto pc) {
                constexpr int P = decltype(pc)::value;
                float3 v[P];
                unsigned i0[P];
                float ax[P];
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const float txs = __builtin_amdgcn_fmed3f(
                        sn_txs<LIN4>(s, nuf[0], muSr[0], txb[0], er[s0 + k + q]), 0.0f, txmax);
                    const unsigned is = min(__float2uint_rz(txs), ismax);
                    ax[q] = txs - (float)is;
                    i0[q] = kbase + is;
                    const float4 *B = Brow + (size_t)(k + q) * per;
                    v[q] = sn_lerp(ax[q], f3(B[i0[q]]), f3(B[i0[q] + 1]));
                }
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    if (fk[0] != 0.0f) {  // off-grid nu: blend with the next slice (k0 + 1 <= RNU - 1)
                        const float4 *B = Brow + (size_t)(k + q) * per;
                        unsigned j0 = i0[q] + (unsigned)s.RMUS;
                        float bx = ax[q];
                        if (j0 > W4m2) { j0 = W4m2; bx = 1.0f; }   // past the row's last column
                        v[q] = lerp3(v[q], sn_lerp(bx, f3(B[j0]), f3(B[j0 + 1])), fk[0]);
                    }
                    const float4 tw = er[s0 + k + q].tw;
                    acc[0][0] = fmaf(v[q].x, tw.x, acc[0][0]);
                    acc[0][1] = fmaf(v[q].y, tw.y, acc[0][1]);
                    acc[0][2] = fmaf(v[q].z, tw.z, acc[0][2]);
                }
            };
            for (int s0 = 0; s0 < nc; s0 += sb) {
                const int ns = min(sb, nc - s0);
                int k = 0;
                for (; k + 1 < ns; k += 2) group(s0, k, std::integral_constant<int, 2>{});
                if (k < ns) group(s0, k, std::integral_constant<int, 1>{});
                __syncthreads();
                k = 0;
                for (; k + 1 < ns; k += 2) lookup(s0, k, std::integral_constant<int, 2>{});
                if (k < ns) lookup(s0, k, std::integral_constant<int, 1>{});
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        if (j >= nmine) break;
        const int e = threadIdx.x + j * (int)blockDim.x;
        if (e >= per) break;
        const int row = e / s.W4, x = e % s.W4;
        const size_t t = ((size_t)l * s.RMU + yb + row) * s.W4 + x;
        // the row's trapezoid step, as the sample staging computes it (mu depends on the row only)
        double mu, muS_u, nu_u;
        mu_mus_nu(s, L, 0, yb + row, mu, muS_u, nu_u);
        const float dx = (float)(limit_d(s, r, mu) / nsamp);
        const float4 o = q16(make_float4(acc[j][0] * dx, acc[j][1] * dx, acc[j][2] * dx, 0.f), s.hs);
        s.dSR4[t] = o;
        if (s.S4) {
            const double nu = nuf[j];
            const float pr = (float)((3.0 / (16.0 * kPI)) * (1.0 + nu * nu));
            float4 S = s.S4[t];
            S.x += o.x / pr;
            S.y += o.y / pr;
            S.z += o.z / pr;
            s.S4[t] = q16(S, s.hs);   // FP16 ADD blend: S = FP16(S + dS / pR)
        }
    }
}

// ---------------------------------------------------------------------------
// Transmittance tables of one precompute (after the T pass, before the passes
// that use them): every ratio T(x)/T(y) of common.glsl:272-282 the later passes
// take depends only on the T table and the geometry, not on the order.
//  k_tview:   T(r, mu, t_i) along the view ray of each (layer, mu row) at the
//             inscatter samples t_i = i limit(r, mu) / nsamp (inscatter1.glsl,
//             inscatterN.glsl), same tiles as k_inscatter1;
//  k_tground: T from the ground point seen in direction theta to the layer
//             radius (inscatterS.glsl:129-138), per (layer, theta row); run on
//             the dJ tiles, by the workgroup of mu_s column 0 of each layer.
// Both use T_ratio_d (FP64 coordinates; underflow-free with the optical-depth
// table).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tview(const SetDesc *__restrict__ sets,
                                               const uint2 *__restrict__ tiles, int nsamp) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    const int RPW = sn_rows_per_wg(s.W4, kNRows);
    const int ngy = (s.RMU + RPW - 1) / RPW;
    const int l = (int)tl.y / ngy;
    const int yb = ((int)tl.y % ngy) * RPW;
    const int nrows = min(RPW, s.RMU - yb);
    const LayerGeom L = s.layers[l];
    const double r = L.r;
    const double rho2 = r * r - s.Rg2;
    for (int e = threadIdx.x; e < nrows * (nsamp + 1); e += blockDim.x) {
        const int row = e / (nsamp + 1), i = e % (nsamp + 1);
        double mu, muS_u, nu_u;
        mu_mus_nu(s, L, 0, yb + row, mu, muS_u, nu_u);
        const double rmu = r * mu;
        const double ti = i * (limit_d(s, r, mu) / nsamp);
        const double q = rho2 + ti * (ti + 2.0 * rmu);   // ri^2 - Rg^2
        const double ri = sqrt(q + s.Rg2);
        const double h1 = q / (ri + s.Rg);
        const double mu1 = (rmu + ti) / ri;
        const float3 t = mu > 0.0 ? T_ratio_d(s, r - s.Rg, mu, h1, mu1)
                                  : T_ratio_d(s, h1, -mu1, r - s.Rg, -mu);
        s.tview[((size_t)l * s.RMU + yb + row) * (nsamp + 1) + i] = make_float4(t.x, t.y, t.z, 0.f);
    }
}

__global__ __launch_bounds__(256) void k_tground(const SetDesc *__restrict__ sets,
                                                 const uint2 *__restrict__ tiles, int NS) {
    const uint2 tl = tiles[blockIdx.x];
    const SetDesc &s = sets[tl.x];
    if ((int)tl.y % s.RMUS != 0) return;   // one workgroup per layer
    const int l = (int)tl.y / s.RMUS;
    const double r = fmin(fmax(s.layers[l].r, s.Rg), s.Rt);   // inscatterS.glsl:112
    for (int it = threadIdx.x; it < NS; it += blockDim.x) {
        const double ct = cos((it + 0.5) * (kPI / NS));
        const double q = s.Rg / r;
        float3 tr = make_float3(0.f, 0.f, 0.f);
        if (ct < -sqrt(1.0 - q * q)) {
            const double dg = -r * ct - sqrt(r * r * (ct * ct - 1.0) + s.Rg2);
            const double mug = -(r * ct + dg) / s.Rg;
            const double r1sq = s.Rg2 + dg * dg + 2.0 * s.Rg * mug * dg;
            const double r1 = sqrt(r1sq);
            const double mu1 = (s.Rg * mug + dg) / r1;
            const double h1 = (r1sq - s.Rg2) / (r1 + s.Rg);
            tr = mug > 0.0 ? T_ratio_d(s, 0.0, mug, h1, mu1) : T_ratio_d(s, h1, -mu1, 0.0, -mug);
        }
        s.tground[(size_t)l * NS + it] = make_float4(tr.x, tr.y, tr.z, 0.f);
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
bool inscatterS_chunked(int NS, int CH, int RNU, bool first) {
    return CH < NS || (first && (size_t)CH * (RNU + 1) > (size_t)kJ1Plane);
}
size_t lds_inscatterS(int NS, int CH, int RNU, int SW, int JX, bool first) {
    const int NT = first ? 2 : 1, PW = first ? 3 : 2;
    const size_t npairs = (size_t)CH * (RNU + 1) + (first ? 0 : (size_t)JX);
    const bool planes = first && !inscatterS_chunked(NS, CH, RNU, first);
    return sizeof(float4) * ((planes ? 2 * (size_t)kJ1Plane + npairs : npairs * PW) +
                             (planes ? 0 : (size_t)CH * (RNU + 2) * NT) +
                             (first ? (size_t)(e_pad(SW) + SW + 1) * 2 : 0) + (size_t)CH * 3);
}
int merged_row_extra(double r, double Rg, double Rt, int SW, int NS, int it) {
    // the kernel's r, clamped as inscatterS.glsl:112 does, and its ground test
    r = std::min(std::max(r, Rg), Rt);
    const double ct = std::cos((it + 0.5) * (kPI / NS));
    if (ct >= 1e-9) return 0;   // theta < 90 deg: never a ground row
    const double q = Rg / r, cthetamin = -std::sqrt(1.0 - q * q);
    // ground distance, or (where the device's ground test could round either
    // way) the tangent distance, which bounds the ground distance of every ray
    // that hits the ground
    const double dg = ct < cthetamin - 1e-9 ? -r * ct - std::sqrt(r * r * (ct * ct - 1.0) + Rg * Rg)
                                            : std::sqrt(std::max(r * r - Rg * Rg, 0.0));
    const double e1 = dg / Rg * (SW / 1.2);
    // the kernel keeps at most ceil(2.05 e1) + 1 extra entries in a row
    return (int)std::ceil(2.05 * e1) + 2;
}
size_t lds_irradianceN(int CH, int RNU, bool first) {
    const int NT = first ? 2 : 1;
    return sizeof(float4) * ((size_t)CH * (RNU + 1) * NT * 2 + (size_t)CH * (RNU + 2) * NT) +
           sizeof(float) * 3 * 16;
}
int chunk_inscatterS(int NS, int RNU, int SW, const std::function<int(int)> &jx_of_ch, bool first) {
    for (int ch = NS; ch >= 1; --ch)
        if (lds_inscatterS(NS, ch, RNU, SW, first ? 0 : jx_of_ch(ch), first) <= kSphLdsBudget) return ch;
    return 0;
}
int chunk_irradianceN(int NI, int RNU, bool first) {
    for (int ch = NI / 2; ch >= 1; --ch)
        if (lds_irradianceN(ch, RNU, first) <= kSphLdsBudget) return ch;
    return 0;
}

hipError_t launch_transmittance(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                                hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_transmittance, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_tview(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tview, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_tground(const SetDesc *sets, const uint2 *tiles, int ntiles, int NS, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tground, dim3(ntiles), dim3(64), 0, st, sets, tiles, NS);
    return hipGetLastError();
}
hipError_t launch_irradiance1(const SetDesc *sets, const uint2 *tiles, int ntiles,
                              hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_irradiance1, dim3(ntiles), dim3(256), 0, st, sets, tiles);
    return hipGetLastError();
}
hipError_t launch_inscatter1(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             bool linT, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (linT)
        hipLaunchKernelGGL(k_inscatter1<true>, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    else
        hipLaunchKernelGGL(k_inscatter1<false>, dim3(ntiles), dim3(256), 0, st, sets, tiles, nsamp);
    return hipGetLastError();
}
hipError_t launch_inscatterS(const SetDesc *sets, const uint2 *tiles, int ntiles, int NS, int CH, int RNU,
                             int JX, const void *dirT, const void *dirW, bool first, size_t lds, int G,
                             hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (CH < 1 || CH > NS || JX < 0) return hipErrorInvalidValue;
#define ATMO_LJ(F, GG, C)                                                                          \
    hipLaunchKernelGGL((k_inscatterS<F, GG, C>), dim3(ntiles), dim3(256), lds, st, sets, tiles, NS, \
                       CH, JX, (const DirJ *)dirT, (const float2 *)dirW)
    if (G != 2 && G != 3 && G != 4 && G != 8) return hipErrorInvalidValue;
    if (inscatterS_chunked(NS, CH, RNU, first)) {   // chunked staging: one variant (all G give the same bits)
        if (first) ATMO_LJ(true, 4, true); else ATMO_LJ(false, 4, true);
    } else if (G == 2) {
        if (first) ATMO_LJ(true, 2, false); else ATMO_LJ(false, 2, false);
    } else if (G == 3) {
        if (first) ATMO_LJ(true, 3, false); else ATMO_LJ(false, 3, false);
    } else if (G == 4) {
        if (first) ATMO_LJ(true, 4, false); else ATMO_LJ(false, 4, false);
    } else if (G == 8) {
        if (first) ATMO_LJ(true, 8, false); else ATMO_LJ(false, 8, false);
    } else {
        return hipErrorInvalidValue;
    }
#undef ATMO_LJ
    return hipGetLastError();
}
hipError_t launch_irradianceN(const SetDesc *sets, const uint2 *tiles, int ntiles, int NI, int CH,
                              const void *dirs, bool first, size_t lds, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (CH < 1 || CH > NI / 2) return hipErrorInvalidValue;
    if (first)
        hipLaunchKernelGGL(k_irradianceN<true>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NI,
                           CH, (const DirE *)dirs);
    else
        hipLaunchKernelGGL(k_irradianceN<false>, dim3(ntiles), dim3(256), lds, st, sets, tiles, NI,
                           CH, (const DirE *)dirs);
    return hipGetLastError();
}
size_t static_lds_inscatterN() { return sizeof(SnEntry) * kNRows * kNChunk; }
size_t lds_inscatterN(int max_row_texels, int sb) {
    return sb > 0 ? sizeof(float4) * (size_t)sb * (size_t)max_row_texels : 0;
}
hipError_t launch_inscatterN(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             int rpw_cap, int sb, size_t lds, bool lin4, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
#define ATMO_LN(S, L, LDS, SB) \
    hipLaunchKernelGGL((k_inscatterN<S, L>), dim3(ntiles), dim3(256), LDS, st, sets, tiles, nsamp, rpw_cap, SB)
    if (sb > 0) {
        if (lin4) ATMO_LN(true, true, lds, sb); else ATMO_LN(true, false, lds, sb);
    } else {
        if (lin4) ATMO_LN(false, true, 0, 0); else ATMO_LN(false, false, 0, 0);
    }
#undef ATMO_LN
    return hipGetLastError();
}
hipError_t launch_check(const SetDesc *sets, const uint2 *tiles, int ntiles, unsigned int *flags,
                        hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check, dim3(ntiles), dim3(256), 0, st, sets, tiles, flags);
    return hipGetLastError();
}

}  // namespace atmo
