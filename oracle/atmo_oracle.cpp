/*
 * atmo_oracle.cpp — FP64 CPU oracle of Proland's atmo preprocess.
 *
 * TEST INFRASTRUCTURE ONLY (the reference results the tests compare against;
 * never part of the product).
 *
 * A double-precision restatement, function by function, of
 *   src/atmo/sources/proland/preprocess/atmo/common.glsl        (shared math)
 *   src/atmo/sources/proland/preprocess/atmo/{transmittance,irradiance1,
 *       inscatter1,copyIrradiance,copyInscatter1,inscatterS,irradianceN,
 *       inscatterN,copyInscatterN}.glsl                          (the passes)
 *   src/atmo/sources/proland/preprocess/atmo/PreprocessAtmo.cpp  (setLayer,
 *       pass order, blending)
 * plus an emulation of the GL texture unit those shaders sample through
 * (GL 3.3 core, 3.8.11: LINEAR filter, texel centres at (i+0.5)/N,
 * CLAMP_TO_EDGE — the wrap the table consumers declare in
 * src/atmo/examples/atmo/helloworld.xml:71-75).
 *
 * Deviations from a literal transcription, each needed only where the GLSL
 * would take sqrt of a negative number or divide 0 by 0 (the GL output there
 * is undefined):
 *   D1 texture4D: rho = sqrt(max(r^2 - Rg^2, 0)) and sqrt(max(delta + H^2, 0)).
 *   D2 texture4D ground branch: uMu uses -rho / (sqrt(delta) - r mu), which is
 *      algebraically equal to (r mu + sqrt(delta)) / rho (multiply by
 *      sqrt(delta) - r mu > 0) and finite at rho = 0.  The literal form is kept
 *      in oracle_texture4D_coords_literal and tested equal away from rho = 0.
 *   D3 getTransmittanceUV: sqrt(max(r - Rg, 0) / (Rt - Rg)).
 *   D4 copy passes read the texel they sample at its centre directly (in exact
 *      arithmetic a LINEAR fetch at a texel centre is that texel).
 *   D5 the transmittance ratios min(T(x)/T(y), 1) (common.glsl:272-282 and the
 *      consumer's) take min as GPUs execute GLSL min, IEEE minNum: a 0/0 ratio
 *      (both lookups 0: FP16 storage, or T underflowing) gives 1.
 *   D6 in oracle_precompute (FP64 tables) the segment ratio is evaluated from
 *      the optical depths tau = -ln T of the T texels: T(x) = e^-m sum w_i
 *      e^(m - tau_i), m the smallest depth among the taps of nonzero weight.
 *      Equal to the ratio of the bilinear lookups in exact arithmetic; finite
 *      where both lookups are below FP64's range (tau > 708).
 * Uniforms are FP32 in the reference; the oracle takes the FP32 parameter
 * values and the FP32-rounded per-layer r and dhdH (PreprocessAtmo.cpp:255-271
 * computes them in double and uploads floats) and does everything else in FP64.
 */
#include "atmo_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* common.glsl:99  (not pi: 3.141592657) */
const double kPI = 3.141592657;

struct P {
    double Rg, Rt, RL, alpha, HR, HM, mieG;
    double bR[3], bMS[3], bME[3];
    int TW, TH, SW, SH, RR, RMU, RMUS, RNU;
    int orders, tS, iS, irrS, sphS;
    bool linT, lin4;   /* linear parameterisations (common.glsl:105-106 undefined) */
    const double *tau; /* optical depths -ln T of the T texels (D6), or null */
};

P mk(const atmo_params *a, const atmo_options *o) {
    P p;
    p.Rg = a->Rg; p.Rt = a->Rt; p.RL = a->RL;
    p.alpha = a->AVERAGE_GROUND_REFLECTANCE;
    p.HR = a->HR; p.HM = a->HM; p.mieG = a->mieG;
    p.bR[0] = a->betaR.x; p.bR[1] = a->betaR.y; p.bR[2] = a->betaR.z;
    p.bMS[0] = a->betaMSca.x; p.bMS[1] = a->betaMSca.y; p.bMS[2] = a->betaMSca.z;
    p.bME[0] = a->betaMEx.x; p.bME[1] = a->betaMEx.y; p.bME[2] = a->betaMEx.z;
    p.TW = a->TRANSMITTANCE_W; p.TH = a->TRANSMITTANCE_H;
    p.SW = a->SKY_W; p.SH = a->SKY_H;
    p.RR = a->RES_R; p.RMU = a->RES_MU; p.RMUS = a->RES_MU_S; p.RNU = a->RES_NU;
    /* common.glsl:94-97 and PreprocessAtmo.cpp:351 */
    p.orders = 4; p.tS = 500; p.iS = 50; p.irrS = 32; p.sphS = 16;
    p.tau = nullptr;
    p.linT = o && (o->flags & ATMO_FLAG_LINEAR_TRANSMITTANCE);
    p.lin4 = o && (o->flags & ATMO_FLAG_LINEAR_INSCATTER);
    if (o) {
        p.orders = o->orders; p.tS = o->transmittance_samples; p.iS = o->inscatter_samples;
        p.irrS = o->irradiance_samples; p.sphS = o->spherical_samples;
    }
    return p;
}

int g_threads = 0;

inline double clampd(double x, double a, double b) { return std::min(std::max(x, a), b); }

/* GL 3.3 3.8.11, one axis of LINEAR + CLAMP_TO_EDGE: texel coordinate
 * u*N - 0.5, lower index floor(.), weight of the upper texel frac(.). */
inline void axis(double u, int N, int &i0, int &i1, double &a) {
    double x = u * N - 0.5;
    x = clampd(x, -4.0, N + 4.0); /* out-of-range coordinates all clamp to an edge texel */
    double f = std::floor(x);
    a = x - f;
    int k = (int)f;
    i0 = std::min(std::max(k, 0), N - 1);
    i1 = std::min(std::max(k + 1, 0), N - 1);
}

void sample2(const double *tex, int W, int H, int C, double s, double t, double *out) {
    int i0, i1, j0, j1;
    double a, b;
    axis(s, W, i0, i1, a);
    axis(t, H, j0, j1, b);
    for (int c = 0; c < C; ++c) {
        double v00 = tex[((size_t)j0 * W + i0) * C + c], v01 = tex[((size_t)j0 * W + i1) * C + c];
        double v10 = tex[((size_t)j1 * W + i0) * C + c], v11 = tex[((size_t)j1 * W + i1) * C + c];
        out[c] = (1 - a) * (1 - b) * v00 + a * (1 - b) * v01 + (1 - a) * b * v10 + a * b * v11;
    }
}

void sample3(const double *tex, int W, int H, int D, int C, double s, double t, double u,
             double *out) {
    int i0, i1, j0, j1, k0, k1;
    double a, b, g;
    axis(s, W, i0, i1, a);
    axis(t, H, j0, j1, b);
    axis(u, D, k0, k1, g);
    for (int c = 0; c < C; ++c) {
        auto at = [&](int k, int j, int i) { return tex[(((size_t)k * H + j) * W + i) * C + c]; };
        out[c] = (1 - g) * ((1 - b) * ((1 - a) * at(k0, j0, i0) + a * at(k0, j0, i1)) +
                            b * ((1 - a) * at(k0, j1, i0) + a * at(k0, j1, i1))) +
                 g * ((1 - b) * ((1 - a) * at(k1, j0, i0) + a * at(k1, j0, i1)) +
                      b * ((1 - a) * at(k1, j1, i0) + a * at(k1, j1, i1)));
    }
}

/* PreprocessAtmo.cpp:255-271 setLayer.  The float*float products are float
 * in the reference (params are float members); r and dhdH go to the shader as
 * FP32 uniforms. */
void layer(const P &p, int l, double &r_out, double dhdH[4]) {
    float Rg = (float)p.Rg, Rt = (float)p.Rt;
    double r = l / (p.RR - 1.0);
    r = r * r;
    float RgRg = Rg * Rg, RtRt = Rt * Rt;
    float dRR = RtRt - RgRg;
    r = std::sqrt((double)RgRg + r * (double)dRR) +
        (l == 0 ? 0.01 : (l == p.RR - 1 ? -0.001 : 0.0));
    double dmin = Rt - r;
    double dmax = std::sqrt(r * r - (double)RgRg) + std::sqrt((double)dRR);
    double dminp = r - Rg;
    double dmaxp = std::sqrt(r * r - (double)RgRg);
    r_out = (double)(float)r;
    dhdH[0] = (double)(float)dmin;
    dhdH[1] = (double)(float)dmax;
    dhdH[2] = (double)(float)dminp;
    dhdH[3] = (double)(float)dmaxp;
}

/* common.glsl:209-219 limit */
double limit(const P &p, double r, double mu) {
    double dout = -r * mu + std::sqrt(r * r * (mu * mu - 1.0) + p.RL * p.RL);
    double delta2 = r * r * (mu * mu - 1.0) + p.Rg * p.Rg;
    if (delta2 >= 0.0) {
        double din = -r * mu - std::sqrt(delta2);
        if (din >= 0.0) dout = std::min(dout, din);
    }
    return dout;
}

/* common.glsl:290-292 */
double phaseR(double mu) { return (3.0 / (16.0 * kPI)) * (1.0 + mu * mu); }
/* common.glsl:295-297 */
double phaseM(const P &p, double mu) {
    double g = p.mieG;
    return 1.5 * 1.0 / (4.0 * kPI) * (1.0 - g * g) * std::pow(1.0 + g * g - 2.0 * g * mu, -1.5) *
           (1.0 + mu * mu) / (2.0 + g * g);
}

/* common.glsl:114-124 getTransmittanceUV, D3 (TRANSMITTANCE_NON_LINEAR; the
 * #else branch :119-120 with linT) */
void transmittanceUV(const P &p, double r, double mu, double &u, double &v) {
    if (p.linT) {
        u = (mu + 0.15) / (1.0 + 0.15);
        v = (r - p.Rg) / (p.Rt - p.Rg);
        return;
    }
    double uR = std::sqrt(std::max(r - p.Rg, 0.0) / (p.Rt - p.Rg));
    double uMu = std::atan((mu + 0.15) / (1.0 + 0.15) * std::tan(1.5)) / 1.5;
    u = uMu;
    v = uR;
}

/* common.glsl:126-136 getTransmittanceRMu, texel (i, j): gl_FragCoord = (i+.5, j+.5) */
void transmittanceRMu(const P &p, int i, int j, double &r, double &mu) {
    double rr = (j + 0.5) / p.TH;
    double m = (i + 0.5) / p.TW;
    if (p.linT) {   /* :132-133 */
        r = p.Rg + rr * (p.Rt - p.Rg);
        mu = -0.15 + m * (1.0 + 0.15);
        return;
    }
    r = p.Rg + (rr * rr) * (p.Rt - p.Rg);
    mu = -0.15 + std::tan(1.5 * m) / std::tan(1.5) * (1.0 + 0.15);
}

/* common.glsl:144-147 getIrradianceRMuS */
void irradianceRMuS(const P &p, int i, int j, double &r, double &muS) {
    r = p.Rg + ((j + 0.5) - 0.5) / (p.SH - 1.0) * (p.Rt - p.Rg);
    muS = -0.2 + ((i + 0.5) - 0.5) / (p.SW - 1.0) * (1.0 + 0.2);
}

/* common.glsl:223-226 transmittance(r, mu): bilinear T lookup */
void T_lookup(const P &p, const double *T, double r, double mu, double out[3]) {
    double u, v;
    transmittanceUV(p, r, mu, u, v);
    sample2(T, p.TW, p.TH, 3, u, v, out);
}

/* ln T(r, mu) per channel from the optical-depth table (D6): -m + ln S */
void T_log_lookup(const P &p, double r, double mu, double m[3], double S[3]) {
    double u, v;
    transmittanceUV(p, r, mu, u, v);
    int i0, i1, j0, j1;
    double a, b;
    axis(u, p.TW, i0, i1, a);
    axis(v, p.TH, j0, j1, b);
    const size_t idx[4] = {(size_t)j0 * p.TW + i0, (size_t)j0 * p.TW + i1, (size_t)j1 * p.TW + i0,
                           (size_t)j1 * p.TW + i1};
    const double w[4] = {(1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b};
    for (int c = 0; c < 3; ++c) {
        double mc = INFINITY, Sc = 0.0;
        for (int q = 0; q < 4; ++q)
            if (w[q] > 0.0) mc = std::min(mc, p.tau[idx[q] * 3 + c]);
        for (int q = 0; q < 4; ++q)
            if (w[q] > 0.0) Sc += w[q] * std::exp(mc - p.tau[idx[q] * 3 + c]);
        m[c] = mc;
        S[c] = Sc;
    }
}

/* common.glsl:272-282 transmittance(r, mu, d) between x and x + d*v */
void T_segment(const P &p, const double *T, double r, double mu, double d, double out[3]) {
    if (p.tau) {   /* D6 */
        const double r1 = std::sqrt(r * r + d * d + 2.0 * r * mu * d);
        const double mu1 = (r * mu + d) / r1;
        double ma[3], Sa[3], mb[3], Sb[3];
        if (mu > 0.0) {
            T_log_lookup(p, r, mu, ma, Sa);
            T_log_lookup(p, r1, mu1, mb, Sb);
        } else {
            T_log_lookup(p, r1, -mu1, ma, Sa);
            T_log_lookup(p, r, -mu, mb, Sb);
        }
        for (int c = 0; c < 3; ++c) {
            const double lr = (mb[c] - ma[c]) + std::log(Sa[c] / Sb[c]);
            out[c] = lr >= 0.0 ? 1.0 : std::exp(lr);
        }
        return;
    }
    double r1 = std::sqrt(r * r + d * d + 2.0 * r * mu * d);
    double mu1 = (r * mu + d) / r1;
    double a[3], b[3];
    if (mu > 0.0) {
        T_lookup(p, T, r, mu, a);
        T_lookup(p, T, r1, mu1, b);
    } else {
        T_lookup(p, T, r1, -mu1, a);
        T_lookup(p, T, r, -mu, b);
    }
    /* GLSL min on a GPU is IEEE minNum (the non-NaN operand): where both
     * lookups are 0 (FP16 storage, or T underflowing far below the horizon)
     * the ratio 0/0 gives 1, not NaN (D5) */
    for (int c = 0; c < 3; ++c) out[c] = std::fmin(a[c] / b[c], 1.0);
}

/* common.glsl:138-142 + 284-287 irradiance(sampler, r, muS) */
void E_lookup(const P &p, const double *E, double r, double muS, double out[3]) {
    double uR = (r - p.Rg) / (p.Rt - p.Rg);
    double uMuS = (muS + 0.2) / (1.0 + 0.2);
    sample2(E, p.SW, p.SH, 3, uMuS, uR, out);
}

/* common.glsl:149-173 texture4D (INSCATTER_NON_LINEAR): the 3D-texture
 * coordinates of the two nu slices and the nu lerp weight.
 * out = {uR, uMu, uMuS, nu-lerp position}. D1, D2. */
void tex4D_coords_linear(const P &p, double r, double mu, double muS, double nu, double out[4]);
void tex4D_coords(const P &p, double r, double mu, double muS, double nu, double out[4]) {
    if (p.lin4) {
        tex4D_coords_linear(p, r, mu, muS, nu, out);
        return;
    }
    double H = std::sqrt(p.Rt * p.Rt - p.Rg * p.Rg);
    double rho = std::sqrt(std::max(r * r - p.Rg * p.Rg, 0.0));
    double rmu = r * mu;
    double delta = rmu * rmu - r * r + p.Rg * p.Rg;
    double uMu;
    if (rmu < 0.0 && delta > 0.0) {
        double f = -rho / (std::sqrt(delta) - rmu);
        uMu = (0.5 - 0.5 / p.RMU) + f * (0.5 - 1.0 / p.RMU);
    } else {
        double f = (-rmu + std::sqrt(std::max(delta + H * H, 0.0))) / (rho + H);
        uMu = (0.5 + 0.5 / p.RMU) + f * (0.5 - 1.0 / p.RMU);
    }
    double uR = 0.5 / p.RR + rho / H * (1.0 - 1.0 / p.RR);
    double uMuS = 0.5 / p.RMUS +
                  (std::atan(std::max(muS, -0.1975) * std::tan(1.26 * 1.1)) / 1.1 + (1.0 - 0.26)) *
                      0.5 * (1.0 - 1.0 / p.RMUS);
    double lerp = (nu + 1.0) / 2.0 * (p.RNU - 1.0);
    out[0] = uR;
    out[1] = uMu;
    out[2] = uMuS;
    out[3] = lerp;
}

/* common.glsl:162-166, the #else branch (no INSCATTER_NON_LINEAR) */
void tex4D_coords_linear(const P &p, double r, double mu, double muS, double nu, double out[4]) {
    double H = std::sqrt(p.Rt * p.Rt - p.Rg * p.Rg);
    double rho = std::sqrt(std::max(r * r - p.Rg * p.Rg, 0.0));
    out[0] = 0.5 / p.RR + rho / H * (1.0 - 1.0 / p.RR);
    out[1] = 0.5 / p.RMU + (mu + 1.0) / 2.0 * (1.0 - 1.0 / p.RMU);
    out[2] = 0.5 / p.RMUS + std::max(muS + 0.2, 0.0) / 1.2 * (1.0 - 1.0 / p.RMUS);
    out[3] = (nu + 1.0) / 2.0 * (p.RNU - 1.0);
}

/* The literal GLSL form of the same coordinates (common.glsl:155-160), kept
 * only so the tests can show D2 changes nothing away from rho = 0. */
void tex4D_coords_literal(const P &p, double r, double mu, double muS, double nu,
                          double out[4]) {
    if (p.lin4) {
        tex4D_coords_linear(p, r, mu, muS, nu, out);
        return;
    }
    double H = std::sqrt(p.Rt * p.Rt - p.Rg * p.Rg);
    double rho = std::sqrt(r * r - p.Rg * p.Rg);
    double rmu = r * mu;
    double delta = rmu * rmu - r * r + p.Rg * p.Rg;
    double c0, c1, c2, c3;
    if (rmu < 0.0 && delta > 0.0) {
        c0 = 1.0; c1 = 0.0; c2 = 0.0; c3 = 0.5 - 0.5 / p.RMU;
    } else {
        c0 = -1.0; c1 = H * H; c2 = H; c3 = 0.5 + 0.5 / p.RMU;
    }
    double uR = 0.5 / p.RR + rho / H * (1.0 - 1.0 / p.RR);
    double uMu = c3 + (rmu * c0 + std::sqrt(delta + c1)) / (rho + c2) * (0.5 - 1.0 / p.RMU);
    double uMuS = 0.5 / p.RMUS +
                  (std::atan(std::max(muS, -0.1975) * std::tan(1.26 * 1.1)) / 1.1 + (1.0 - 0.26)) *
                      0.5 * (1.0 - 1.0 / p.RMUS);
    out[0] = uR;
    out[1] = uMu;
    out[2] = uMuS;
    out[3] = (nu + 1.0) / 2.0 * (p.RNU - 1.0);
}

/* common.glsl:149-173 texture4D on a [RES_R][RES_MU][RES_NU*RES_MU_S][C] table */
void tex4D(const P &p, const double *tab, int C, double r, double mu, double muS, double nu,
           double *out) {
    double c[4];
    tex4D_coords(p, r, mu, muS, nu, c);
    double lerp = c[3];
    double uNu = std::floor(lerp);
    lerp = lerp - uNu;
    double a[4], b[4];
    int W = p.RMUS * p.RNU;
    sample3(tab, W, p.RMU, p.RR, C, (uNu + c[2]) / p.RNU, c[1], c[0], a);
    sample3(tab, W, p.RMU, p.RR, C, (uNu + c[2] + 1.0) / p.RNU, c[1], c[0], b);
    for (int k = 0; k < C; ++k) out[k] = a[k] * (1.0 - lerp) + b[k] * lerp;
}

/* common.glsl:175-201 getMuMuSNu (INSCATTER_NON_LINEAR) for texel (x, y) */
void muMuSNu(const P &p, double r, const double *dhdH, int xi, int yi, double &mu, double &muS,
             double &nu) {
    double x = xi;  /* gl_FragCoord.x - 0.5 */
    double y = yi;
    if (p.lin4) {   /* :194-199 */
        mu = -1.0 + 2.0 * y / (p.RMU - 1.0);
        muS = std::fmod(x, (double)p.RMUS) / (p.RMUS - 1.0);
        muS = -0.2 + muS * 1.2;
        nu = -1.0 + std::floor(x / p.RMUS) / (p.RNU - 1.0) * 2.0;
        return;
    }
    if (y < p.RMU / 2.0) {
        double d = 1.0 - y / (p.RMU / 2.0 - 1.0);
        d = std::min(std::max(dhdH[2], d * dhdH[3]), dhdH[3] * 0.999);
        mu = (p.Rg * p.Rg - r * r - d * d) / (2.0 * r * d);
        mu = std::min(mu, -std::sqrt(1.0 - (p.Rg / r) * (p.Rg / r)) - 0.001);
    } else {
        double d = (y - p.RMU / 2.0) / (p.RMU / 2.0 - 1.0);
        d = std::min(std::max(dhdH[0], d * dhdH[1]), dhdH[1] * 0.999);
        mu = (p.Rt * p.Rt - r * r - d * d) / (2.0 * r * d);
    }
    muS = std::fmod(x, (double)p.RMUS) / (p.RMUS - 1.0);
    muS = std::tan((2.0 * muS - 1.0 + 0.26) * 1.1) / std::tan(1.26 * 1.1);
    nu = -1.0 + std::floor(x / p.RMUS) / (p.RNU - 1.0) * 2.0;
}

/* transmittance.glsl:78-91 opticalDepth: 500-step trapezoid of exp(-h/H) to limit() */
double opticalDepth(const P &p, double H, double r, double mu) {
    double result = 0.0;
    double dx = limit(p, r, mu) / p.tS;
    double yi = std::exp(-(r - p.Rg) / H);
    for (int i = 1; i <= p.tS; ++i) {
        double xj = i * dx;
        double yj = std::exp(-(std::sqrt(r * r + xj * xj + 2.0 * xj * r * mu) - p.Rg) / H);
        result += (yi + yj) / 2.0 * dx;
        yi = yj;
    }
    return mu < -std::sqrt(1.0 - (p.Rg / r) * (p.Rg / r)) ? 1e9 : result;
}

void set_threads();

/* D6: the optical depths tau = -ln T of the T texels (the exponent of
 * transmittance.glsl:93-98), [TH][TW][3] */
void tau_table(const P &p, double *tau) {
    set_threads();
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < p.TH; ++j)
        for (int i = 0; i < p.TW; ++i) {
            double r, mu;
            transmittanceRMu(p, i, j, r, mu);
            const double dR = opticalDepth(p, p.HR, r, mu), dM = opticalDepth(p, p.HM, r, mu);
            for (int c = 0; c < 3; ++c)
                tau[((size_t)j * p.TW + i) * 3 + c] = p.bR[c] * dR + p.bME[c] * dM;
        }
}

void set_threads() {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

inline size_t idx4(const P &p, int l, int y, int x) {
    return ((size_t)l * p.RMU + y) * (size_t)(p.RMUS * p.RNU) + x;
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void) {
#ifdef _OPENMP
    set_threads();
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_layer(const atmo_params *a, int l, double *r, double dhdH[4]) {
    P p = mk(a, nullptr);
    layer(p, l, *r, dhdH);
}
double oracle_limit(const atmo_params *a, double r, double mu) { return limit(mk(a, nullptr), r, mu); }
double oracle_phaseR(double mu) { return phaseR(mu); }
double oracle_phaseM(const atmo_params *a, double mu) { return phaseM(mk(a, nullptr), mu); }
void oracle_transmittance_uv(const atmo_params *a, const atmo_options *o, double r, double mu,
                             double uv[2]) {
    transmittanceUV(mk(a, o), r, mu, uv[0], uv[1]);
}
void oracle_transmittance_rmu(const atmo_params *a, const atmo_options *o, int i, int j,
                              double rmu[2]) {
    transmittanceRMu(mk(a, o), i, j, rmu[0], rmu[1]);
}
void oracle_irradiance_rmus(const atmo_params *a, int i, int j, double rmus[2]) {
    irradianceRMuS(mk(a, nullptr), i, j, rmus[0], rmus[1]);
}
void oracle_mumusnu(const atmo_params *a, const atmo_options *o, double r, const double dhdH[4],
                    int x, int y, double out[3]) {
    muMuSNu(mk(a, o), r, dhdH, x, y, out[0], out[1], out[2]);
}
void oracle_texture4D_coords(const atmo_params *a, const atmo_options *o, double r, double mu,
                             double muS, double nu, double out[4]) {
    tex4D_coords(mk(a, o), r, mu, muS, nu, out);
}
void oracle_texture4D_coords_literal(const atmo_params *a, const atmo_options *o, double r,
                                     double mu, double muS, double nu, double out[4]) {
    tex4D_coords_literal(mk(a, o), r, mu, muS, nu, out);
}
void oracle_sample2D(const double *tex, int W, int H, int C, double s, double t, double *out) {
    sample2(tex, W, H, C, s, t, out);
}
void oracle_sample3D(const double *tex, int W, int H, int D, int C, double s, double t, double u,
                     double *out) {
    sample3(tex, W, H, D, C, s, t, u, out);
}
void oracle_texture4D(const atmo_params *a, const atmo_options *o, const double *tab, int C,
                      double r, double mu, double muS, double nu, double *out) {
    tex4D(mk(a, o), tab, C, r, mu, muS, nu, out);
}
void oracle_transmittance_lookup(const atmo_params *a, const double *T, double r, double mu,
                                 double out[3]) {
    T_lookup(mk(a, nullptr), T, r, mu, out);
}
void oracle_transmittance_segment(const atmo_params *a, const double *T, double r, double mu,
                                  double d, double out[3]) {
    T_segment(mk(a, nullptr), T, r, mu, d, out);
}
void oracle_transmittance_segment_tau(const atmo_params *a, const atmo_options *o,
                                      const double *tau, double r, double mu, double d,
                                      double out[3]) {
    P p = mk(a, o);
    p.tau = tau;
    T_segment(p, nullptr, r, mu, d, out);
}
void oracle_optical_depth_table(const atmo_params *a, const atmo_options *o, double *tau) {
    tau_table(mk(a, o), tau);
}
double oracle_optical_depth(const atmo_params *a, const atmo_options *o, double H, double r,
                            double mu) {
    return opticalDepth(mk(a, o), H, r, mu);
}

/* transmittance.glsl:93-98 */
void oracle_pass_transmittance(const atmo_params *a, const atmo_options *o, double *T) {
    P p = mk(a, o);
    set_threads();
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < p.TH; ++j) {
        for (int i = 0; i < p.TW; ++i) {
            double r, mu;
            transmittanceRMu(p, i, j, r, mu);
            double dR = opticalDepth(p, p.HR, r, mu);
            double dM = opticalDepth(p, p.HM, r, mu);
            for (int c = 0; c < 3; ++c)
                T[((size_t)j * p.TW + i) * 3 + c] = std::exp(-(p.bR[c] * dR + p.bME[c] * dM));
        }
    }
}

/* irradiance1.glsl:78-82 */
void oracle_pass_irradiance1(const atmo_params *a, const atmo_options *o, const double *T,
                             double *dE) {
    P p = mk(a, o);
    for (int j = 0; j < p.SH; ++j)
        for (int i = 0; i < p.SW; ++i) {
            double r, muS, t[3];
            irradianceRMuS(p, i, j, r, muS);
            T_lookup(p, T, r, muS, t);
            for (int c = 0; c < 3; ++c) dE[((size_t)j * p.SW + i) * 3 + c] = t[c] * std::max(muS, 0.0);
        }
}

namespace {
/* inscatter1.glsl:104-115 integrand */
void inscatter1_integrand(const P &p, const double *T, double r, double mu, double muS, double nu,
                          double t, double ray[3], double mie[3]) {
    ray[0] = ray[1] = ray[2] = 0.0;
    mie[0] = mie[1] = mie[2] = 0.0;
    double ri = std::sqrt(r * r + t * t + 2.0 * r * mu * t);
    double muSi = (nu * t + muS * r) / ri;
    ri = std::max(p.Rg, ri);
    if (muSi >= -std::sqrt(1.0 - p.Rg * p.Rg / (ri * ri))) {
        double t1[3], t2[3];
        T_segment(p, T, r, mu, t, t1);
        T_lookup(p, T, ri, muSi, t2);
        double er = std::exp(-(ri - p.Rg) / p.HR), em = std::exp(-(ri - p.Rg) / p.HM);
        for (int c = 0; c < 3; ++c) {
            double ti = t1[c] * t2[c];
            ray[c] = er * ti;
            mie[c] = em * ti;
        }
    }
}
}  // namespace

/* inscatter1.glsl:117-150, one draw per layer (PreprocessAtmo.cpp:285-297) */
static void pass_inscatter1(const P &p, const double *T, double *dSR, double *dSM) {
    int W = p.RMUS * p.RNU;
    set_threads();
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int l = 0; l < p.RR; ++l) {
        for (int y = 0; y < p.RMU; ++y) {
            double r, dhdH[4];
            layer(p, l, r, dhdH);
            for (int x = 0; x < W; ++x) {
                double mu, muS, nu;
                muMuSNu(p, r, dhdH, x, y, mu, muS, nu);
                double ray[3] = {0, 0, 0}, mie[3] = {0, 0, 0};
                double dx = limit(p, r, mu) / p.iS;
                double rayi[3], miei[3];
                inscatter1_integrand(p, T, r, mu, muS, nu, 0.0, rayi, miei);
                for (int i = 1; i <= p.iS; ++i) {
                    double xj = i * dx;
                    double rayj[3], miej[3];
                    inscatter1_integrand(p, T, r, mu, muS, nu, xj, rayj, miej);
                    for (int c = 0; c < 3; ++c) {
                        ray[c] += (rayi[c] + rayj[c]) / 2.0 * dx;
                        mie[c] += (miei[c] + miej[c]) / 2.0 * dx;
                        rayi[c] = rayj[c];
                        miei[c] = miej[c];
                    }
                }
                size_t k = idx4(p, l, y, x) * 3;
                for (int c = 0; c < 3; ++c) {
                    dSR[k + c] = ray[c] * p.bR[c];
                    dSM[k + c] = mie[c] * p.bMS[c];
                }
            }
        }
    }
}

void oracle_pass_inscatter1(const atmo_params *a, const atmo_options *o, const double *T,
                            double *dSR, double *dSM) {
    pass_inscatter1(mk(a, o), T, dSR, dSM);
}

/* copyInscatter1.glsl:104-109 (D4) */
void oracle_pass_copy_inscatter1(const atmo_params *a, const double *dSR, const double *dSM,
                                 double *S) {
    P p = mk(a, nullptr);
    size_t n = (size_t)p.RR * p.RMU * p.RMUS * p.RNU;
    for (size_t k = 0; k < n; ++k) {
        S[k * 4 + 0] = dSR[k * 3 + 0];
        S[k * 4 + 1] = dSR[k * 3 + 1];
        S[k * 4 + 2] = dSR[k * 3 + 2];
        S[k * 4 + 3] = dSM[k * 3 + 0];
    }
}

/* inscatterS.glsl:111-189 */
static void pass_inscatterS(const P &p, const double *T, const double *dE, const double *dSR,
                            const double *dSM, int first, double *dJ) {
    int W = p.RMUS * p.RNU;
    const int NS = p.sphS;
    const double dphi = kPI / NS, dtheta = kPI / NS;
    set_threads();
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int l = 0; l < p.RR; ++l) {
        for (int y = 0; y < p.RMU; ++y) {
            double r0, dhdH[4];
            layer(p, l, r0, dhdH);
            for (int x = 0; x < W; ++x) {
                double mu, muS, nu;
                muMuSNu(p, r0, dhdH, x, y, mu, muS, nu);
                double r = clampd(r0, p.Rg, p.Rt);
                mu = clampd(mu, -1.0, 1.0);
                muS = clampd(muS, -1.0, 1.0);
                double var = std::sqrt(1.0 - mu * mu) * std::sqrt(1.0 - muS * muS);
                nu = clampd(nu, muS * mu - var, muS * mu + var);
                double cthetamin = -std::sqrt(1.0 - (p.Rg / r) * (p.Rg / r));
                double v[3] = {std::sqrt(1.0 - mu * mu), 0.0, mu};
                double sx = v[0] == 0.0 ? 0.0 : (nu - muS * mu) / v[0];
                double s[3] = {sx, std::sqrt(std::max(0.0, 1.0 - sx * sx - muS * muS)), muS};
                double raymie[3] = {0, 0, 0};
                for (int it = 0; it < NS; ++it) {
                    double theta = (it + 0.5) * dtheta;
                    double ctheta = std::cos(theta);
                    double greflectance = 0.0, dground = 0.0, gtransp[3] = {0, 0, 0};
                    if (ctheta < cthetamin) {
                        greflectance = p.alpha / kPI;
                        dground = -r * ctheta - std::sqrt(r * r * (ctheta * ctheta - 1.0) + p.Rg * p.Rg);
                        T_segment(p, T, p.Rg, -(r * ctheta + dground) / p.Rg, dground, gtransp);
                    }
                    for (int ip = 0; ip < 2 * NS; ++ip) {
                        double phi = (ip + 0.5) * dphi;
                        double dw = dtheta * dphi * std::sin(theta);
                        double w[3] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta),
                                       ctheta};
                        double nu1 = s[0] * w[0] + s[1] * w[1] + s[2] * w[2];
                        double nu2 = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
                        double pr2 = phaseR(nu2), pm2 = phaseM(p, nu2);
                        double gn[3] = {dground * w[0] / p.Rg, dground * w[1] / p.Rg,
                                        (r + dground * w[2]) / p.Rg};
                        double girr[3];
                        E_lookup(p, dE, p.Rg, gn[0] * s[0] + gn[1] * s[1] + gn[2] * s[2], girr);
                        double rm1[3];
                        for (int c = 0; c < 3; ++c) rm1[c] = greflectance * girr[c] * gtransp[c];
                        if (first) {
                            double pr1 = phaseR(nu1), pm1 = phaseM(p, nu1);
                            double ray1[3], mie1[3];
                            tex4D(p, dSR, 3, r, w[2], muS, nu1, ray1);
                            tex4D(p, dSM, 3, r, w[2], muS, nu1, mie1);
                            for (int c = 0; c < 3; ++c) rm1[c] += ray1[c] * pr1 + mie1[c] * pm1;
                        } else {
                            double ray1[3];
                            tex4D(p, dSR, 3, r, w[2], muS, nu1, ray1);
                            for (int c = 0; c < 3; ++c) rm1[c] += ray1[c];
                        }
                        for (int c = 0; c < 3; ++c)
                            raymie[c] += rm1[c] *
                                         (p.bR[c] * std::exp(-(r - p.Rg) / p.HR) * pr2 +
                                          p.bMS[c] * std::exp(-(r - p.Rg) / p.HM) * pm2) *
                                         dw;
                    }
                }
                size_t k = idx4(p, l, y, x) * 3;
                for (int c = 0; c < 3; ++c) dJ[k + c] = raymie[c];
            }
        }
    }
}

void oracle_pass_inscatterS(const atmo_params *a, const atmo_options *o, const double *T,
                            const double *dE, const double *dSR, const double *dSM, int first,
                            double *dJ) {
    pass_inscatterS(mk(a, o), T, dE, dSR, dSM, first, dJ);
}

/* irradianceN.glsl:85-114 */
void oracle_pass_irradianceN(const atmo_params *a, const atmo_options *o, const double *dSR,
                             const double *dSM, int first, double *dE) {
    P p = mk(a, o);
    const int NI = p.irrS;
    const double dphi = kPI / NI, dtheta = kPI / NI;
    set_threads();
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int j = 0; j < p.SH; ++j) {
        for (int i = 0; i < p.SW; ++i) {
            double r, muS;
            irradianceRMuS(p, i, j, r, muS);
            double s[3] = {std::sqrt(std::max(1.0 - muS * muS, 0.0)), 0.0, muS};
            double res[3] = {0, 0, 0};
            for (int ip = 0; ip < 2 * NI; ++ip) {
                double phi = (ip + 0.5) * dphi;
                for (int it = 0; it < NI / 2; ++it) {
                    double theta = (it + 0.5) * dtheta;
                    double dw = dtheta * dphi * std::sin(theta);
                    double w[3] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta),
                                   std::cos(theta)};
                    double nu = s[0] * w[0] + s[1] * w[1] + s[2] * w[2];
                    if (first) {
                        double pr1 = phaseR(nu), pm1 = phaseM(p, nu);
                        double ray1[3], mie1[3];
                        tex4D(p, dSR, 3, r, w[2], muS, nu, ray1);
                        tex4D(p, dSM, 3, r, w[2], muS, nu, mie1);
                        for (int c = 0; c < 3; ++c) res[c] += (ray1[c] * pr1 + mie1[c] * pm1) * w[2] * dw;
                    } else {
                        double ray1[3];
                        tex4D(p, dSR, 3, r, w[2], muS, nu, ray1);
                        for (int c = 0; c < 3; ++c) res[c] += ray1[c] * w[2] * dw;
                    }
                }
            }
            for (int c = 0; c < 3; ++c) dE[((size_t)j * p.SW + i) * 3 + c] = res[c];
        }
    }
}

/* inscatterN.glsl:105-131 */
static void pass_inscatterN(const P &p, const double *T, const double *dJ, double *dSR) {
    int W = p.RMUS * p.RNU;
    set_threads();
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int l = 0; l < p.RR; ++l) {
        for (int y = 0; y < p.RMU; ++y) {
            double r, dhdH[4];
            layer(p, l, r, dhdH);
            for (int x = 0; x < W; ++x) {
                double mu, muS, nu;
                muMuSNu(p, r, dhdH, x, y, mu, muS, nu);
                auto integrand = [&](double t, double out[3]) {
                    double ri = std::sqrt(r * r + t * t + 2.0 * r * mu * t);
                    double mui = (r * mu + t) / ri;
                    double muSi = (nu * t + muS * r) / ri;
                    double j4[3], tt[3];
                    tex4D(p, dJ, 3, ri, mui, muSi, nu, j4);
                    T_segment(p, T, r, mu, t, tt);
                    for (int c = 0; c < 3; ++c) out[c] = j4[c] * tt[c];
                };
                double raymie[3] = {0, 0, 0};
                double dx = limit(p, r, mu) / p.iS;
                double ri_[3];
                integrand(0.0, ri_);
                for (int i = 1; i <= p.iS; ++i) {
                    double xj = i * dx;
                    double rj[3];
                    integrand(xj, rj);
                    for (int c = 0; c < 3; ++c) {
                        raymie[c] += (ri_[c] + rj[c]) / 2.0 * dx;
                        ri_[c] = rj[c];
                    }
                }
                size_t k = idx4(p, l, y, x) * 3;
                for (int c = 0; c < 3; ++c) dSR[k + c] = raymie[c];
            }
        }
    }
}

void oracle_pass_inscatterN(const atmo_params *a, const atmo_options *o, const double *T,
                            const double *dJ, double *dSR) {
    pass_inscatterN(mk(a, o), T, dJ, dSR);
}

/* copyInscatterN.glsl:105-110 with the ADD blend of PreprocessAtmo.cpp:336,350 (D4) */
void oracle_pass_copy_inscatterN(const atmo_params *a, const double *dS, double *S) {
    P p = mk(a, nullptr);
    int W = p.RMUS * p.RNU;
    for (int l = 0; l < p.RR; ++l) {
        double r, dhdH[4];
        layer(p, l, r, dhdH);
        for (int y = 0; y < p.RMU; ++y)
            for (int x = 0; x < W; ++x) {
                double mu, muS, nu;
                muMuSNu(p, r, dhdH, x, y, mu, muS, nu);
                double pr = phaseR(nu);
                size_t k = idx4(p, l, y, x);
                for (int c = 0; c < 3; ++c) S[k * 4 + c] += dS[k * 3 + c] / pr;
                S[k * 4 + 3] += 0.0;
            }
    }
}

/* copyIrradiance.glsl:81-84: E = k * dE (k = 0, PreprocessAtmo.cpp:302) or,
 * with the ADD blend, E += dE (k = 1, :336-341). */
void oracle_pass_copy_irradiance(const atmo_params *a, double k, const double *dE, double *E) {
    P p = mk(a, nullptr);
    size_t n = (size_t)p.SW * p.SH * 3;
    if (k == 0.0) {
        for (size_t i = 0; i < n; ++i) E[i] = 0.0 * dE[i];
    } else {
        for (size_t i = 0; i < n; ++i) E[i] += k * dE[i];
    }
}

/* Nearest IEEE binary16 value (round to nearest even, gradual underflow below
 * 2^-14 on the quantum 2^-24, overflow to +-inf from 65520): what storing a
 * value into an RGB16F / RGBA16F texture keeps (GL 3.3 core 2.1.2). */
double oracle_q16(double x) {
    const double ax = std::fabs(x);
    if (std::isnan(x)) return x;
    if (!(ax < 65520.0)) return std::copysign(INFINITY, x);
    if (ax < 0x1p-14) return std::nearbyint(x * 0x1p24) * 0x1p-24;
    int e;
    std::frexp(ax, &e);                      // ax in [2^(e-1), 2^e)
    const double q = std::ldexp(1.0, e - 11);  // 10 fraction bits
    return std::nearbyint(x / q) * q;
}

static void q16_all(double *v, size_t n) {
    for (size_t i = 0; i < n; ++i) v[i] = oracle_q16(v[i]);
}

/* PreprocessAtmo.cpp:273-354: T, dE1, dS1, E=0, S=dS1, then for each order
 * n = 2..orders: dJn, dEn, dSn, E += dEn, S += dSn/phaseR.  With
 * ATMO_FLAG_HALF_STORAGE every table is rounded to FP16 after the pass that
 * writes it (the reference's RGB16F / RGBA16F textures, PreprocessAtmo.cpp:
 * 130-143) and the two ADD blends round FP16(dst + src) (:336, :350). */
void oracle_precompute(const atmo_params *a, const atmo_options *o, double *T, double *E,
                       double *S) {
    P p = mk(a, o);
    size_t n4 = (size_t)p.RR * p.RMU * p.RMUS * p.RNU;
    size_t nE = (size_t)p.SW * p.SH;
    const bool h = o && (o->flags & ATMO_FLAG_HALF_STORAGE);
    const size_t nT = (size_t)p.TW * p.TH * 3;
    std::vector<double> dE(nE * 3), dSR(n4 * 3), dSM(n4 * 3), dJ(n4 * 3);
    oracle_pass_transmittance(a, o, T);
    if (h) q16_all(T, nT);
    /* D6: optical depths of the T texels, for the segment ratios of the chain
     * (FP16 storage keeps the reference's quotient of stored values) */
    std::vector<double> tau;
    if (!h) {
        tau.resize(nT);
        tau_table(p, tau.data());
    }
    P pc = p;
    pc.tau = h ? nullptr : tau.data();
    oracle_pass_irradiance1(a, o, T, dE.data());
    if (h) q16_all(dE.data(), dE.size());
    pass_inscatter1(pc, T, dSR.data(), dSM.data());
    if (h) {
        q16_all(dSR.data(), dSR.size());
        q16_all(dSM.data(), dSM.size());
    }
    oracle_pass_copy_irradiance(a, 0.0, dE.data(), E);
    oracle_pass_copy_inscatter1(a, dSR.data(), dSM.data(), S);
    for (int order = 2; order <= p.orders; ++order) {
        int first = order == 2;
        pass_inscatterS(pc, T, dE.data(), dSR.data(), dSM.data(), first, dJ.data());
        if (h) q16_all(dJ.data(), dJ.size());
        oracle_pass_irradianceN(a, o, dSR.data(), dSM.data(), first, dE.data());
        if (h) q16_all(dE.data(), dE.size());
        pass_inscatterN(pc, T, dJ.data(), dSR.data());
        if (h) q16_all(dSR.data(), dSR.size());
        oracle_pass_copy_irradiance(a, 1.0, dE.data(), E);
        if (h) q16_all(E, nE * 3);
        oracle_pass_copy_inscatterN(a, dSR.data(), S);
        if (h) q16_all(S, n4 * 4);
    }
}


/* ------------------------------------------------------------------------
 * Table consumer: the public functions of src/atmo/examples/atmo/
 * atmosphereShader.glsl (built with OPTIMIZE, ATMO_FULL, HORIZON_HACK,
 * TRANSMITTANCE_NON_LINEAR and INSCATTER_NON_LINEAR as declared at :62-64,
 * :127-128) reading the tables this preprocess writes.  The shader hard-codes
 * Earth in metres (:72-125); here every constant comes from atmo_params, and
 * lengths are in the parameters' unit (km for Earth), so the shader's 600 m
 * near-ground offset (:472) is 600 / meters_per_unit.
 * Deviation C1: with OPTIMIZE, SQRT(f, err) is sqrt(f) (:324-330), NaN for a
 * ray that misses the top sphere from outside; GLSL max(NaN, 0) is undefined
 * and GPUs return 0, so din = 0 there.
 * ------------------------------------------------------------------------ */
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static const double kSunIntensity = 100.0; /* atmosphereShader.glsl:66 */

/* :300-302 getMie */
static void get_mie(const P &p, const double rm[4], double out[3]) {
    for (int c = 0; c < 3; ++c) out[c] = rm[c] * rm[3] / std::max(rm[0], 1e-4) * (p.bR[0] / p.bR[c]);
}

static void transmittance_ratio(const P &p, const double *T, double r0, double mu0, double r,
                                double mu, double out[3]) {
    double a[3], b[3];
    if (mu > 0.0) {
        T_lookup(p, T, r0, mu0, a);
        T_lookup(p, T, r, mu, b);
    } else {
        T_lookup(p, T, r, -mu, a);
        T_lookup(p, T, r0, -mu0, b);
    }
    for (int c = 0; c < 3; ++c) out[c] = std::fmin(a[c] / b[c], 1.0);   /* D5 */
}

/* :371-418 skyRadiance */
void oracle_sky_radiance(const atmo_params *a, const double *T, const double *S,
                         const double camera[3], const double viewdir[3],
                         const double sundir[3], double shaft_width, double rad[3],
                         double ext[3]) {
    P p = mk(a, nullptr);
    double cam[3];
    for (int c = 0; c < 3; ++c) cam[c] = camera[c] + viewdir[c] * std::max(shaft_width, 0.0);
    double r = std::sqrt(dot3(cam, cam));
    double rMu = dot3(cam, viewdir);
    double mu = rMu / r;
    const double r0 = r, mu0 = mu;
    const double q = rMu * rMu - r * r + p.Rt * p.Rt;
    const double din = q >= 0.0 ? std::max(-rMu - std::sqrt(q), 0.0) : 0.0; /* C1 */
    if (din > 0.0) {
        for (int c = 0; c < 3; ++c) cam[c] += din * viewdir[c];
        rMu += din;
        mu = rMu / p.Rt;
        r = p.Rt;
    }
    if (r <= p.Rt) {
        const double nu = dot3(viewdir, sundir);
        const double muS = dot3(cam, sundir) / r;
        double is[4];
        tex4D(p, S, 4, r, rMu / r, muS, nu, is);
        if (shaft_width > 0.0) {
            double f[3];
            transmittance_ratio(p, T, r0, mu0, r, mu, f);
            for (int c = 0; c < 3; ++c) is[c] *= f[c];
            is[3] *= f[0];
        }
        T_lookup(p, T, r, mu, ext);
        double m[3];
        get_mie(p, is, m);
        const double pr = phaseR(nu), pm = phaseM(p, nu);
        for (int c = 0; c < 3; ++c) rad[c] = (is[c] * pr + m[c] * pm) * kSunIntensity;
    } else {
        for (int c = 0; c < 3; ++c) {
            rad[c] = 0.0;
            ext[c] = 1.0;
        }
    }
}

/* :425-533 inScattering (aerial perspective between camera and point) */
void oracle_in_scattering(const atmo_params *a, const double *T, const double *S,
                          const double camera[3], const double point_in[3],
                          const double sundir[3], double shaft_width, double meters_per_unit,
                          double rad[3], double ext[3]) {
    P p = mk(a, nullptr);
    double cam[3] = {camera[0], camera[1], camera[2]};
    double point[3] = {point_in[0], point_in[1], point_in[2]};
    double vd[3] = {point[0] - cam[0], point[1] - cam[1], point[2] - cam[2]};
    double d = std::sqrt(dot3(vd, vd));
    for (int c = 0; c < 3; ++c) vd[c] /= d;
    double r = std::sqrt(dot3(cam, cam));
    if (r < 0.9 * p.Rg) { /* flat-terrain coordinates */
        cam[2] += p.Rg;
        point[2] += p.Rg;
        r = std::sqrt(dot3(cam, cam));
    }
    double rMu = dot3(cam, vd);
    double mu = rMu / r;
    const double sw = clampd(shaft_width, 0.0, d);
    for (int c = 0; c < 3; ++c) point[c] -= vd[c] * sw;
    const double q = rMu * rMu - r * r + p.Rt * p.Rt;
    const double din = q >= 0.0 ? std::max(-rMu - std::sqrt(q), 0.0) : 0.0; /* C1 */
    if (din > 0.0 && din < d) {
        for (int c = 0; c < 3; ++c) cam[c] += din * vd[c];
        rMu += din;
        mu = rMu / p.Rt;
        r = p.Rt;
        d -= din;
    }
    if (r <= p.Rt) {
        const double nu = dot3(vd, sundir);
        const double muS = dot3(cam, sundir) / r;
        const double near = p.Rg + 600.0 / meters_per_unit;
        if (r < near) {
            const double f = near / r;
            r *= f;
            rMu *= f;
            for (int c = 0; c < 3; ++c) point[c] *= f;
        }
        double r1 = std::sqrt(dot3(point, point));
        const double rMu1 = dot3(point, vd);
        double mu1 = rMu1 / r1;
        const double muS1 = dot3(point, sundir) / r1;
        if (mu > 0.0) {
            double x[3], y[3];
            T_lookup(p, T, r, mu, x);
            T_lookup(p, T, r1, mu1, y);
            for (int c = 0; c < 3; ++c) ext[c] = std::fmin(x[c] / y[c], 1.0);   /* D5 */
        } else {
            double x[3], y[3];
            T_lookup(p, T, r1, -mu1, x);
            T_lookup(p, T, r, -mu, y);
            for (int c = 0; c < 3; ++c) ext[c] = std::fmin(x[c] / y[c], 1.0);   /* D5 */
        }
        auto diff = [&](double m, double rr1, double mm1, double out[4]) {
            double s0[4], s1[4];
            tex4D(p, S, 4, r, m, muS, nu, s0);
            tex4D(p, S, 4, rr1, mm1, muS1, nu, s1);
            const double e[4] = {ext[0], ext[1], ext[2], ext[0]};
            for (int c = 0; c < 4; ++c) out[c] = std::max(s0[c] - s1[c] * e[c], 0.0);
        };
        double is[4];
        const double EPS = 0.004; /* HORIZON_HACK, :486-510 */
        const double lim = -std::sqrt(1.0 - (p.Rg / r) * (p.Rg / r));
        if (std::fabs(mu - lim) < EPS) {
            const double w = ((mu - lim) + EPS) / (2.0 * EPS);
            double A[4], B[4];
            double m = lim - EPS;
            r1 = std::sqrt(r * r + d * d + 2.0 * r * d * m);
            mu1 = (r * m + d) / r1;
            diff(m, r1, mu1, A);
            m = lim + EPS;
            r1 = std::sqrt(r * r + d * d + 2.0 * r * d * m);
            mu1 = (r * m + d) / r1;
            diff(m, r1, mu1, B);
            for (int c = 0; c < 4; ++c) is[c] = A[c] * (1.0 - w) + B[c] * w;
        } else {
            diff(mu, r1, mu1, is);
        }
        /* :519 smoothstep(0.00, 0.02, muS) */
        const double t = clampd(muS / 0.02, 0.0, 1.0);
        is[3] *= t * t * (3.0 - 2.0 * t);
        double m[3];
        get_mie(p, is, m);
        const double pr = phaseR(nu), pm = phaseM(p, nu);
        for (int c = 0; c < 3; ++c) rad[c] = (is[c] * pr + m[c] * pm) * kSunIntensity;
    } else {
        for (int c = 0; c < 3; ++c) {
            rad[c] = 0.0;
            ext[c] = 1.0;
        }
    }
}

/* :342-350 sunRadiance (transmittanceWithShadow, :263-265) and :355-361
 * skyIrradiance, for (r, muS) pairs */
void oracle_sun_sky_irradiance(const atmo_params *a, const double *T, const double *E, double r,
                               double muS, double sun[3], double sky[3]) {
    P p = mk(a, nullptr);
    if (muS < -std::sqrt(1.0 - (p.Rg / r) * (p.Rg / r))) {
        sun[0] = sun[1] = sun[2] = 0.0;
    } else {
        T_lookup(p, T, r, muS, sun);
        for (int c = 0; c < 3; ++c) sun[c] *= kSunIntensity;
    }
    E_lookup(p, E, r, muS, sky);
    for (int c = 0; c < 3; ++c) sky[c] *= kSunIntensity;
}

}  // extern "C"
