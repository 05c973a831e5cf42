/*
 * atmo_oracle.h — FP64 CPU oracle of Proland's atmo preprocess.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP product library's
 * tables with its results; it is never linked into, loaded by or called from
 * the product path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it.
 *
 * It is a double-precision restatement of the reference GLSL passes in
 * src/atmo/sources/proland/preprocess/atmo/ together with an emulation
 * of the OpenGL texture sampling those shaders rely on (GL 3.3 core spec
 * section 3.8.11 LINEAR filtering, CLAMP_TO_EDGE wrap).  Tables are stored
 * in FP64; with ATMO_FLAG_HALF_STORAGE, oracle_precompute rounds every stored
 * table to FP16 as the reference's RGB16F/RGBA16F textures do.
 *
 * The coordinate primitives take the options only for their parameterisation
 * flags (NULL: the reference's non-linear parameterisations).
 *
 * Table layouts (doubles, rows as in the reference's .raw files):
 *   T   [TRANSMITTANCE_H][TRANSMITTANCE_W][3]
 *   E   [SKY_H][SKY_W][3]          (also dE)
 *   dS  [RES_R][RES_MU][RES_NU*RES_MU_S][3]   (dSR, dSM, dJ)
 *   S   [RES_R][RES_MU][RES_NU*RES_MU_S][4]
 */
#ifndef ATMO_ORACLE_H_
#define ATMO_ORACLE_H_

#include "../include/atmo.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_API __attribute__((visibility("default")))

/* --- primitives (each cites the reference function it restates) --- */
ORACLE_API void oracle_layer(const atmo_params *p, int layer, double *r, double dhdH[4]);
ORACLE_API double oracle_limit(const atmo_params *p, double r, double mu);
ORACLE_API double oracle_phaseR(double mu);
ORACLE_API double oracle_phaseM(const atmo_params *p, double mu);
ORACLE_API void oracle_transmittance_uv(const atmo_params *p, const atmo_options *o, double r,
                                        double mu, double uv[2]);
ORACLE_API void oracle_transmittance_rmu(const atmo_params *p, const atmo_options *o, int i, int j,
                                         double rmu[2]);
ORACLE_API void oracle_irradiance_rmus(const atmo_params *p, int i, int j, double rmus[2]);
ORACLE_API void oracle_mumusnu(const atmo_params *p, const atmo_options *o, double r,
                               const double dhdH[4], int x, int y, double out[3]);
ORACLE_API void oracle_texture4D_coords(const atmo_params *p, const atmo_options *o, double r,
                                        double mu, double muS, double nu, double out[4]);
ORACLE_API void oracle_texture4D_coords_literal(const atmo_params *p, const atmo_options *o,
                                                double r, double mu, double muS, double nu,
                                                double out[4]);
ORACLE_API void oracle_sample2D(const double *tex, int W, int H, int C, double s, double t,
                                double *out);
ORACLE_API void oracle_sample3D(const double *tex, int W, int H, int D, int C, double s, double t,
                                double u, double *out);
ORACLE_API void oracle_texture4D(const atmo_params *p, const atmo_options *o, const double *tab,
                                 int C, double r, double mu, double muS, double nu, double *out);
ORACLE_API void oracle_transmittance_lookup(const atmo_params *p, const double *T, double r,
                                            double mu, double out[3]);
ORACLE_API void oracle_transmittance_segment(const atmo_params *p, const double *T, double r,
                                             double mu, double d, double out[3]);
/* the same segment ratio from a table of optical depths -ln T, [TH][TW][3]
 * (deviation D6: how oracle_precompute evaluates it) */
ORACLE_API void oracle_transmittance_segment_tau(const atmo_params *p, const atmo_options *o,
                                                 const double *tau, double r, double mu,
                                                 double d, double out[3]);
ORACLE_API void oracle_optical_depth_table(const atmo_params *p, const atmo_options *o,
                                           double *tau);
ORACLE_API double oracle_optical_depth(const atmo_params *p, const atmo_options *o, double H,
                                       double r, double mu);

/* --- passes (one reference shader program each) --- */
ORACLE_API void oracle_pass_transmittance(const atmo_params *p, const atmo_options *o, double *T);
ORACLE_API void oracle_pass_irradiance1(const atmo_params *p, const atmo_options *o,
                                        const double *T, double *dE);
ORACLE_API void oracle_pass_inscatter1(const atmo_params *p, const atmo_options *o,
                                       const double *T, double *dSR, double *dSM);
ORACLE_API void oracle_pass_copy_inscatter1(const atmo_params *p, const double *dSR,
                                            const double *dSM, double *S);
ORACLE_API void oracle_pass_inscatterS(const atmo_params *p, const atmo_options *o,
                                       const double *T, const double *dE, const double *dSR,
                                       const double *dSM, int first, double *dJ);
ORACLE_API void oracle_pass_irradianceN(const atmo_params *p, const atmo_options *o,
                                        const double *dSR, const double *dSM, int first,
                                        double *dE);
ORACLE_API void oracle_pass_inscatterN(const atmo_params *p, const atmo_options *o,
                                       const double *T, const double *dJ, double *dSR);
ORACLE_API void oracle_pass_copy_inscatterN(const atmo_params *p, const double *dS, double *S);
ORACLE_API void oracle_pass_copy_irradiance(const atmo_params *p, double k, const double *dE,
                                            double *E);

/* Nearest FP16 value of x (round to nearest even, gradual underflow). */
ORACLE_API double oracle_q16(double x);

/* Whole precompute (PreprocessAtmo.cpp:273-354 pass order).  T, E, S as above. */
ORACLE_API void oracle_precompute(const atmo_params *p, const atmo_options *o, double *T,
                                  double *E, double *S);

/* --- table consumer (src/atmo/examples/atmo/atmosphereShader.glsl) ---
 * T [TH][TW][3], S [RES_R][RES_MU][RES_NU*RES_MU_S][4], E [SH][SW][3];
 * positions in the parameters' length unit, planet centre at the origin. */
ORACLE_API void oracle_sky_radiance(const atmo_params *p, const double *T, const double *S,
                                    const double camera[3], const double viewdir[3],
                                    const double sundir[3], double shaft_width, double rad[3],
                                    double ext[3]);
ORACLE_API void oracle_in_scattering(const atmo_params *p, const double *T, const double *S,
                                     const double camera[3], const double point[3],
                                     const double sundir[3], double shaft_width,
                                     double meters_per_unit, double rad[3], double ext[3]);
ORACLE_API void oracle_sun_sky_irradiance(const atmo_params *p, const double *T, const double *E,
                                          double r, double muS, double sun[3], double sky[3]);

/* Number of OpenMP threads the passes use (0 = library default). */
ORACLE_API void oracle_set_threads(int n);
ORACLE_API int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
