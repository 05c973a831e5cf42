/*
 * atmo_consumer.h — C ABI of the table consumer: the radiance functions that
 * Proland's rendering shaders evaluate on the tables this library writes.
 *
 * Reference: src/atmo/examples/atmo/atmosphereShader.glsl (built with
 * OPTIMIZE, ATMO_FULL, HORIZON_HACK, TRANSMITTANCE_NON_LINEAR and
 * INSCATTER_NON_LINEAR, :62-64, :127-128), the include every atmosphere,
 * terrain and ocean shader of the demos uses.  The shader hard-codes Earth in
 * metres and the reference table sizes (:72-125); here every constant comes
 * from atmo_params, so non-Earth and non-default-resolution tables render
 * too.  Positions are in the parameters' length unit (km for Earth), with the
 * planet centre at the origin.
 *
 * Tables are read in their .raw layouts, exactly as written by
 * atmo_precompute / atmo_preprocess (device pointers, float32):
 *   T  [TRANSMITTANCE_H][TRANSMITTANCE_W][3]
 *   S  [RES_R][RES_MU][RES_NU*RES_MU_S][4]
 *   E  [SKY_H][SKY_W][3]
 * with GL LINEAR filtering and CLAMP_TO_EDGE (the sampler state of
 * src/atmo/examples/atmo/helloworld.xml:71-75).  Results are multiplied by
 * the shader's SUN_INTENSITY (100, :66).
 */
#ifndef ATMO_CONSUMER_H_
#define ATMO_CONSUMER_H_

#include "atmo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct atmo_view {
    double camera[3];        /* observer position */
    double sun[3];           /* unit vector towards the sun */
    double shaft_width;      /* the shaders' shaftWidth argument (0: no light shafts) */
    double meters_per_unit;  /* 1000 for km: sets inScattering's 600 m near-ground offset */
} atmo_view;

/* skyRadiance (atmosphereShader.glsl:371-418) along n view directions
 * dirs[n][4] (unit xyz, w ignored) -> radiance[n][4], extinction[n][4]
 * (rgb, w = 0).  extinction may be NULL. */
ATMO_API atmo_status atmo_sky_radiance(const atmo_params *p, const float *T, const float *S,
                                       const atmo_view *view, const float *dirs, int n,
                                       float *radiance, float *extinction, void *stream);

/* inScattering (atmosphereShader.glsl:425-533): light scattered between the
 * camera and n points points[n][4] (xyz, w ignored) -> radiance[n][4],
 * extinction[n][4] (the camera-to-point transmittance).  extinction may be
 * NULL. */
ATMO_API atmo_status atmo_in_scattering(const atmo_params *p, const float *T, const float *S,
                                        const atmo_view *view, const float *points, int n,
                                        float *radiance, float *extinction, void *stream);

/* sunRadiance (:342-350) and skyIrradiance (:355-361) at n pairs
 * rmus[n][2] = (r, muS) -> sun[n][4], sky[n][4]. */
ATMO_API atmo_status atmo_sun_sky_irradiance(const atmo_params *p, const float *T, const float *E,
                                             const float *rmus, int n, float *sun, float *sky,
                                             void *stream);

#ifdef __cplusplus
}
#endif

#endif
