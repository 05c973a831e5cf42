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

/* ---------------------------------------------------------------------- */
/* ork texture descriptors                                                 */
/* ---------------------------------------------------------------------- */

/* One texture2D / texture3D element of an ork resource file, the way the
 * renderers load the tables: src/demo/textures/atmo/{transmittance,
 * irradiance,inscatter}.xml:25 and src/atmo/examples/atmo/helloworld.xml:71-75
 * (source="<table>.raw" internalformat="RGBA16F" min="LINEAR" mag="LINEAR"
 * wraps/wrapt/wrapr="CLAMP_TO_EDGE"), with the sizes from the .raw trailer of
 * its source (PreprocessAtmo.cpp:355-405). */
typedef struct atmo_texture_desc {
    char name[64];
    char source[512];        /* the .raw path, resolved against the descriptor's directory */
    char format[16];         /* internalformat */
    int kind;                /* 2: texture2D, 3: texture3D */
    int format_channels;     /* channels of the internal format */
    int half;                /* 1: a 16F format (values stored as IEEE binary16) */
    int min_linear, mag_linear;
    int clamp_to_edge[3];    /* wraps, wrapt, wrapr == CLAMP_TO_EDGE */
    int width, height, depth, components;   /* trailer: depth 0 for a 2D table */
} atmo_texture_desc;

/* Parses the first texture2D / texture3D element of the descriptor file and
 * validates it against its source's trailer: magic and size, texture3D <=>
 * depth > 0, a known float internalformat with at least as many channels as
 * the payload, an explicit wrap mode on every axis. */
ATMO_API atmo_status atmo_texture_load_desc(const char *xml_path, atmo_texture_desc *out);

/* The payload of the descriptor's source (.raw layout: components floats per
 * texel, x fastest) as the GPU texture holds it: with a 16F internalformat
 * every value rounded to the nearest binary16 (round to nearest even, gradual
 * underflow, +-inf from 65520), a 32F one unchanged.  data: host buffer of
 * capacity_floats; out (may be NULL) receives the descriptor. */
ATMO_API atmo_status atmo_texture_load(const char *xml_path, float *data, size_t capacity_floats,
                                       atmo_texture_desc *out);

#ifdef __cplusplus
}
#endif

#endif
