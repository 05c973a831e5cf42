/*
 * atmo.h — C ABI of the MI355X atmospheric-scattering precompute.
 *
 * Drop-in boundary for Proland's atmo preprocess
 * (src/atmo/sources/proland/preprocess/atmo/PreprocessAtmo.{h,cpp}).
 * Every entry point takes plain pointers and sizes; device pointers are
 * HIP device addresses on the engine's device, `stream` is a hipStream_t
 * passed as void* (NULL = the null stream).
 *
 * Table layouts (identical to the reference's .raw files, little-endian FP32):
 *   transmittance  float[TRANSMITTANCE_H][TRANSMITTANCE_W][3]      (RGB)
 *   irradiance     float[SKY_H][SKY_W][3]                          (RGB)
 *   inscatter      float[RES_R][RES_MU][RES_NU*RES_MU_S][4]        (RGBA, x = nu*RES_MU_S + mu_s)
 * (PreprocessAtmo.cpp:355-405 writes exactly these, followed by a 20-byte
 *  trailer {0xCAFEBABE, width, height*depth, depth, components}.)
 *
 * Pass-level entry points (atmo_pass_*) work on the engine's internal
 * "texture" layout: one float4 per texel (RGB tables leave .w unused), rows
 * as above.  They mirror one reference shader program each and exist so a
 * caller (or a test) can drive a single pass on its own inputs.
 */
#ifndef ATMO_H_
#define ATMO_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATMO_API __attribute__((visibility("default")))

/* Status codes.  The reference reports no errors at all
 * (PreprocessAtmo.cpp:367 uses fopen/fwrite unchecked); every entry point
 * here returns one of these and sets atmo_last_error(). */
typedef enum atmo_status {
    ATMO_OK = 0,
    ATMO_ERR_INVALID_ARGUMENT = 1,
    ATMO_ERR_DEVICE = 2,        /* HIP runtime error, or no usable gfx950 device */
    ATMO_ERR_OUT_OF_MEMORY = 3,
    ATMO_ERR_IO = 4,            /* .raw file could not be written/read/validated */
    ATMO_ERR_NUMERIC = 5,       /* NaN/Inf or out-of-range table values */
    ATMO_ERR_UNSUPPORTED = 6
} atmo_status;

typedef struct atmo_vec3f { float x, y, z; } atmo_vec3f;

/* Byte-compatible with proland::AtmoParameters (PreprocessAtmo.h:44-67):
 * 24 x 4 bytes = 96 bytes, members in the same order. */
typedef struct atmo_params {
    float Rg;
    float Rt;
    float RL;
    int TRANSMITTANCE_W;
    int TRANSMITTANCE_H;
    int SKY_W;
    int SKY_H;
    int RES_R;
    int RES_MU;
    int RES_MU_S;
    int RES_NU;
    float AVERAGE_GROUND_REFLECTANCE;
    float HR;
    atmo_vec3f betaR;
    float HM;
    atmo_vec3f betaMSca;
    atmo_vec3f betaMEx;
    float mieG;
} atmo_params;

/* Extensions the reference hard-codes (common.glsl:94-97 sample counts,
 * PreprocessAtmo.cpp:351 order count).  Defaults reproduce the reference. */
typedef struct atmo_options {
    int orders;                 /* highest scattering order, >= 1 (reference: 4) */
    int transmittance_samples;  /* TRANSMITTANCE_INTEGRAL_SAMPLES (500) */
    int inscatter_samples;      /* INSCATTER_INTEGRAL_SAMPLES (50) */
    int irradiance_samples;     /* IRRADIANCE_INTEGRAL_SAMPLES (32) */
    int spherical_samples;      /* INSCATTER_SPHERICAL_INTEGRAL_SAMPLES (16) */
    int flags;                  /* ATMO_FLAG_* */
    int texels_per_thread;      /* k_inscatterS variant 2, 3, 4 or 8 (bit-identical results);
                                   0 = per set by workgroup occupancy for the pass-level
                                   entry points, autotuned in atmo_engine_precompute */
    int inscatterN_form;        /* k_inscatterN form: 1 direct gathers, 2 LDS-staged rows
                                   (2 samples per LDS round), 3 LDS-staged rows (1 sample
                                   per round), all bit-identical; 0 = form 3 for the
                                   pass-level entry points, autotuned in
                                   atmo_engine_precompute */
    int reserved[8];
} atmo_options;

#define ATMO_FLAG_CHECK_NUMERICS 1  /* validate outputs (finite, T in [0,1], E,S >= 0) */
#define ATMO_FLAG_NO_GRAPH       2  /* launch passes eagerly instead of replaying a hipGraph */
#define ATMO_FLAG_TIME_PASSES    4  /* eager launches, a hipEvent pair around each pass
                                       (read with atmo_engine_pass_times) */
#define ATMO_FLAG_NO_AUTOTUNE    8  /* keep the default kernel variants */
#define ATMO_FLAG_HALF_STORAGE  16  /* store every table value rounded to the nearest FP16
                                       (round to nearest even, gradual underflow), and every
                                       ADD blend as FP16(dst + src), as the reference's
                                       RGB16F / RGBA16F textures and blending do
                                       (PreprocessAtmo.cpp:130-143, :336, :350); the
                                       default keeps FP32 tables */
#define ATMO_FLAG_LINEAR_TRANSMITTANCE 32  /* the linear transmittance parameterisation: the
                                              reference built without TRANSMITTANCE_NON_LINEAR
                                              (common.glsl:105, :113-136) */
#define ATMO_FLAG_LINEAR_INSCATTER     64  /* the linear inscatter parameterisation: the
                                              reference built without INSCATTER_NON_LINEAR
                                              (common.glsl:106, :149-201); the table consumer
                                              (atmo_consumer.h) reads the non-linear tables */

/* Reference defaults (PreprocessAtmo.cpp:90-98; note betaMEx = 4.44e-3). */
ATMO_API void atmo_params_default(atmo_params *p);
ATMO_API void atmo_options_default(atmo_options *o);

/* Last error message of the calling thread ("" when none). */
ATMO_API const char *atmo_last_error(void);

/* Library version string. */
ATMO_API const char *atmo_version(void);

/* ---------------------------------------------------------------------- */
/* Whole-precompute entry points                                           */
/* ---------------------------------------------------------------------- */

/* Replaces proland::preprocessAtmo(const AtmoParameters&, const char*)
 * (PreprocessAtmo.h:75, PreprocessAtmo.cpp:424-436): if <outdir>/inscatter.raw
 * exists (and its trailer validates) return ATMO_OK at once; otherwise run the
 * 4-order precompute on the default device and write transmittance.raw,
 * irradiance.raw and inscatter.raw (each via a temporary file + rename).
 * Returns instead of calling exit(0). */
ATMO_API atmo_status atmo_preprocess(const atmo_params *params, const char *outdir);

/* Same, with explicit options; skip_if_exists=0 always recomputes. */
ATMO_API atmo_status atmo_preprocess_ex(const atmo_params *params, const atmo_options *opts,
                                        const char *outdir, int skip_if_exists);

/* One parameter set, host output buffers in the .raw layouts above
 * (any of the three pointers may be NULL to skip that table's copy-back). */
ATMO_API atmo_status atmo_precompute_host(const atmo_params *params, const atmo_options *opts,
                                          float *transmittance, float *irradiance,
                                          float *inscatter);

/* Engine: owns the device work tables and launch descriptors.  One engine per
 * thread; calls on one engine are serialised by the caller. */
typedef struct atmo_engine atmo_engine;

ATMO_API atmo_status atmo_engine_create(atmo_engine **out, int device);
ATMO_API atmo_status atmo_engine_destroy(atmo_engine *e);

/* Device output pointers of one parameter set (.raw layouts). */
typedef struct atmo_tables {
    float *transmittance;
    float *irradiance;
    float *inscatter;
} atmo_tables;

/* Batched precompute of nsets independent parameter sets (each with its own
 * table resolutions) into caller-allocated device tables out[0..nsets).
 * Work tables are allocated (and kept) by the engine.  Replaces running the
 * reference's preprocess once per set. */
ATMO_API atmo_status atmo_engine_precompute(atmo_engine *e, const atmo_params *sets, int nsets,
                                            const atmo_options *opts, const atmo_tables *out,
                                            void *stream);

/* Per-pass GPU times of the last atmo_engine_precompute run with
 * ATMO_FLAG_TIME_PASSES (hipEvents on the caller's stream).  pass: 0 transmittance,
 * 1 irradiance1, 2 inscatter1, 3 inscatterS, 4 irradianceN, 5 inscatterN;
 * order: the scattering order the launch computes (1 for the first three). */
typedef struct atmo_pass_time {
    int pass;
    int order;
    float ms;
} atmo_pass_time;
ATMO_API atmo_status atmo_engine_pass_times(atmo_engine *e, atmo_pass_time *out, int capacity,
                                            int *count);

/* Kernel variants the autotuner picked for the engine's current plan.  After
 * the first run of a batch shape not seen before in the process, the engine
 * times every variant of the two dominant passes and keeps the fastest.
 * k_inscatterS variants (2 / 3 / 4 / 8 texels per thread, or each set with
 * the G whose workgroup-occupancy cost ceil(RES_MU * ceil(RES_NU / G) / 256) * G
 * is lowest, reported as 0) and the k_inscatterN forms (direct gathers /
 * LDS-staged blended rows) give bit-identical tables: tuning changes the time
 * only.  Times in ms: inscatterS_ms[i] for
 * 2 / 3 / 4 / 8 / per-set texels per thread (one order-2 + (orders-2) later
 * launches), inscatterN_ms[9f + 3c + x] for form f (direct, staged with 1
 * sample per LDS round, staged with 2), at most 1 / 2 / 4 mu rows per
 * workgroup (c), tile order x (0 mu
 * rows, 1 XCD-banded mu rows, 2 XCD-banded view lines sorted by impact
 * parameter); -1 = not timed (a k_inscatterN form that does not fit the
 * device's LDS, or the per-set k_inscatterS choice when every set makes the
 * same one). */
typedef struct atmo_tuning {
    int tuned;
    int inscatterS_texels_per_thread;  /* 2, 3, 4, 8; 0: per set */
    int inscatterN_rows_per_workgroup;
    int inscatterN_xcd_order;      /* 0, 1 or 2 (see inscatterN_ms) */
    float inscatterS_ms[5];
    float inscatterN_ms[27];
    int inscatterN_staged_samples;  /* 0: direct gathers; > 0: staged, samples per LDS round */
} atmo_tuning;
ATMO_API atmo_status atmo_engine_tuning(atmo_engine *e, atmo_tuning *out);

/* Bytes of device work memory the engine would hold for this batch. */
ATMO_API atmo_status atmo_engine_workspace_size(const atmo_params *sets, int nsets,
                                                const atmo_options *opts, size_t *bytes);

/* Allocation-free precompute.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, tables and semantics as
 * atmo_engine_precompute, but every device buffer it needs (work tables,
 * set descriptors, tile and direction tables, check counters) is a part of
 * one static device array of ATMO_NOALLOC_MAX_BYTES per GPU, which the
 * library holds from its first call on a GPU until the process ends; it
 * calls no hipMalloc / hipFree.  A batch that needs more returns
 * ATMO_ERR_INVALID_ARGUMENT before any device work (one set at the reference
 * configuration, 32x128x32x8 with T 256x64 and E 64x16, needs 48.4 MiB).
 * The caller passes the output tables.  Runs on the current device and the
 * caller's stream, launches eagerly (no hipGraph) and never autotunes: it
 * takes the kernel variants an engine in this process tuned for the same
 * shape class (then its tables equal that engine's bit for bit), else the
 * defaults (ATMO_FLAG_NO_AUTOTUNE's tables).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads, other
 * streams) race; calls on one stream are ordered. */
#define ATMO_NOALLOC_MAX_BYTES (64u * 1024u * 1024u)
ATMO_API atmo_status atmo_precompute_noalloc(const atmo_params *sets, int nsets,
                                             const atmo_options *opts, const atmo_tables *out,
                                             void *stream);

/* ---------------------------------------------------------------------- */
/* Pass-level entry points (one reference shader program each)             */
/* All pointers: device float4 tables on the engine's device.              */
/* The segment transmittances T(x)/T(y) (common.glsl:272-282) are the      */
/* plain FP32 quotient of two lookups in the caller's T table, as in the   */
/* reference shaders.  Where both lookups fall below FP32's range (grazing */
/* paths with optical depths above ~87: dense or very large atmospheres)   */
/* that quotient is 0/0 and the pass differs from an FP64 evaluation;      */
/* atmo_engine_precompute evaluates these ratios from optical depths and   */
/* has no such limit.                                                      */
/* ---------------------------------------------------------------------- */

/* transmittance.glsl:78-98 -> T[TRANSMITTANCE_H][TRANSMITTANCE_W] */
ATMO_API atmo_status atmo_pass_transmittance(atmo_engine *e, const atmo_params *p,
                                             const atmo_options *o, float *T4, void *stream);
/* irradiance1.glsl:78-82 -> dE[SKY_H][SKY_W] */
ATMO_API atmo_status atmo_pass_irradiance1(atmo_engine *e, const atmo_params *p,
                                           const atmo_options *o, const float *T4, float *dE4,
                                           void *stream);
/* inscatter1.glsl:104-150 -> dSR, dSM [RES_R][RES_MU][RES_NU*RES_MU_S];
 * if S4 != NULL also copyInscatter1.glsl:104-109: S = (dSR.rgb, dSM.r). */
ATMO_API atmo_status atmo_pass_inscatter1(atmo_engine *e, const atmo_params *p,
                                          const atmo_options *o, const float *T4, float *dSR4,
                                          float *dSM4, float *S4, void *stream);
/* inscatterS.glsl:111-189 -> dJ (first != 0: order 2, Rayleigh/Mie separate) */
ATMO_API atmo_status atmo_pass_inscatterS(atmo_engine *e, const atmo_params *p,
                                          const atmo_options *o, const float *T4,
                                          const float *dE4, const float *dSR4, const float *dSM4,
                                          int first, float *dJ4, void *stream);
/* irradianceN.glsl:85-114 -> dE; if E4 != NULL also copyIrradiance k=1 + ADD blend:
 * E += dE (PreprocessAtmo.cpp:335-341). */
ATMO_API atmo_status atmo_pass_irradianceN(atmo_engine *e, const atmo_params *p,
                                           const atmo_options *o, const float *dSR4,
                                           const float *dSM4, int first, float *dE4, float *E4,
                                           void *stream);
/* inscatterN.glsl:105-131 -> dS (written to dSR4); if S4 != NULL also
 * copyInscatterN.glsl:105-110 + ADD blend: S.rgb += dS / phaseR(nu). */
ATMO_API atmo_status atmo_pass_inscatterN(atmo_engine *e, const atmo_params *p,
                                          const atmo_options *o, const float *T4,
                                          const float *dJ4, float *dSR4, float *S4, void *stream);

/* ---------------------------------------------------------------------- */
/* .raw files (PreprocessAtmo.cpp:355-405)                                  */
/* ---------------------------------------------------------------------- */

/* Writes width*height*depth(or 1)*components floats + the 20-byte trailer
 * {0xCAFEBABE, width, height(*depth), depth (0 for 2D), components}. */
ATMO_API atmo_status atmo_write_raw(const char *path, const float *data, int width, int height,
                                    int depth, int components);
/* Reads the trailer of a .raw file (header[5]) and validates magic + size. */
ATMO_API atmo_status atmo_read_raw_header(const char *path, int32_t header[5]);
/* Reads the payload of a validated .raw file into data (capacity in floats). */
ATMO_API atmo_status atmo_read_raw(const char *path, float *data, size_t capacity_floats,
                                   int32_t header[5]);

#ifdef __cplusplus
}
#endif

#endif /* ATMO_H_ */
