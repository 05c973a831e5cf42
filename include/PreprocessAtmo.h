/*
 * PreprocessAtmo.h — C++ drop-in for proland::preprocessAtmo.
 *
 * Same declarations as src/atmo/sources/proland/preprocess/atmo/PreprocessAtmo.h:35-77:
 * class AtmoParameters (same members, same order, same defaults, 96 bytes) and
 *     void preprocessAtmo(const AtmoParameters &params, const char *output);
 * so callers (e.g. src/atmo/examples/atmo/HelloWorld.cpp:240,
 * src/demo/ProlandDemo.cpp:404) compile and link unchanged against
 * libproland_atmo.so.  Deviation: the reference never returns (it runs a
 * GLUT main loop and calls exit(0) after writing the tables,
 * PreprocessAtmo.cpp:406-408); this one returns, and preprocessAtmoStatus()
 * reports the outcome (0 = OK, else an atmo_status code from atmo.h).
 */
#ifndef PROLAND_PREPROCESS_ATMO_MI355X_
#define PROLAND_PREPROCESS_ATMO_MI355X_

#ifndef PROLAND_API
#define PROLAND_API __attribute__((visibility("default")))
#endif

namespace proland
{

/* Layout-compatible stand-in for ork::vec3f (three floats x, y, z). It
 * converts from any type with float-convertible x, y, z members. */
struct vec3f_compat
{
    float x, y, z;
    vec3f_compat() : x(0), y(0), z(0) {}
    vec3f_compat(float x, float y, float z) : x(x), y(y), z(z) {}
    template <class V> vec3f_compat(const V &v) : x(v.x), y(v.y), z(v.z) {}
};

class PROLAND_API AtmoParameters
{
public:
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
    vec3f_compat betaR;
    float HM;
    vec3f_compat betaMSca;
    vec3f_compat betaMEx;
    float mieG;

    AtmoParameters();
};

PROLAND_API void preprocessAtmo(const AtmoParameters &params, const char *output);

/* Status of the last preprocessAtmo call on this thread (0 = OK). */
PROLAND_API int preprocessAtmoStatus();

}

#endif
