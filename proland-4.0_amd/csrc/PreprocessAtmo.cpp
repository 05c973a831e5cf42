// PreprocessAtmo.cpp — proland::preprocessAtmo on top of the C ABI.
// Replaces src/atmo/sources/proland/preprocess/atmo/PreprocessAtmo.cpp:90-98 (defaults)
// and :424-436 (entry point: skip if <output>/inscatter.raw exists, else
// precompute and write the three .raw tables).
#include "../../include/PreprocessAtmo.h"

#include <cstdio>
#include <cstring>

#include "../../include/atmo.h"

static_assert(sizeof(proland::AtmoParameters) == sizeof(atmo_params),
              "AtmoParameters must stay byte-compatible with atmo_params");

namespace proland
{

static thread_local int g_status = 0;

AtmoParameters::AtmoParameters() : Rg(6360.0), Rt(6420.0), RL(6421.0),
    TRANSMITTANCE_W(256), TRANSMITTANCE_H(64),
    SKY_W(64), SKY_H(16),
    RES_R(32), RES_MU(128), RES_MU_S(32), RES_NU(8),
    AVERAGE_GROUND_REFLECTANCE(0.1f),
    HR(8.0), betaR(5.8e-3f, 1.35e-2f, 3.31e-2f),
    HM(1.2f), betaMSca(4e-3f, 4e-3f, 4e-3f), betaMEx(4.44e-3f, 4.44e-3f, 4.44e-3f), mieG(0.8f)
{
}

void preprocessAtmo(const AtmoParameters &params, const char *output)
{
    atmo_params p;
    std::memcpy(&p, &params, sizeof p);
    atmo_status st = atmo_preprocess(&p, output);
    g_status = (int)st;
    if (st != ATMO_OK) {
        std::fprintf(stderr, "preprocessAtmo: %s\n", atmo_last_error());
    }
}

int preprocessAtmoStatus()
{
    return g_status;
}

}
