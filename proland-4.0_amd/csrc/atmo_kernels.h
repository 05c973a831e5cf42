// atmo_kernels.h — host-callable launchers of the gfx950 pass kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include "atmo_device.h"

namespace atmo {

// Per-theta constants of the inscatterS direction grid (NS thetas).
struct DirJHost {
    float ct, st, dw, pad;
};
// Per-direction constants of the irradianceN hemisphere grid, index ip*(NI/2)+it.
struct DirEHost {
    float wx, wz, wzdw, pad;
};

// The two spherical-integral kernels stage their theta rows in LDS, CH rows at
// a time: CH is the largest count (up to all of them) whose staging fits
// kSphLdsBudget, the dynamic LDS a launch gets without raising the kernel's
// limit; 0 = not even one row fits.  lds_* are the bytes for CH rows.
constexpr size_t kSphLdsBudget = 64 * 1024;
// JX: extra LDS entries of the merged ground rows of one staging round of the
// orders >= 3 pass (the sum over its theta rows of merged_row_extra, maximum
// over the layers and sets of a launch)
// Order 2 stages its 48-byte entries as three planes of float4, kJ1Plane
// entries apart, in unchunked launches whose staging round (all NS rows) fits
// one plane; other launches keep each entry whole.  544 = 32 x 64 directions at
// RES_NU 16.  Chunked: CH < NS, or an order-2 round too large for a plane
// (those launches use the chunked variant, also for a single round).
constexpr int kJ1Plane = 544;
bool inscatterS_chunked(int NS, int CH, int RNU, bool first);
size_t lds_inscatterS(int NS, int CH, int RNU, int SW, int JX, bool first);
// bound on the extra entries of theta row `it` (NS rows) of a layer at radius r
int merged_row_extra(double r, double Rg, double Rt, int SW, int NS, int it);
size_t lds_irradianceN(int CH, int RNU, bool first);
int chunk_inscatterS(int NS, int RNU, int SW, const std::function<int(int)> &jx_of_ch, bool first);
int chunk_irradianceN(int NI, int RNU, bool first);

hipError_t launch_transmittance(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                                hipStream_t st);
hipError_t launch_irradiance1(const SetDesc *sets, const uint2 *tiles, int ntiles, hipStream_t st);
// transmittance tables of a precompute (SetDesc::tview on the inscatter1 tiles,
// SetDesc::tground on the dJ tiles), after the T pass
hipError_t launch_tview(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp, hipStream_t st);
hipError_t launch_tground(const SetDesc *sets, const uint2 *tiles, int ntiles, int NS, hipStream_t st);
// linT / lin4: the sets' parameterisation flags (one per batch), compiled in
hipError_t launch_inscatter1(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             bool linT, hipStream_t st);
// G: texels per thread (2, 3, 4 or 8); CH: theta rows per staging round; RNU:
// the largest RES_NU of the launch's sets
hipError_t launch_inscatterS(const SetDesc *sets, const uint2 *tiles, int ntiles, int NS, int CH, int RNU,
                             int JX, const void *dirT, const void *dirW, bool first, size_t lds, int G,
                             hipStream_t st);
hipError_t launch_irradianceN(const SetDesc *sets, const uint2 *tiles, int ntiles, int NI, int CH,
                              const void *dirs, bool first, size_t lds, hipStream_t st);
// rpw_cap: at most this many mu rows per workgroup (1, 2 or 4); the tile table
// must have been built with the same cap.  sb = 0: direct gathers; sb > 0:
// LDS-staged blended rows, sb samples per round, lds = lds_inscatterN(max over
// the batch of rows-per-workgroup x RNU*RMU_S, sb) bytes of dynamic LDS.
size_t lds_inscatterN(int max_row_texels, int sb);
size_t static_lds_inscatterN();   // bytes of k_inscatterN's static LDS
hipError_t launch_inscatterN(const SetDesc *sets, const uint2 *tiles, int ntiles, int nsamp,
                             int rpw_cap, int sb, size_t lds, bool lin4, hipStream_t st);
hipError_t launch_check(const SetDesc *sets, const uint2 *tiles, int ntiles, unsigned int *flags,
                        hipStream_t st);

}  // namespace atmo
