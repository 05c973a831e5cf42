// atmo_engine.cpp — host side of the MI355X atmo precompute: parameter
// validation, per-set descriptors, batch plans (tile tables + work tables +
// captured hipGraph), the C ABI of include/atmo.h and the .raw file format.
//
// Pass order and blending follow PreprocessAtmo.cpp:273-354 (13-step state
// machine), re-expressed as one kernel launch per pass for a whole batch:
//   T, dE1 (+E=0), dS1 (+S=dS1), then per order n = 2..orders:
//   dJn, dEn (+E+=dEn), dSn (+S+=dSn/phaseR).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/atmo.h"
#include "atmo_kernels.h"

using namespace atmo;

namespace {

thread_local std::string g_err;

atmo_status fail(atmo_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? ATMO_ERR_OUT_OF_MEMORY : ATMO_ERR_DEVICE, \
                        "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

enum Pass { P_T = 0, P_E1, P_S1, P_J, P_EN, P_SN, P_CHECK, P_COUNT };

constexpr size_t kXcds = 8;  // MI355X accelerator complex dies
constexpr size_t kJBlock = 16;   // dJ tiles (mu_s columns) dealt to one XCD together

atmo_options default_options() {
    atmo_options o;
    memset(&o, 0, sizeof o);
    o.orders = 4;
    o.transmittance_samples = 500;
    o.inscatter_samples = 50;
    o.irradiance_samples = 32;
    o.spherical_samples = 16;
    return o;
}

atmo_status validate(const atmo_params &p, int idx) {
    auto bad = [&](const char *what) {
        return fail(ATMO_ERR_INVALID_ARGUMENT, "parameter set %d: %s", idx, what);
    };
    if (!(p.Rg > 0.f) || !(p.Rt > p.Rg) || !(p.RL >= p.Rt)) return bad("need 0 < Rg < Rt <= RL");
    if (p.TRANSMITTANCE_W < 2 || p.TRANSMITTANCE_H < 2) return bad("transmittance table < 2x2");
    if (p.SKY_W < 2 || p.SKY_H < 2) return bad("irradiance table < 2x2");
    if (p.RES_R < 2 || p.RES_MU < 4 || (p.RES_MU & 1) || p.RES_MU_S < 2 || p.RES_NU < 2)
        return bad("inscatter resolution: need RES_R>=2, even RES_MU>=4, RES_MU_S>=2, RES_NU>=2");
    if ((long long)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU > (1LL << 27))
        return bad("inscatter table larger than 2^27 texels (2 GiB per work table)");
    if (p.RES_MU_S * p.RES_NU > 1024) return bad("RES_MU_S * RES_NU must be <= 1024");
    if (p.RES_NU > 64) return bad("RES_NU must be <= 64");
    if ((long long)p.TRANSMITTANCE_W * p.TRANSMITTANCE_H > (1LL << 26) ||
        (long long)p.SKY_W * p.SKY_H > (1LL << 24))
        return bad("transmittance/irradiance table too large");
    if (!(p.HR > 0.f) || !(p.HM > 0.f)) return bad("scale heights must be > 0");
    if (!(p.mieG > -1.f && p.mieG < 1.f)) return bad("mieG must be in (-1, 1)");
    const float *v = &p.betaR.x;
    for (int i = 0; i < 3; ++i)
        if (!(v[i] >= 0.f) || !(p.betaMSca.x >= 0.f) || !(p.betaMEx.x >= 0.f)) return bad("negative beta");
    if (!(p.betaMSca.y >= 0.f && p.betaMSca.z >= 0.f && p.betaMEx.y >= 0.f && p.betaMEx.z >= 0.f))
        return bad("negative beta");
    if (!(p.AVERAGE_GROUND_REFLECTANCE >= 0.f)) return bad("negative ground reflectance");
    return ATMO_OK;
}

atmo_status validate_opts(const atmo_options &o) {
    if (o.orders < 1 || o.orders > 64) return fail(ATMO_ERR_INVALID_ARGUMENT, "orders must be in [1, 64]");
    if (o.transmittance_samples < 1 || o.transmittance_samples > 1000000 || o.inscatter_samples < 1 ||
        o.inscatter_samples > 100000)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "integral sample counts out of range");
    if (o.irradiance_samples < 2 || (o.irradiance_samples & 1) || o.irradiance_samples > 256)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "irradiance_samples must be even, in [2, 256]");
    if (o.spherical_samples < 1 || o.spherical_samples > 128)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "spherical_samples must be in [1, 128]");
    if (o.texels_per_thread != 0 && o.texels_per_thread != 2 && o.texels_per_thread != 3 &&
        o.texels_per_thread != 4 && o.texels_per_thread != 8)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "texels_per_thread must be 0, 2, 3, 4 or 8");
    if (o.inscatterN_form < 0 || o.inscatterN_form > 3)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "inscatterN_form must be 0, 1, 2 or 3");
    return ATMO_OK;
}

// PreprocessAtmo.cpp:255-271 setLayer: computed in double from the float
// parameters (float*float products stay float as in C++), uploaded as FP32.
LayerGeom layer_geom(const atmo_params &p, int l) {
    double r = l / (p.RES_R - 1.0);
    r = r * r;
    float RgRg = p.Rg * p.Rg, RtRt = p.Rt * p.Rt;
    float dRR = RtRt - RgRg;
    r = std::sqrt((double)RgRg + r * (double)dRR) +
        (l == 0 ? 0.01 : (l == p.RES_R - 1 ? -0.001 : 0.0));
    double dmin = p.Rt - r;
    double dmax = std::sqrt(r * r - (double)RgRg) + std::sqrt((double)dRR);
    double dminp = r - p.Rg;
    double dmaxp = std::sqrt(r * r - (double)RgRg);
    LayerGeom g;
    g.r = (double)(float)r;
    g.dhdH[0] = (double)(float)dmin;
    g.dhdH[1] = (double)(float)dmax;
    g.dhdH[2] = (double)(float)dminp;
    g.dhdH[3] = (double)(float)dmaxp;
    return g;
}

// Extra LDS entries the merged ground rows of the orders >= 3 dJ pass need in
// one staging round of ch theta rows: the largest sum of merged_row_extra over
// the rows of a round, over every layer of every set.
int merged_extra(const atmo_params *sets, int nsets, int NS, int ch) {
    int best = 0;
    std::vector<int> ex(NS);
    for (int i = 0; i < nsets; ++i) {
        const atmo_params &p = sets[i];
        for (int l = 0; l < p.RES_R; ++l) {
            const double r = layer_geom(p, l).r;
            for (int it = 0; it < NS; ++it) ex[it] = merged_row_extra(r, p.Rg, p.Rt, p.SKY_W, NS, it);
            for (int c0 = 0; c0 < NS; c0 += ch) {
                int sum = 0;
                for (int it = c0; it < std::min(NS, c0 + ch); ++it) sum += ex[it];
                best = std::max(best, sum);
            }
        }
    }
    return best;
}

// mu of texel row y of a layer (getMuMuSNu, common.glsl:175-190), as the kernels compute it
double row_mu(const atmo_params &p, const LayerGeom &g, int y) {
    const double r = g.r, Rg = p.Rg, Rt = p.Rt, half = p.RES_MU * 0.5;
    double mu;
    if (y < half) {
        double d = 1.0 - y / (half - 1.0);
        d = std::min(std::max(g.dhdH[2], d * g.dhdH[3]), g.dhdH[3] * 0.999);
        mu = (Rg * Rg - r * r - d * d) / (2.0 * r * d);
        const double q = Rg / r;
        mu = std::min(mu, -std::sqrt(1.0 - q * q) - 0.001);
    } else {
        double d = (y - half) / (half - 1.0);
        d = std::min(std::max(g.dhdH[0], d * g.dhdH[1]), g.dhdH[1] * 0.999);
        mu = (Rt * Rt - r * r - d * d) / (2.0 * r * d);
    }
    return std::min(std::max(mu, -1.0), 1.0);
}

void fill_desc(SetDesc &s, const atmo_params &p) {
    memset(&s, 0, sizeof s);
    s.Rg = p.Rg; s.Rt = p.Rt; s.RL = p.RL;
    s.HR = p.HR; s.HM = p.HM; s.mieG = p.mieG; s.alpha = p.AVERAGE_GROUND_REFLECTANCE;
    s.H = std::sqrt(s.Rt * s.Rt - s.Rg * s.Rg);
    s.Rg2 = s.Rg * s.Rg; s.Rt2 = s.Rt * s.Rt; s.RL2 = s.RL * s.RL;
    s.dJb = 16;
    s.bR[0] = p.betaR.x; s.bR[1] = p.betaR.y; s.bR[2] = p.betaR.z;
    s.bMS[0] = p.betaMSca.x; s.bMS[1] = p.betaMSca.y; s.bMS[2] = p.betaMSca.z;
    s.bME[0] = p.betaMEx.x; s.bME[1] = p.betaMEx.y; s.bME[2] = p.betaMEx.z;
    s.f_Rg = (float)s.Rg;
    s.f_dRt = (float)(s.Rt - s.Rg);
    s.f_inv_dRt = (float)(1.0 / (s.Rt - s.Rg));
    s.f_H = (float)s.H;
    s.f_invH = (float)(1.0 / s.H);
    s.tMu = std::tan(1.5) / 1.15;
    s.f_tMu = (float)s.tMu;
    s.f_tMuS = (float)std::tan(1.26 * 1.1);
    s.kR = (float)(3.0 / (16.0 * kPI));
    double g = s.mieG;
    s.kM = (float)(1.5 * 1.0 / (4.0 * kPI) * (1.0 - g * g) / (2.0 + g * g));
    s.g2p1 = (float)(1.0 + g * g);
    s.m2g = (float)(-2.0 * g);
    s.TW = p.TRANSMITTANCE_W; s.TH = p.TRANSMITTANCE_H;
    s.SW = p.SKY_W; s.SH = p.SKY_H;
    s.RR = p.RES_R; s.RMU = p.RES_MU; s.RMUS = p.RES_MU_S; s.RNU = p.RES_NU;
    s.W4 = s.RMUS * s.RNU;
    s.n4 = s.RR * s.RMU * s.W4;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct SetSizes {
    size_t nT, nE, n4;
};
SetSizes sizes(const atmo_params &p) {
    return {(size_t)p.TRANSMITTANCE_W * p.TRANSMITTANCE_H, (size_t)p.SKY_W * p.SKY_H,
            (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU};
}
// work tables per set: T4, dE4, E4, dSR4, dSM4, dJ4
size_t work_bytes(const atmo_params &p, const atmo_options &o) {
    SetSizes z = sizes(p);
    size_t b = 2 * align256(z.nT * 16) + 2 * align256(z.nE * 16) + 2 * align256(z.n4 * 16);   // T, tau, ...
    b += align256((size_t)p.RES_R * p.RES_MU * (o.inscatter_samples + 1) * 16);   // tview
    if (o.orders >= 2) {
        b += align256(z.n4 * 12);                                         // dJ, packed rgb
        b += align256((size_t)p.RES_R * o.spherical_samples * 16);        // tground
    }
    return b;
}

// Flags that change what the tables hold (part of a plan's identity).
constexpr int kTableFlags = ATMO_FLAG_HALF_STORAGE | ATMO_FLAG_LINEAR_TRANSMITTANCE | ATMO_FLAG_LINEAR_INSCATTER;
bool lin_t(const atmo_options &o) { return (o.flags & ATMO_FLAG_LINEAR_TRANSMITTANCE) != 0; }
bool lin_4(const atmo_options &o) { return (o.flags & ATMO_FLAG_LINEAR_INSCATTER) != 0; }
void set_table_flags(SetDesc &s, const atmo_options &o) {
    s.hs = (o.flags & ATMO_FLAG_HALF_STORAGE) ? 1 : 0;
    s.linT = (o.flags & ATMO_FLAG_LINEAR_TRANSMITTANCE) ? 1 : 0;
    s.lin4 = (o.flags & ATMO_FLAG_LINEAR_INSCATTER) ? 1 : 0;
}

bool same_params(const atmo_params &a, const atmo_params &b) { return memcmp(&a, &b, sizeof a) == 0; }
bool same_opts(const atmo_options &a, const atmo_options &b) {
    return a.orders == b.orders && a.transmittance_samples == b.transmittance_samples &&
           a.inscatter_samples == b.inscatter_samples && a.irradiance_samples == b.irradiance_samples &&
           a.spherical_samples == b.spherical_samples && a.texels_per_thread == b.texels_per_thread &&
           a.inscatterN_form == b.inscatterN_form &&
           (a.flags & kTableFlags) == (b.flags & kTableFlags);
}

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    bool owned = true;   // false: a part of memory the buffer neither frees nor grows
    ~DevBuf() {
        if (p && owned) (void)hipFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (!owned) return hipErrorOutOfMemory;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
        if (e == hipSuccess) n = bytes;
        return e;
    }
};

// Where a plan's device buffers come from: hipMalloc (base == nullptr), or
// consecutive 256-byte aligned parts of one caller-provided device region
// (the allocation-free entry point).
struct Region {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    hipError_t get(DevBuf &b, size_t bytes) {
        if (!base) return b.ensure(bytes);
        bytes = align256(std::max<size_t>(bytes, 1));
        if (used + bytes > cap) return hipErrorOutOfMemory;
        if (b.p && b.owned) (void)hipFree(b.p);
        b.p = base + used;
        b.n = bytes;
        b.owned = false;
        used += bytes;
        return hipSuccess;
    }
};

// k_inscatterS texels per thread for one set: a workgroup integrates the
// RES_MU x RES_NU texels of one (r, mu_s) column, in RES_MU * ceil(RES_NU/G)
// groups of G texels over 256 threads, so its time is about
// ceil(groups / 256) * G; the smallest such G in {8, 4, 3, 2} (the largest on
// ties, whose per-direction terms are shared most widely).
int jg_for_set(const atmo_params &p) {
    int best = 8;
    long best_t = 1L << 40;
    for (int g : {8, 4, 3, 2}) {
        const long groups = (long)p.RES_MU * ((p.RES_NU + g - 1) / g);
        const long t = ((groups + 255) / 256) * g;
        if (t < best_t) {
            best_t = t;
            best = g;
        }
    }
    return best;
}

struct JSeg {   // a run of P_J tiles whose sets share one G
    int G, off, cnt;
};

// Kernel variants chosen by the autotuner, with the times measured for each.
struct Tuning {
    int jG = 4, snCap = 4, snXcd = 1, snSb = 1;   // jG = 0: per-set choice
    float msJ[5] = {0, 0, 0, 0, 0};  // G = 2, 3, 4, 8, per set (order-2 + order->=3 launches)
    float msN[27] = {};   // [direct, staged 1, staged 2][caps 1, 2, 4][plain, XCD, XCD + line order]
};

// staged samples per LDS round of a k_inscatterN form (1 direct: 0; 0, the
// default before tuning: the one-sample rounds of form 3)
int form_sb(int form) { return form == 1 ? 0 : (form == 2 ? 2 : 1); }
// largest rows-per-workgroup x RNU*RMU_S of a batch (k_inscatterN staging)
int sn_max_row_texels(const std::vector<atmo_params> &ps, int snCap) {
    int m = 0;
    for (const atmo_params &p : ps) {
        const int W4 = p.RES_MU_S * p.RES_NU;
        const int rpw = std::max(1, std::min(std::min(snCap, 4), 256 / W4));
        m = std::max(m, rpw * W4);
    }
    return m;
}

struct Plan {
    std::vector<atmo_params> params;
    atmo_options opts;
    std::vector<SetDesc> hsets;
    int off[P_COUNT], cnt[P_COUNT];
    DevBuf dsets, dlayers, dtiles, ddirs, dflags, arena;
    hipStream_t xs = nullptr;   // the engine's non-blocking stream (uploads)
    size_t dirJT_off = 0, dirJW_off = 0, dirE_off = 0;
    size_t ldsJ1 = 0, ldsJ = 0, ldsE1 = 0, ldsEN = 0;
    int chJ1 = 0, chJ = 0, chE1 = 0, chEN = 0;   // theta rows per LDS staging round
    int maxNU = 0;     // largest RES_NU of the batch (the dJ staging's entry count)
    int jx = 0;   // extra LDS entries of the merged ground rows of an orders >= 3 dJ round
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // kernel variants (picked by the autotuner)
    int jG = 4;        // k_inscatterS texels per thread (0: per set, see jg_for_set)
    std::vector<JSeg> jseg;   // P_J tiles grouped by the per-set G
    int snCap = 4;     // k_inscatterN max mu rows per workgroup
    int snXcd = 1;     // k_inscatterN XCD-banded tile order
    int snSb = 1;      // k_inscatterN: 0 direct gathers, > 0 LDS-staged rows (samples per round)
    bool tuned = false;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~Plan() { drop_graph(); }
};

}  // namespace

namespace atmo {
// Sets the calling thread's atmo_last_error() message (used by the surrogate).
atmo_status set_error(atmo_status st, const char *msg) {
    g_err = msg;
    return st;
}
}  // namespace atmo

struct atmo_engine {
    int device = 0;
    hipStream_t cap = nullptr;  // capture stream
    std::unique_ptr<Plan> plan;
    DevBuf pass_sets, pass_layers, pass_tiles, pass_dirs;
    std::vector<atmo_pass_time> times;
    Tuning last_tuning;
    std::string last_key;
};

namespace {

atmo_status check_device(int device) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ATMO_ERR_DEVICE, "device %d not present (%d devices)", device, n);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ATMO_ERR_DEVICE, "device %d is %s; this library is built for gfx950 only", device,
                    prop.gcnArchName);
    return ATMO_OK;
}

// Direction tables of the two spherical integrals (inscatterS.glsl:109-110,
// 127-145; irradianceN.glsl:83-101), computed in double, stored as float.
void build_dirs(int NS, int NI, std::vector<DirJHost> &jt, std::vector<float> &jw,
                std::vector<DirEHost> &de) {
    const double dth = kPI / NS, dph = kPI / NS;
    jt.resize(NS);
    jw.resize((size_t)NS * 2 * NS * 2);
    for (int it = 0; it < NS; ++it) {
        double th = (it + 0.5) * dth;
        jt[it] = {(float)std::cos(th), (float)std::sin(th), (float)(dth * dph * std::sin(th)), 0.f};
        for (int ip = 0; ip < 2 * NS; ++ip) {
            double ph = (ip + 0.5) * dph;
            jw[((size_t)it * 2 * NS + ip) * 2 + 0] = (float)(std::cos(ph) * std::sin(th));
            jw[((size_t)it * 2 * NS + ip) * 2 + 1] = (float)(std::sin(ph) * std::sin(th));
        }
    }
    const double eth = kPI / NI, eph = kPI / NI;
    const int NTH = NI / 2;
    de.resize((size_t)2 * NI * NTH);
    for (int ip = 0; ip < 2 * NI; ++ip) {
        double ph = (ip + 0.5) * eph;
        for (int it = 0; it < NTH; ++it) {
            double th = (it + 0.5) * eth;
            double dw = eth * eph * std::sin(th);
            double wz = std::cos(th);
            de[(size_t)ip * NTH + it] = {(float)(std::cos(ph) * std::sin(th)), (float)wz,
                                         (float)(wz * dw), 0.f};
        }
    }
}

// Host -> device copy on the engine's own non-blocking stream, waited for on
// the host: no work on the legacy null stream, which other threads' graph
// captures must not be made to depend on.
hipError_t h2d(void *dst, const void *src, size_t n, hipStream_t xs) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, xs);
    return e != hipSuccess ? e : hipStreamSynchronize(xs);
}

// Bytes of the direction tables (build_dirs) and the offsets of their parts.
size_t dirs_bytes(const atmo_options &o, size_t &offJT, size_t &offJW, size_t &offE) {
    const size_t NS = o.spherical_samples, NI = o.irradiance_samples;
    offJT = 0;
    offJW = align256(NS * sizeof(DirJHost));
    offE = offJW + align256(NS * 2 * NS * 2 * sizeof(float));
    return offE + align256(2 * NI * (NI / 2) * sizeof(DirEHost));
}

atmo_status upload_dirs(DevBuf &buf, hipStream_t xs, const atmo_options &o, size_t &offJT, size_t &offJW,
                        size_t &offE, Region *rg = nullptr) {
    std::vector<DirJHost> jt;
    std::vector<float> jw;
    std::vector<DirEHost> de;
    build_dirs(o.spherical_samples, o.irradiance_samples, jt, jw, de);
    const size_t total = dirs_bytes(o, offJT, offJW, offE);
    Region none;
    HIPCHK((rg ? rg : &none)->get(buf, total));
    std::vector<char> h(total, 0);
    memcpy(h.data() + offJT, jt.data(), jt.size() * sizeof(DirJHost));
    memcpy(h.data() + offJW, jw.data(), jw.size() * sizeof(float));
    memcpy(h.data() + offE, de.data(), de.size() * sizeof(DirEHost));
    HIPCHK(h2d(buf.p, h.data(), total, xs));
    return ATMO_OK;
}

// Tile tables of every pass for a batch.  The spherical-integral passes list
// the most expensive tiles first so mixed-resolution batches drain evenly.
void build_tiles(const std::vector<atmo_params> &ps, const atmo_options &o,
                 std::vector<uint2> &tiles, int off[P_COUNT], int cnt[P_COUNT], int snCap = 4,
                 int snXcd = 1, std::vector<JSeg> *jseg = nullptr) {
    tiles.clear();
    const int n = (int)ps.size();
    for (int pass = 0; pass < P_COUNT; ++pass) {
        off[pass] = (int)tiles.size();
        std::vector<std::pair<double, uint2>> v;
        for (int i = 0; i < n; ++i) {
            const atmo_params &p = ps[i];
            SetSizes z = sizes(p);
            long long nt = 0;
            double cost = 0;
            switch (pass) {
            case P_T: nt = (z.nT + 15) / 16; break;
            case P_E1: nt = (z.nE + 255) / 256; break;
            case P_S1: {
                // workgroups of up to 4 mu rows (k_inscatter1)
                const int W4 = p.RES_MU_S * p.RES_NU;
                const int rpw = std::max(1, std::min(4, 256 / W4));
                nt = (long long)p.RES_R * ((p.RES_MU + rpw - 1) / rpw);
                break;
            }
            case P_SN: {
                // workgroups of up to snCap mu rows (k_inscatterN)
                const int W4 = p.RES_MU_S * p.RES_NU;
                const int rpw = std::max(1, std::min(snCap, 256 / W4));
                nt = (long long)p.RES_R * ((p.RES_MU + rpw - 1) / rpw);
                break;
            }
            case P_J:
                // grouped by the per-set G (descending), most expensive first inside a group
                nt = (long long)p.RES_R * p.RES_MU_S;
                cost = jg_for_set(p) * 1e12 + (double)p.RES_MU * p.RES_NU;
                break;
            case P_EN: nt = (long long)z.nE; cost = (double)p.RES_NU; break;
            case P_CHECK: nt = (std::max(std::max(z.n4, z.nT), z.nE) + 255) / 256; break;
            }
            if (pass == P_J && o.orders < 2) nt = 0;
            if (pass == P_EN && o.orders < 2) nt = 0;
            if (pass == P_SN && o.orders < 2) nt = 0;
            for (long long t = 0; t < nt; ++t) v.push_back({cost, make_uint2((unsigned)i, (unsigned)t)});
        }
        if (pass == P_J || pass == P_EN)
            std::stable_sort(v.begin(), v.end(),
                             [](const std::pair<double, uint2> &a, const std::pair<double, uint2> &b) {
                                 return a.first > b.first;
                             });
        if (pass == P_SN && snXcd == 2) {
            // "line order": a view ray is a straight line, fixed by its impact
            // parameter r sqrt(1 - mu^2) and its half (up / down); rays of different
            // layers on nearby lines gather the same dJ rows.  Sort the tiles by
            // (set, half, impact parameter) before dealing them to the XCDs.
            std::vector<std::pair<std::pair<double, double>, uint2>> k;
            k.reserve(v.size());
            for (auto &e : v) {
                const atmo_params &p = ps[e.second.x];
                const int W4 = p.RES_MU_S * p.RES_NU;
                const int rpw = std::max(1, std::min(std::min(snCap, 4), 256 / W4));
                const int ngy = (p.RES_MU + rpw - 1) / rpw;
                const int l = (int)e.second.y / ngy, yb = ((int)e.second.y % ngy) * rpw;
                const LayerGeom g = layer_geom(p, l);
                const double mu = row_mu(p, g, yb);
                const double half = yb < p.RES_MU / 2 ? 0.0 : 1.0;
                k.push_back({{e.second.x * 2.0 + half, g.r * std::sqrt(std::max(0.0, 1.0 - mu * mu))}, e.second});
            }
            std::stable_sort(k.begin(), k.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            for (size_t i = 0; i < v.size(); ++i) v[i].second = k[i].second;
        }
        if (pass == P_SN && snXcd) {
            // Workgroups are dealt round-robin to the 8 XCDs (each with its own L2);
            // deal them so that XCD x runs a contiguous band of tiles (mu rows, or
            // lines in line order), whose rays gather from overlapping parts of the
            // dJ table.
            const size_t nt = v.size(), chunk = (nt + kXcds - 1) / kXcds;
            std::vector<std::pair<double, uint2>> w;
            w.reserve(nt);
            for (size_t p = 0; p < chunk * kXcds; ++p) {
                const size_t src = (p % kXcds) * chunk + p / kXcds;
                if (src < nt) w.push_back(v[src]);
            }
            v.swap(w);
        }
        if (pass == P_J && jseg) {
            jseg->clear();
            for (size_t i = 0; i < v.size(); ++i) {
                const int g = jg_for_set(ps[v[i].second.x]);
                if (jseg->empty() || jseg->back().G != g) jseg->push_back({g, (int)i, 0});
                jseg->back().cnt += 1;
            }
        }
        if (pass == P_J) {
            // Consecutive tiles are neighbouring mu_s columns of one layer, whose
            // 12-byte dJ texels share cache lines.  Deal them to the XCDs in blocks
            // of kJBlock, so that a block runs on one XCD (one L2) at about the same
            // time and the partial lines merge there before they are written back;
            // block-cyclic, so the order (per-set G groups, expensive sets first)
            // holds at block granularity.  Per launch segment (each jseg is its own
            // launch, whose workgroup ids start at 0).
            auto deal = [&](size_t b0, size_t n) {
                std::vector<std::vector<std::pair<double, uint2>>> q(kXcds);
                for (size_t i = 0; i < n; ++i) q[(i / kJBlock) % kXcds].push_back(v[b0 + i]);
                size_t p = b0;
                for (size_t k = 0; p < b0 + n; ++k)
                    for (size_t x = 0; x < kXcds; ++x)
                        if (k < q[x].size()) v[p++] = q[x][k];
            };
            if (jseg && !jseg->empty())
                for (const JSeg &g : *jseg) deal((size_t)g.off, (size_t)g.cnt);
            else
                deal(0, v.size());
        }
        for (auto &e : v) tiles.push_back(e.second);
        cnt[pass] = (int)tiles.size() - off[pass];
    }
}

// The dJ pass of a plan: one launch with G texels per thread, or with G = 0
// one launch per run of sets sharing their own G.
hipError_t launch_dj(const Plan &pl, const SetDesc *S, const uint2 *T, const char *dirs, bool first, int G,
                     hipStream_t st) {
    const atmo_options &o = pl.opts;
    const size_t lds = first ? pl.ldsJ1 : pl.ldsJ;
    const int ch = first ? pl.chJ1 : pl.chJ;
    if (G != 0)
        return launch_inscatterS(S, T + pl.off[P_J], pl.cnt[P_J], o.spherical_samples, ch, pl.maxNU, pl.jx,
                                 dirs + pl.dirJT_off, dirs + pl.dirJW_off, first, lds, G, st);
    for (const JSeg &g : pl.jseg) {
        hipError_t e = launch_inscatterS(S, T + pl.off[P_J] + g.off, g.cnt, o.spherical_samples, ch, pl.maxNU, pl.jx,
                                         dirs + pl.dirJT_off, dirs + pl.dirJW_off, first, lds, g.G, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Event pairs around each pass launch (ATMO_FLAG_TIME_PASSES).
struct PassTimer {
    struct Rec {
        int pass, order;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    hipStream_t st = nullptr;
    bool on = false;
    ~PassTimer() {
        for (auto &r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    hipError_t begin(int pass, int order) {
        if (!on) return hipSuccess;
        Rec r{pass, order, nullptr, nullptr};
        hipError_t e = hipEventCreate(&r.a);
        if (e == hipSuccess) e = hipEventCreate(&r.b);
        if (e == hipSuccess) e = hipEventRecord(r.a, st);
        recs.push_back(r);
        return e;
    }
    hipError_t end() { return on ? hipEventRecord(recs.back().b, st) : hipSuccess; }
};

atmo_status launch_all(Plan &pl, hipStream_t st, PassTimer *tm = nullptr) {
    PassTimer none;
    PassTimer &t = tm ? *tm : none;
    t.st = st;
    const SetDesc *S = (const SetDesc *)pl.dsets.p;
    const uint2 *T = (const uint2 *)pl.dtiles.p;
    const atmo_options &o = pl.opts;
    const char *dirs = (const char *)pl.ddirs.p;
    HIPCHK(t.begin(0, 1));
    HIPCHK(launch_transmittance(S, T + pl.off[P_T], pl.cnt[P_T], o.transmittance_samples, st));
    // the transmittance tables of the later passes (timed with the T pass)
    HIPCHK(launch_tview(S, T + pl.off[P_S1], pl.cnt[P_S1], o.inscatter_samples, st));
    if (o.orders >= 2) HIPCHK(launch_tground(S, T + pl.off[P_J], pl.cnt[P_J], o.spherical_samples, st));
    HIPCHK(t.end());
    HIPCHK(t.begin(1, 1));
    HIPCHK(launch_irradiance1(S, T + pl.off[P_E1], pl.cnt[P_E1], st));
    HIPCHK(t.end());
    HIPCHK(t.begin(2, 1));
    HIPCHK(launch_inscatter1(S, T + pl.off[P_S1], pl.cnt[P_S1], o.inscatter_samples, lin_t(o), st));
    HIPCHK(t.end());
    for (int order = 2; order <= o.orders; ++order) {
        const bool first = order == 2;
        HIPCHK(t.begin(3, order));
        HIPCHK(launch_dj(pl, S, T, dirs, first, pl.jG, st));
        HIPCHK(t.end());
        HIPCHK(t.begin(4, order));
        HIPCHK(launch_irradianceN(S, T + pl.off[P_EN], pl.cnt[P_EN], o.irradiance_samples,
                                  first ? pl.chE1 : pl.chEN, dirs + pl.dirE_off, first,
                                  first ? pl.ldsE1 : pl.ldsEN, st));
        HIPCHK(t.end());
        HIPCHK(t.begin(5, order));
        HIPCHK(launch_inscatterN(S, T + pl.off[P_SN], pl.cnt[P_SN], o.inscatter_samples, pl.snCap, pl.snSb,
                                 lds_inscatterN(sn_max_row_texels(pl.params, pl.snCap), pl.snSb), lin_4(o),
                                 st));
        HIPCHK(t.end());
    }
    return ATMO_OK;
}

// ---------------------------------------------------------------------------
// Autotuner.  After the first run of a plan whose shape class has not been
// tuned in this process, every variant of the two dominant passes is timed on
// the plan's own (by then scratch) work tables: k_inscatterS with 2, 3, 4 or 8
// texels per thread (orders 2 and >= 3), k_inscatterN in its three forms with
// up to 1, 2 or 4 mu rows per workgroup and three tile orders (its
// descriptors carry no S pointer, so the finished tables are not touched).
// Every variant computes bit-identical tables (the k_inscatterN forms blend the
// same columns in the same FP32 order), so tuning changes the time only.  The
// fastest is kept per shape class and applied before the first launch of any
// later plan of that class.
// ---------------------------------------------------------------------------
std::string shape_key(const std::vector<atmo_params> &ps, const atmo_options &o) {
    std::vector<std::string> parts;
    for (const atmo_params &p : ps) {
        char b[96];
        snprintf(b, sizeof b, "%d.%d.%d.%d.%d", p.RES_R, p.RES_MU, p.RES_MU_S, p.RES_NU, p.SKY_W);
        parts.push_back(b);
    }
    std::sort(parts.begin(), parts.end());
    std::string k;
    char b[96];
    snprintf(b, sizeof b, "NS%d.IS%d.N%zu|", o.spherical_samples, o.inscatter_samples, ps.size());
    k = b;
    // count of each resolution class
    for (size_t i = 0; i < parts.size();) {
        size_t j = i;
        while (j < parts.size() && parts[j] == parts[i]) ++j;
        k += parts[i] + "x" + std::to_string(j - i) + ";";
        i = j;
    }
    return k;
}

std::map<std::string, Tuning> &tuning_cache() {
    static std::map<std::string, Tuning> c;
    return c;
}
std::mutex &tuning_mutex() {
    static std::mutex m;
    return m;
}

atmo_status time_launch(hipStream_t st, float &best, const std::function<hipError_t()> &f) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    best = 1e30f;
    atmo_status rc = ATMO_OK;
    for (int rep = 0; rep < 2 && rc == ATMO_OK; ++rep) {
        hipError_t he = hipEventRecord(a, st);
        if (he == hipSuccess) he = f();
        if (he == hipSuccess) he = hipEventRecord(b, st);
        if (he == hipSuccess) he = hipEventSynchronize(b);
        float ms = 0.f;
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
        if (he != hipSuccess) rc = fail(ATMO_ERR_DEVICE, "autotune: %s", hipGetErrorString(he));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

atmo_status autotune(Plan &pl, hipStream_t st, Tuning &tu) {
    const atmo_options &o = pl.opts;
    const SetDesc *S = (const SetDesc *)pl.dsets.p;
    const uint2 *T = (const uint2 *)pl.dtiles.p;
    const char *dirs = (const char *)pl.ddirs.p;
    HIPCHK(hipStreamSynchronize(st));
    // 0: each set with its own G (jg_for_set), a candidate only when the sets'
    // choices differ (otherwise it is the same launch as one of the others)
    const int Gs[5] = {2, 3, 4, 8, 0};
    const bool mixedG = pl.jseg.size() > 1;
    float bestJ = 1e30f;
    for (int gi = 0; gi < 5; ++gi) {
        if (Gs[gi] == 0 && !mixedG) {
            tu.msJ[gi] = -1.f;
            continue;
        }
        float m1 = 0.f, m2 = 0.f;
        atmo_status rc = time_launch(st, m1, [&]() { return launch_dj(pl, S, T, dirs, true, Gs[gi], st); });
        if (rc != ATMO_OK) return rc;
        rc = time_launch(st, m2, [&]() { return launch_dj(pl, S, T, dirs, false, Gs[gi], st); });
        if (rc != ATMO_OK) return rc;
        // weight: one order-2 launch, (orders - 2) later ones
        tu.msJ[gi] = m1 + m2 * std::max(0, o.orders - 2);
        if (tu.msJ[gi] < bestJ) {
            bestJ = tu.msJ[gi];
            tu.jG = Gs[gi];
        }
    }
    // k_inscatterN variants: descriptors without S (the accumulation is skipped)
    std::vector<SetDesc> hs = pl.hsets;
    for (auto &d : hs) d.S4 = nullptr;
    DevBuf ds, dt;
    HIPCHK(ds.ensure(hs.size() * sizeof(SetDesc)));
    HIPCHK(h2d(ds.p, hs.data(), hs.size() * sizeof(SetDesc), pl.xs));
    const int caps[3] = {1, 2, 4}, sbs[3] = {0, 1, 2};   // sb: staged samples per LDS round
    int dev = 0, maxLds = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    float bestN = 1e30f;
    for (int si = 0; si < 3; ++si)
    for (int ci = 0; ci < 3; ++ci)
        for (int xo = 0; xo < 3; ++xo) {
            const size_t lds = lds_inscatterN(sn_max_row_texels(pl.params, caps[ci]), sbs[si]);
            if (lds + static_lds_inscatterN() > (size_t)maxLds) {  // does not fit this device's LDS
                tu.msN[si * 9 + ci * 3 + xo] = -1.f;
                continue;
            }
            std::vector<uint2> tiles;
            int off[P_COUNT], cnt[P_COUNT];
            build_tiles(pl.params, o, tiles, off, cnt, caps[ci], xo);
            HIPCHK(dt.ensure(tiles.size() * sizeof(uint2)));
            HIPCHK(h2d(dt.p, tiles.data(), tiles.size() * sizeof(uint2), pl.xs));
            float ms = 0.f;
            atmo_status rc = time_launch(st, ms, [&]() {
                return launch_inscatterN((const SetDesc *)ds.p, (const uint2 *)dt.p + off[P_SN], cnt[P_SN],
                                         o.inscatter_samples, caps[ci], sbs[si], lds, lin_4(o), st);
            });
            if (rc != ATMO_OK) return rc;
            tu.msN[si * 9 + ci * 3 + xo] = ms;
            if (ms < bestN) {
                bestN = ms;
                tu.snCap = caps[ci];
                tu.snXcd = xo;
                tu.snSb = sbs[si];
            }
        }
    return ATMO_OK;
}

// Applies a tuning to a plan: rebuilds the tile table if the k_inscatterN
// variant changed and drops the captured graph.
atmo_status apply_tuning(Plan &pl, const Tuning &tu) {
    const bool retile = tu.snCap != pl.snCap || tu.snXcd != pl.snXcd;
    pl.jG = pl.opts.texels_per_thread ? pl.opts.texels_per_thread : tu.jG;
    pl.snCap = tu.snCap;
    pl.snXcd = tu.snXcd;
    pl.snSb = pl.opts.inscatterN_form ? form_sb(pl.opts.inscatterN_form) : tu.snSb;
    pl.tuned = true;
    if (retile) {
        std::vector<uint2> tiles;
        build_tiles(pl.params, pl.opts, tiles, pl.off, pl.cnt, pl.snCap, pl.snXcd, &pl.jseg);
        HIPCHK(pl.dtiles.ensure(tiles.size() * sizeof(uint2)));
        HIPCHK(h2d(pl.dtiles.p, tiles.data(), tiles.size() * sizeof(uint2), pl.xs));
    }
    pl.drop_graph();
    return ATMO_OK;
}

// rg: where the device buffers come from (nullptr: hipMalloc); tu: kernel
// variants to build the plan with (nullptr: the defaults, tuned later).
atmo_status build_plan(hipStream_t xs, const atmo_params *sets, int nsets, const atmo_options &o,
                       std::unique_ptr<Plan> &out, Region *rg = nullptr, const Tuning *tu = nullptr) {
    std::unique_ptr<Plan> pl(new Plan);
    Region heap;
    Region &R = rg ? *rg : heap;
    pl->xs = xs;
    pl->params.assign(sets, sets + nsets);
    pl->opts = o;
    pl->jG = o.texels_per_thread;   // 0: each set with its own G until the tuner has run
    pl->snSb = form_sb(o.inscatterN_form);
    if (tu) {
        pl->jG = o.texels_per_thread ? o.texels_per_thread : tu->jG;
        pl->snCap = tu->snCap;
        pl->snXcd = tu->snXcd;
        pl->snSb = o.inscatterN_form ? form_sb(o.inscatterN_form) : tu->snSb;
        pl->tuned = true;
    }
    // work arena
    size_t total = 0;
    std::vector<size_t> base(nsets);
    for (int i = 0; i < nsets; ++i) {
        base[i] = total;
        total += work_bytes(sets[i], o);
    }
    HIPCHK(R.get(pl->arena, total));
    // layers
    std::vector<LayerGeom> layers;
    std::vector<size_t> loff(nsets);
    for (int i = 0; i < nsets; ++i) {
        loff[i] = layers.size();
        for (int l = 0; l < sets[i].RES_R; ++l) layers.push_back(layer_geom(sets[i], l));
    }
    HIPCHK(R.get(pl->dlayers, layers.size() * sizeof(LayerGeom)));
    HIPCHK(h2d(pl->dlayers.p, layers.data(), layers.size() * sizeof(LayerGeom), pl->xs));
    // descriptors (output pointers are filled in per call)
    pl->hsets.resize(nsets);
    char *ar = (char *)pl->arena.p;
    for (int i = 0; i < nsets; ++i) {
        SetDesc &s = pl->hsets[i];
        fill_desc(s, sets[i]);
        SetSizes z = sizes(sets[i]);
        char *b = ar + base[i];
        s.T4 = (float4 *)b; b += align256(z.nT * 16);
        s.tau4 = (float4 *)b; b += align256(z.nT * 16);
        s.tview = (float4 *)b; b += align256((size_t)sets[i].RES_R * sets[i].RES_MU * (o.inscatter_samples + 1) * 16);
        s.dE4 = (float4 *)b; b += align256(z.nE * 16);
        s.E4 = (float4 *)b; b += align256(z.nE * 16);
        s.dSR4 = (float4 *)b; b += align256(z.n4 * 16);
        s.dSM4 = (float4 *)b; b += align256(z.n4 * 16);
        s.dJ4 = o.orders >= 2 ? (float4 *)b : nullptr;
        if (o.orders >= 2) {
            b += align256(z.n4 * 12);
            s.tground = (float4 *)b;
        }
        s.dJb = 12;  // the ray gathers of dSn read 12 bytes per texel
        set_table_flags(s, o);
        s.layers = (const LayerGeom *)pl->dlayers.p + loff[i];
    }
    HIPCHK(R.get(pl->dsets, nsets * sizeof(SetDesc)));
    // tiles
    std::vector<uint2> tiles;
    build_tiles(pl->params, o, tiles, pl->off, pl->cnt, pl->snCap, pl->snXcd, &pl->jseg);
    HIPCHK(R.get(pl->dtiles, tiles.size() * sizeof(uint2)));
    HIPCHK(h2d(pl->dtiles.p, tiles.data(), tiles.size() * sizeof(uint2), pl->xs));
    HIPCHK(R.get(pl->dflags, nsets * sizeof(unsigned)));
    atmo_status st = upload_dirs(pl->ddirs, pl->xs, o, pl->dirJT_off, pl->dirJW_off, pl->dirE_off, &R);
    if (st != ATMO_OK) return st;
    int maxNU = 0, maxSW = 0;
    for (int i = 0; i < nsets; ++i) {
        maxNU = std::max(maxNU, sets[i].RES_NU);
        maxSW = std::max(maxSW, sets[i].SKY_W);
    }
    auto jx_of_ch = [&](int ch) { return merged_extra(sets, nsets, o.spherical_samples, ch); };
    pl->chJ1 = chunk_inscatterS(o.spherical_samples, maxNU, maxSW, jx_of_ch, true);
    pl->chJ = chunk_inscatterS(o.spherical_samples, maxNU, maxSW, jx_of_ch, false);
    pl->jx = pl->chJ ? jx_of_ch(pl->chJ) : 0;
    pl->chE1 = chunk_irradianceN(o.irradiance_samples, maxNU, true);
    pl->chEN = chunk_irradianceN(o.irradiance_samples, maxNU, false);
    if (!pl->chJ1 || !pl->chJ || !pl->chE1 || !pl->chEN)
        return fail(ATMO_ERR_UNSUPPORTED, "SKY_W %d: the ground irradiance row does not fit in LDS", maxSW);
    pl->maxNU = maxNU;
    pl->ldsJ1 = lds_inscatterS(o.spherical_samples, pl->chJ1, maxNU, maxSW, 0, true);
    pl->ldsJ = lds_inscatterS(o.spherical_samples, pl->chJ, maxNU, maxSW, pl->jx, false);
    pl->ldsE1 = lds_irradianceN(pl->chE1, maxNU, true);
    pl->ldsEN = lds_irradianceN(pl->chEN, maxNU, false);
    out = std::move(pl);
    return ATMO_OK;
}

// Device bytes build_plan takes from a Region for this batch (every part
// rounded to 256 bytes, as Region::get does).
size_t plan_bytes(const std::vector<atmo_params> &ps, const atmo_options &o, int snCap, int snXcd) {
    size_t work = 0, nl = 0;
    for (const atmo_params &p : ps) {
        work += work_bytes(p, o);
        nl += (size_t)p.RES_R;
    }
    std::vector<uint2> tiles;
    int off[P_COUNT], cnt[P_COUNT];
    build_tiles(ps, o, tiles, off, cnt, snCap, snXcd);
    size_t a, b, c;
    return align256(work) + align256(nl * sizeof(LayerGeom)) + align256(ps.size() * sizeof(SetDesc)) +
           align256(tiles.size() * sizeof(uint2)) + align256(ps.size() * sizeof(unsigned)) +
           align256(dirs_bytes(o, a, b, c));
}

// Runs every pass of a plan on the caller's stream: with pass timing (eager
// launches between hipEvent pairs), eagerly, or by replaying the plan's graph
// (captured on first use).
atmo_status run_plan(atmo_engine *e, Plan &pl, const atmo_options &o, hipStream_t hs) {
    atmo_status st;
    if (o.flags & ATMO_FLAG_TIME_PASSES) {
        PassTimer tm;
        tm.on = true;
        if ((st = launch_all(pl, hs, &tm)) != ATMO_OK) return st;
        HIPCHK(hipStreamSynchronize(hs));
        e->times.clear();
        for (auto &r : tm.recs) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            e->times.push_back({r.pass, r.order, ms});
        }
    } else if (o.flags & ATMO_FLAG_NO_GRAPH) {
        if ((st = launch_all(pl, hs)) != ATMO_OK) return st;
    } else {
        bool eager = false;
        if (!pl.gexec) {
            HIPCHK(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
            atmo_status cs = launch_all(pl, e->cap);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(e->cap, &g);
            if (ce == hipErrorStreamCaptureInvalidated) {
                // another thread's work on the legacy stream can invalidate a
                // capture in progress: run this call's launches directly and
                // capture again on the next call
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                eager = true;
            } else {
                if (cs != ATMO_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return cs;
                }
                HIPCHK(ce);
                pl.graph = g;
                HIPCHK(hipGraphInstantiate(&pl.gexec, g, nullptr, nullptr, 0));
            }
        }
        if (eager) {
            if ((st = launch_all(pl, hs)) != ATMO_OK) return st;
        } else {
            HIPCHK(hipGraphLaunch(pl.gexec, hs));
        }
    }
    return ATMO_OK;
}

atmo_status check_out(const atmo_tables &t, int i) {
    if (!t.transmittance || !t.irradiance || !t.inscatter)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "output tables of set %d: NULL pointer", i);
    if (((uintptr_t)t.inscatter & 15) != 0)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "inscatter output of set %d is not 16-byte aligned", i);
    if (((uintptr_t)t.transmittance & 3) || ((uintptr_t)t.irradiance & 3))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "outputs of set %d are not 4-byte aligned", i);
    return ATMO_OK;
}

// ATMO_FLAG_CHECK_NUMERICS: counts, per set, the texels of the finished tables
// that are non-finite or out of range (k_check) and reports the first bad set.
atmo_status check_numerics(Plan &pl, hipStream_t hs) {
    const int nsets = (int)pl.params.size();
    HIPCHK(hipMemsetAsync(pl.dflags.p, 0, nsets * sizeof(unsigned), hs));
    HIPCHK(launch_check((const SetDesc *)pl.dsets.p, (const uint2 *)pl.dtiles.p + pl.off[P_CHECK],
                        pl.cnt[P_CHECK], (unsigned *)pl.dflags.p, hs));
    std::vector<unsigned> flags(nsets);
    HIPCHK(hipMemcpyAsync(flags.data(), pl.dflags.p, nsets * sizeof(unsigned), hipMemcpyDeviceToHost, hs));
    HIPCHK(hipStreamSynchronize(hs));
    for (int i = 0; i < nsets; ++i)
        if (flags[i]) return fail(ATMO_ERR_NUMERIC, "set %d: %u texels non-finite or out of range", i, flags[i]);
    return ATMO_OK;
}

}  // namespace

// ===========================================================================
// Work memory of the allocation-free precompute (atmo_precompute_noalloc):
// one file-scope device array, which every GPU the library runs on holds (its
// own copy, at its own address) until the process ends.  Host code never
// passes `scratch` itself to a kernel or a copy (in host code the name is a
// host-side stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device.
// ===========================================================================
static __device__ __align__(256) unsigned char scratch[ATMO_NOALLOC_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device.
atmo_status scratch_ptr(char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ATMO_ERR_DEVICE, "device %d: no work array slot", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ATMO_ERR_DEVICE, "device %d: hipGetSymbolAddress gave a null work array", dev);
        g_scratch[dev] = (char *)p;
    }
    *out = g_scratch[dev];
    return ATMO_OK;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void atmo_params_default(atmo_params *p) {
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->Rg = 6360.0f; p->Rt = 6420.0f; p->RL = 6421.0f;
    p->TRANSMITTANCE_W = 256; p->TRANSMITTANCE_H = 64;
    p->SKY_W = 64; p->SKY_H = 16;
    p->RES_R = 32; p->RES_MU = 128; p->RES_MU_S = 32; p->RES_NU = 8;
    p->AVERAGE_GROUND_REFLECTANCE = 0.1f;
    p->HR = 8.0f;
    p->betaR = {5.8e-3f, 1.35e-2f, 3.31e-2f};
    p->HM = 1.2f;
    p->betaMSca = {4e-3f, 4e-3f, 4e-3f};
    p->betaMEx = {4.44e-3f, 4.44e-3f, 4.44e-3f};
    p->mieG = 0.8f;
}

void atmo_options_default(atmo_options *o) {
    if (o) *o = default_options();
}

const char *atmo_last_error(void) { return g_err.c_str(); }

const char *atmo_version(void) { return "proland-atmo-mi355x 0.1 (gfx950)"; }

atmo_status atmo_engine_create(atmo_engine **out, int device) {
    if (!out) return fail(ATMO_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    atmo_status st = check_device(device);
    if (st != ATMO_OK) return st;
    HIPCHK(hipSetDevice(device));
    atmo_engine *e = new atmo_engine;
    e->device = device;
    hipError_t he = hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return fail(ATMO_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    *out = e;
    g_err.clear();
    return ATMO_OK;
}

atmo_status atmo_engine_destroy(atmo_engine *e) {
    if (!e) return ATMO_OK;
    (void)hipSetDevice(e->device);
    (void)hipDeviceSynchronize();
    e->plan.reset();
    if (e->cap) (void)hipStreamDestroy(e->cap);
    delete e;
    return ATMO_OK;
}

atmo_status atmo_engine_pass_times(atmo_engine *e, atmo_pass_time *out, int capacity, int *count) {
    if (!e || !count || capacity < 0 || (capacity > 0 && !out))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    int n = std::min<int>(capacity, (int)e->times.size());
    for (int i = 0; i < n; ++i) out[i] = e->times[i];
    *count = (int)e->times.size();
    return ATMO_OK;
}

atmo_status atmo_engine_tuning(atmo_engine *e, atmo_tuning *out) {
    if (!e || !out) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    memset(out, 0, sizeof *out);
    out->tuned = !e->last_key.empty();
    out->inscatterS_texels_per_thread = e->last_tuning.jG;
    out->inscatterN_rows_per_workgroup = e->last_tuning.snCap;
    out->inscatterN_xcd_order = e->last_tuning.snXcd;
    for (int i = 0; i < 5; ++i) out->inscatterS_ms[i] = e->last_tuning.msJ[i];
    for (int i = 0; i < 27; ++i) out->inscatterN_ms[i] = e->last_tuning.msN[i];
    out->inscatterN_staged_samples = e->last_tuning.snSb;
    return ATMO_OK;
}

atmo_status atmo_engine_workspace_size(const atmo_params *sets, int nsets, const atmo_options *opts,
                                       size_t *bytes) {
    if (!sets || nsets <= 0 || !bytes) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    atmo_options o = opts ? *opts : default_options();
    atmo_status st = validate_opts(o);
    if (st != ATMO_OK) return st;
    size_t t = 0;
    for (int i = 0; i < nsets; ++i) {
        if ((st = validate(sets[i], i)) != ATMO_OK) return st;
        t += work_bytes(sets[i], o);
    }
    *bytes = t;
    return ATMO_OK;
}

atmo_status atmo_engine_precompute(atmo_engine *e, const atmo_params *sets, int nsets,
                                   const atmo_options *opts, const atmo_tables *out, void *stream) {
    if (!e || !sets || nsets <= 0 || !out) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    atmo_options o = opts ? *opts : default_options();
    atmo_status st = validate_opts(o);
    if (st != ATMO_OK) return st;
    for (int i = 0; i < nsets; ++i) {
        if ((st = validate(sets[i], i)) != ATMO_OK) return st;
        if ((st = check_out(out[i], i)) != ATMO_OK) return st;
    }
    HIPCHK(hipSetDevice(e->device));
    hipStream_t hs = (hipStream_t)stream;
    bool reuse = e->plan && (int)e->plan->params.size() == nsets && same_opts(e->plan->opts, o);
    if (reuse)
        for (int i = 0; i < nsets && reuse; ++i) reuse = same_params(e->plan->params[i], sets[i]);
    if (!reuse) {
        HIPCHK(hipStreamSynchronize(hs));
        e->plan.reset();
        if ((st = build_plan(e->cap, sets, nsets, o, e->plan)) != ATMO_OK) return st;
    }
    Plan &pl = *e->plan;
    bool changed = !reuse;
    for (int i = 0; i < nsets; ++i) {
        SetDesc &s = pl.hsets[i];
        if (s.S4 != (float4 *)out[i].inscatter || s.outT != out[i].transmittance ||
            s.outE != out[i].irradiance) {
            s.S4 = (float4 *)out[i].inscatter;
            s.outT = out[i].transmittance;
            s.outE = out[i].irradiance;
            changed = true;
        }
    }
    if (changed) {
        HIPCHK(hipStreamSynchronize(hs));
        HIPCHK(h2d(pl.dsets.p, pl.hsets.data(), nsets * sizeof(SetDesc), e->cap));
    }
    // A plan that has not been tuned takes the cached tuning of its shape class
    // before its first launch; with none cached, it is tuned after this run.
    bool tune_after = false;
    std::string key;
    if (!pl.tuned && (o.orders >= 2) && !(o.flags & ATMO_FLAG_NO_AUTOTUNE)) {
        key = shape_key(pl.params, o);
        Tuning tu;
        bool have = false;
        {
            std::lock_guard<std::mutex> lk(tuning_mutex());
            auto it = tuning_cache().find(key);
            if (it != tuning_cache().end()) {
                tu = it->second;
                have = true;
            }
        }
        if (have) {
            if ((st = apply_tuning(pl, tu)) != ATMO_OK) return st;
            e->last_tuning = tu;
            e->last_key = key;
        } else {
            tune_after = true;
        }
    }
    if ((st = run_plan(e, pl, o, hs)) != ATMO_OK) return st;
    if (tune_after) {
        Tuning tu;
        if ((st = autotune(pl, hs, tu)) != ATMO_OK) return st;
        {
            std::lock_guard<std::mutex> lk(tuning_mutex());
            tuning_cache()[key] = tu;
        }
        if ((st = apply_tuning(pl, tu)) != ATMO_OK) return st;
        e->last_tuning = tu;
        e->last_key = key;
        // the tables of this call from the variants every later call runs
        if ((st = run_plan(e, pl, o, hs)) != ATMO_OK) return st;
    }
    if (o.flags & ATMO_FLAG_CHECK_NUMERICS)
        if ((st = check_numerics(pl, hs)) != ATMO_OK) return st;
    g_err.clear();
    return ATMO_OK;
}

atmo_status atmo_precompute_noalloc(const atmo_params *sets, int nsets, const atmo_options *opts,
                                    const atmo_tables *out, void *stream) {
    if (!sets || nsets <= 0 || !out) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    atmo_options o = opts ? *opts : default_options();
    atmo_status st = validate_opts(o);
    if (st != ATMO_OK) return st;
    for (int i = 0; i < nsets; ++i) {
        if ((st = validate(sets[i], i)) != ATMO_OK) return st;
        if ((st = check_out(out[i], i)) != ATMO_OK) return st;
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if ((st = check_device(dev)) != ATMO_OK) return st;
    // the variants of this shape class's cached tuning (an engine in this
    // process tuned it), else the defaults; no autotuning here (it allocates)
    std::vector<atmo_params> ps(sets, sets + nsets);
    Tuning tu;
    bool have = false;
    if (o.orders >= 2 && !(o.flags & ATMO_FLAG_NO_AUTOTUNE)) {
        std::lock_guard<std::mutex> lk(tuning_mutex());
        auto it = tuning_cache().find(shape_key(ps, o));
        if (it != tuning_cache().end()) {
            tu = it->second;
            have = true;
        }
    }
    const int snCap = have ? tu.snCap : 4, snXcd = have ? tu.snXcd : 1;
    const size_t need = plan_bytes(ps, o, snCap, snXcd);
    if (need > (size_t)ATMO_NOALLOC_MAX_BYTES)
        return fail(ATMO_ERR_INVALID_ARGUMENT,
                    "this batch needs %zu bytes of device work memory; the allocation-free precompute "
                    "holds at most %zu (ATMO_NOALLOC_MAX_BYTES): use atmo_engine_precompute",
                    need, (size_t)ATMO_NOALLOC_MAX_BYTES);
    char *base = nullptr;
    if ((st = scratch_ptr(&base)) != ATMO_OK) return st;
    hipStream_t hs = (hipStream_t)stream;
    // every upload below is ordered after the earlier work of the caller's
    // stream, which may still read the array
    Region rg;
    rg.base = base;
    rg.cap = ATMO_NOALLOC_MAX_BYTES;
    std::unique_ptr<Plan> pl;
    if ((st = build_plan(hs, sets, nsets, o, pl, &rg, have ? &tu : nullptr)) != ATMO_OK) return st;
    for (int i = 0; i < nsets; ++i) {
        SetDesc &d = pl->hsets[i];
        d.S4 = (float4 *)out[i].inscatter;
        d.outT = out[i].transmittance;
        d.outE = out[i].irradiance;
    }
    HIPCHK(h2d(pl->dsets.p, pl->hsets.data(), nsets * sizeof(SetDesc), hs));
    if ((st = launch_all(*pl, hs)) != ATMO_OK) return st;
    if (o.flags & ATMO_FLAG_CHECK_NUMERICS)
        if ((st = check_numerics(*pl, hs)) != ATMO_OK) return st;
    g_err.clear();
    return ATMO_OK;
}

atmo_status atmo_precompute_host(const atmo_params *params, const atmo_options *opts, float *T,
                                 float *E, float *S) {
    if (!params) return fail(ATMO_ERR_INVALID_ARGUMENT, "params is NULL");
    atmo_status st = validate(*params, 0);
    if (st != ATMO_OK) return st;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    atmo_engine *e = nullptr;
    if ((st = atmo_engine_create(&e, dev)) != ATMO_OK) return st;
    std::unique_ptr<atmo_engine, atmo_status (*)(atmo_engine *)> guard(e, atmo_engine_destroy);
    SetSizes z = sizes(*params);
    DevBuf out;
    size_t bT = align256(z.nT * 12), bE = align256(z.nE * 12), bS = z.n4 * 16;
    HIPCHK(out.ensure(bT + bE + bS));
    atmo_tables t;
    t.transmittance = (float *)out.p;
    t.irradiance = (float *)((char *)out.p + bT);
    t.inscatter = (float *)((char *)out.p + bT + bE);
    atmo_options o = opts ? *opts : default_options();
    // a one-off engine: no graph to replay and no later call to tune for
    o.flags |= ATMO_FLAG_NO_GRAPH | ATMO_FLAG_NO_AUTOTUNE;
    if ((st = atmo_engine_precompute(e, params, 1, &o, &t, nullptr)) != ATMO_OK) return st;
    HIPCHK(hipDeviceSynchronize());
    if (T) HIPCHK(hipMemcpy(T, t.transmittance, z.nT * 12, hipMemcpyDeviceToHost));
    if (E) HIPCHK(hipMemcpy(E, t.irradiance, z.nE * 12, hipMemcpyDeviceToHost));
    if (S) HIPCHK(hipMemcpy(S, t.inscatter, z.n4 * 16, hipMemcpyDeviceToHost));
    return ATMO_OK;
}

// ---------------------------------------------------------------------------
// Pass-level entry points: one set, caller's float4 tables, synchronous.
// ---------------------------------------------------------------------------
static atmo_status pass_setup(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                              SetDesc &s, atmo_options &oo) {
    if (!e || !p) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    oo = o ? *o : default_options();
    atmo_status st = validate_opts(oo);
    if (st != ATMO_OK) return st;
    if ((st = validate(*p, 0)) != ATMO_OK) return st;
    HIPCHK(hipSetDevice(e->device));
    fill_desc(s, *p);
    set_table_flags(s, oo);
    std::vector<LayerGeom> layers;
    for (int l = 0; l < p->RES_R; ++l) layers.push_back(layer_geom(*p, l));
    HIPCHK(e->pass_layers.ensure(layers.size() * sizeof(LayerGeom)));
    HIPCHK(h2d(e->pass_layers.p, layers.data(), layers.size() * sizeof(LayerGeom), e->cap));
    s.layers = (const LayerGeom *)e->pass_layers.p;
    return ATMO_OK;
}

static atmo_status pass_run(atmo_engine *e, const atmo_params *p, const atmo_options &o,
                            const SetDesc &s, int pass, bool first, void *stream) {
    hipStream_t hs = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(hs));
    HIPCHK(e->pass_sets.ensure(sizeof(SetDesc)));
    HIPCHK(h2d(e->pass_sets.p, &s, sizeof(SetDesc), e->cap));
    std::vector<uint2> tiles;
    int off[P_COUNT], cnt[P_COUNT];
    std::vector<atmo_params> ps(1, *p);
    atmo_options ot = o;
    ot.orders = std::max(ot.orders, 2);  // every pass gets its tiles
    build_tiles(ps, ot, tiles, off, cnt);
    HIPCHK(e->pass_tiles.ensure(tiles.size() * sizeof(uint2)));
    HIPCHK(h2d(e->pass_tiles.p, tiles.data(), tiles.size() * sizeof(uint2), e->cap));
    size_t oJT, oJW, oE;
    atmo_status st = upload_dirs(e->pass_dirs, e->cap, o, oJT, oJW, oE);
    if (st != ATMO_OK) return st;
    const SetDesc *S = (const SetDesc *)e->pass_sets.p;
    const uint2 *T = (const uint2 *)e->pass_tiles.p;
    const char *d = (const char *)e->pass_dirs.p;
    switch (pass) {
    case P_T: HIPCHK(launch_transmittance(S, T + off[P_T], cnt[P_T], o.transmittance_samples, hs)); break;
    case P_E1: HIPCHK(launch_irradiance1(S, T + off[P_E1], cnt[P_E1], hs)); break;
    case P_S1: HIPCHK(launch_inscatter1(S, T + off[P_S1], cnt[P_S1], o.inscatter_samples, lin_t(o), hs)); break;
    case P_J: {
        auto jx_of_ch = [&](int c) { return merged_extra(p, 1, o.spherical_samples, c); };
        const int ch = chunk_inscatterS(o.spherical_samples, p->RES_NU, p->SKY_W, jx_of_ch, first);
        if (!ch) return fail(ATMO_ERR_UNSUPPORTED, "SKY_W %d: the ground irradiance row does not fit in LDS", p->SKY_W);
        const int jx = first ? 0 : jx_of_ch(ch);
        const size_t lds = lds_inscatterS(o.spherical_samples, ch, p->RES_NU, p->SKY_W, jx, first);
        HIPCHK(launch_inscatterS(S, T + off[P_J], cnt[P_J], o.spherical_samples, ch, p->RES_NU, jx, d + oJT, d + oJW, first,
                                 lds, o.texels_per_thread ? o.texels_per_thread : jg_for_set(*p), hs));
        break;
    }
    case P_EN: {
        const int ch = chunk_irradianceN(o.irradiance_samples, p->RES_NU, first);
        if (!ch) return fail(ATMO_ERR_UNSUPPORTED, "irradiance staging does not fit in LDS");
        const size_t lds = lds_irradianceN(ch, p->RES_NU, first);
        HIPCHK(launch_irradianceN(S, T + off[P_EN], cnt[P_EN], o.irradiance_samples, ch, d + oE,
                                  first, lds, hs));
        break;
    }
    case P_SN: {
        const int sb = form_sb(o.inscatterN_form);
        const size_t lds = lds_inscatterN(sn_max_row_texels(std::vector<atmo_params>(1, *p), 4), sb);
        HIPCHK(launch_inscatterN(S, T + off[P_SN], cnt[P_SN], o.inscatter_samples, 4, sb, lds, lin_4(o), hs));
        break;
    }
    }
    HIPCHK(hipStreamSynchronize(hs));
    g_err.clear();
    return ATMO_OK;
}

static bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

atmo_status atmo_pass_transmittance(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                    float *T4, void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!T4 || !al16(T4)) return fail(ATMO_ERR_INVALID_ARGUMENT, "T4 must be a 16-byte aligned device pointer");
    s.T4 = (float4 *)T4;
    return pass_run(e, p, oo, s, P_T, false, stream);
}

atmo_status atmo_pass_irradiance1(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                  const float *T4, float *dE4, void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!T4 || !dE4 || !al16(T4) || !al16(dE4)) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad table pointers");
    s.T4 = (float4 *)T4;
    s.dE4 = (float4 *)dE4;
    return pass_run(e, p, oo, s, P_E1, false, stream);
}

atmo_status atmo_pass_inscatter1(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                 const float *T4, float *dSR4, float *dSM4, float *S4, void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!T4 || !dSR4 || !dSM4 || !al16(T4) || !al16(dSR4) || !al16(dSM4) || (S4 && !al16(S4)))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad table pointers");
    s.T4 = (float4 *)T4;
    s.dSR4 = (float4 *)dSR4;
    s.dSM4 = (float4 *)dSM4;
    s.S4 = (float4 *)S4;
    return pass_run(e, p, oo, s, P_S1, false, stream);
}

atmo_status atmo_pass_inscatterS(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                 const float *T4, const float *dE4, const float *dSR4,
                                 const float *dSM4, int first, float *dJ4, void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!T4 || !dE4 || !dSR4 || !dJ4 || (first && !dSM4) || !al16(T4) || !al16(dE4) || !al16(dSR4) ||
        !al16(dJ4) || (dSM4 && !al16(dSM4)))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad table pointers");
    s.T4 = (float4 *)T4;
    s.dE4 = (float4 *)dE4;
    s.dSR4 = (float4 *)dSR4;
    s.dSM4 = (float4 *)(dSM4 ? dSM4 : dSR4);
    s.dJ4 = dJ4 ? (float4 *)dJ4 : nullptr;
    return pass_run(e, p, oo, s, P_J, first != 0, stream);
}

atmo_status atmo_pass_irradianceN(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                  const float *dSR4, const float *dSM4, int first, float *dE4,
                                  float *E4, void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!dSR4 || !dE4 || (first && !dSM4) || !al16(dSR4) || !al16(dE4) || (dSM4 && !al16(dSM4)) ||
        (E4 && !al16(E4)))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad table pointers");
    s.dSR4 = (float4 *)dSR4;
    s.dSM4 = (float4 *)(dSM4 ? dSM4 : dSR4);
    s.dE4 = (float4 *)dE4;
    s.E4 = (float4 *)E4;
    return pass_run(e, p, oo, s, P_EN, first != 0, stream);
}

atmo_status atmo_pass_inscatterN(atmo_engine *e, const atmo_params *p, const atmo_options *o,
                                 const float *T4, const float *dJ4, float *dSR4, float *S4,
                                 void *stream) {
    SetDesc s;
    atmo_options oo;
    atmo_status st = pass_setup(e, p, o, s, oo);
    if (st != ATMO_OK) return st;
    if (!T4 || !dJ4 || !dSR4 || !al16(T4) || !al16(dJ4) || !al16(dSR4) || (S4 && !al16(S4)))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad table pointers");
    s.T4 = (float4 *)T4;
    s.dJ4 = (float4 *)dJ4;
    s.dSR4 = (float4 *)dSR4;
    s.S4 = (float4 *)S4;
    return pass_run(e, p, oo, s, P_SN, false, stream);
}

// ---------------------------------------------------------------------------
// .raw files (PreprocessAtmo.cpp:355-405): FP32 payload + 20-byte trailer.
// ---------------------------------------------------------------------------
atmo_status atmo_write_raw(const char *path, const float *data, int width, int height, int depth,
                           int components) {
    if (!path || !data || width <= 0 || height <= 0 || depth < 0 || components <= 0)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    size_t n = (size_t)width * height * (depth > 0 ? depth : 1) * components;
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(ATMO_ERR_IO, "cannot open %s: %s", tmp.c_str(), strerror(errno));
    int32_t trailer[5] = {(int32_t)0xCAFEBABE, width, height * (depth > 0 ? depth : 1), depth,
                          components};
    bool ok = fwrite(data, sizeof(float), n, f) == n && fwrite(trailer, sizeof trailer, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        unlink(tmp.c_str());
        return fail(ATMO_ERR_IO, "write failed: %s", tmp.c_str());
    }
    if (rename(tmp.c_str(), path) != 0) {
        unlink(tmp.c_str());
        return fail(ATMO_ERR_IO, "rename to %s failed: %s", path, strerror(errno));
    }
    return ATMO_OK;
}

atmo_status atmo_read_raw_header(const char *path, int32_t header[5]) {
    if (!path || !header) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(ATMO_ERR_IO, "cannot open %s", path);
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return fail(ATMO_ERR_IO, "seek failed: %s", path);
    }
    long sz = ftell(f);
    if (sz < 20) {
        fclose(f);
        return fail(ATMO_ERR_IO, "%s: too short for a trailer", path);
    }
    fseek(f, sz - 20, SEEK_SET);
    size_t got = fread(header, sizeof(int32_t), 5, f);
    fclose(f);
    if (got != 5) return fail(ATMO_ERR_IO, "%s: cannot read trailer", path);
    if ((uint32_t)header[0] != 0xCAFEBABEu) return fail(ATMO_ERR_IO, "%s: bad magic", path);
    if (header[1] <= 0 || header[2] <= 0 || header[3] < 0 || header[4] <= 0)
        return fail(ATMO_ERR_IO, "%s: bad trailer", path);
    long long n = (long long)header[1] * header[2] * header[4];
    if (header[3] > 0 && header[2] % header[3] != 0) return fail(ATMO_ERR_IO, "%s: bad depth", path);
    if ((long long)sz != n * 4 + 20) return fail(ATMO_ERR_IO, "%s: size %ld does not match trailer", path, sz);
    return ATMO_OK;
}

atmo_status atmo_read_raw(const char *path, float *data, size_t capacity, int32_t header[5]) {
    atmo_status st = atmo_read_raw_header(path, header);
    if (st != ATMO_OK) return st;
    size_t n = (size_t)header[1] * header[2] * header[4];
    if (!data || capacity < n) return fail(ATMO_ERR_INVALID_ARGUMENT, "buffer too small (%zu < %zu)", capacity, n);
    FILE *f = fopen(path, "rb");
    if (!f) return fail(ATMO_ERR_IO, "cannot open %s", path);
    size_t got = fread(data, sizeof(float), n, f);
    fclose(f);
    if (got != n) return fail(ATMO_ERR_IO, "%s: short read", path);
    return ATMO_OK;
}

// ---------------------------------------------------------------------------
// preprocessAtmo (PreprocessAtmo.cpp:424-436)
// ---------------------------------------------------------------------------
atmo_status atmo_preprocess_ex(const atmo_params *params, const atmo_options *opts, const char *outdir,
                               int skip_if_exists) {
    if (!params || !outdir) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    std::string dir(outdir);
    std::string pS = dir + "/inscatter.raw", pT = dir + "/transmittance.raw", pE = dir + "/irradiance.raw";
    if (skip_if_exists) {
        int32_t h[5];
        if (atmo_read_raw_header(pS.c_str(), h) == ATMO_OK) {
            g_err.clear();
            return ATMO_OK;  // cache hit, as the reference (inscatter.raw is written last)
        }
    }
    atmo_status st = validate(*params, 0);
    if (st != ATMO_OK) return st;
    SetSizes z = sizes(*params);
    std::vector<float> T(z.nT * 3), E(z.nE * 3), S(z.n4 * 4);
    atmo_options o = opts ? *opts : default_options();
    o.flags |= ATMO_FLAG_CHECK_NUMERICS;
    if ((st = atmo_precompute_host(params, &o, T.data(), E.data(), S.data())) != ATMO_OK) return st;
    const atmo_params &p = *params;
    if ((st = atmo_write_raw(pT.c_str(), T.data(), p.TRANSMITTANCE_W, p.TRANSMITTANCE_H, 0, 3)) != ATMO_OK)
        return st;
    if ((st = atmo_write_raw(pE.c_str(), E.data(), p.SKY_W, p.SKY_H, 0, 3)) != ATMO_OK) return st;
    if ((st = atmo_write_raw(pS.c_str(), S.data(), p.RES_MU_S * p.RES_NU, p.RES_MU, p.RES_R, 4)) != ATMO_OK)
        return st;
    g_err.clear();
    return ATMO_OK;
}

atmo_status atmo_preprocess(const atmo_params *params, const char *outdir) {
    return atmo_preprocess_ex(params, nullptr, outdir, 1);
}

}  // extern "C"
