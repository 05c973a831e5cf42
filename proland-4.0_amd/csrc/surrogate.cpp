// surrogate.cpp — host side of the learned multiple-scattering surrogate
// (include/atmo_surrogate.h): parameters and Adam state, dataset assembly from
// plain-method tables, training steps, inference and the judge/fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/atmo_surrogate.h"
#include "surrogate_kernels.h"

using namespace atmo::sg;

extern "C" const char *atmo_last_error(void);
// error reporting shared with atmo_engine.cpp
namespace atmo {
atmo_status set_error(atmo_status st, const char *msg);
}

namespace {

atmo_status fail(atmo_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return atmo::set_error(st, buf);
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(e_ == hipErrorOutOfMemory ? ATMO_ERR_OUT_OF_MEMORY : ATMO_ERR_DEVICE,     \
                        "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);      \
    } while (0)
#define STCHK(expr)                        \
    do {                                   \
        atmo_status s_ = (expr);           \
        if (s_ != ATMO_OK) return s_;      \
    } while (0)

// parameter layout
constexpr size_t OFF_W1 = 0, OFF_B1 = kNIN * kNH;
constexpr size_t OFF_L0 = OFF_B1 + kNH;                 // W2
constexpr size_t LAYER_SZ = (size_t)kNH * kNH + kNH;
constexpr size_t OFF_W5 = OFF_L0 + kNG * LAYER_SZ;   // the output layer
constexpr size_t OFF_B5 = OFF_W5 + (size_t)kNH * kNOUT;
constexpr size_t NPARAM = OFF_B5 + kNOUT;               // 136452 for kNL = 3
constexpr size_t offW(int L) { return OFF_L0 + (size_t)(L - 2) * LAYER_SZ; }   // L = 2..kNL
constexpr size_t offB(int L) { return offW(L) + (size_t)kNH * kNH; }

struct Buf {
    void *p = nullptr;
    size_t n = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 256));
        if (e == hipSuccess) n = bytes;
        return e;
    }
    template <class T> T *as() const { return (T *)p; }
};

uint64_t splitmix(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

FeatSet feat_set(const atmo_params &p) {
    FeatSet f;
    f.RR = p.RES_R; f.RMU = p.RES_MU; f.RMUS = p.RES_MU_S; f.RNU = p.RES_NU;
    f.SW = p.SKY_W; f.SH = p.SKY_H;
    f.n4 = p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU;
    f.pad = 0;
    // centred and scaled over the ranges of the batch distribution
    f.f[0] = (p.Rg - 5000.0f) / 2000.0f;
    f.f[1] = ((p.Rt - p.Rg) - 80.0f) / 40.0f;
    f.f[2] = (p.HR - 8.5f) / 3.5f;
    f.f[3] = (p.betaR.x / 5.8e-3f - 1.65f) / 1.35f;
    f.f[4] = (p.HM - 1.5f) / 0.9f;
    f.f[5] = (float)((std::log((double)std::max(p.betaMSca.x, 1e-12f)) + 5.2072) / 1.7006);
    f.f[6] = (p.mieG - 0.75f) / 0.15f;
    f.Rg = p.Rg;
    f.Rt = p.Rt;
    return f;
}

struct Tabs {  // device tables of one set in .raw layouts
    float *T, *E, *S;
};

}  // namespace

struct atmo_surrogate {
    int device = 0;
    hipStream_t st = nullptr;
    Buf p, g, m, v, wb, wt;   // FP32 params/grads/Adam; bf16 W and W^T of layers 2..kNL
    Buf fz;                   // fused-inference weights (prep_fused)
    Buf iset;                 // InferSet array of the last fused launch
    TargetNorm tn{{0, 0, 0}, {1, 1, 1}, 1e-6f};
    int adam_t = 0;
    int exact_orders = 1;     // orders computed by the plain passes; the network predicts the rest
    Buf Xa, Ta;
    size_t nrows = 0;
    // work buffers: X, T, H, Y for `cap` rows; Z, dZ, dY for `zcap` rows
    int cap = 0, zcap = 0;
    Buf X, T, Z, H, dZa, dZb, dZ1, Y, dY, part, loss;
    atmo_engine *e1 = nullptr, *ej = nullptr, *ef = nullptr;
    ~atmo_surrogate() {
        if (e1) atmo_engine_destroy(e1);
        if (ej) atmo_engine_destroy(ej);
        if (ef) atmo_engine_destroy(ef);
        if (st) (void)hipStreamDestroy(st);
    }
    __bf16 *Wb(int L) { return wb.as<__bf16>() + (size_t)(L - 2) * kNH * kNH; }
    __bf16 *Wt(int L) { return wt.as<__bf16>() + (size_t)(L - 2) * kNH * kNH; }
    float *Zl(int i) { return Z.as<float>() + (size_t)i * zcap * kNH; }      // i = 0..kNL-1: Z1..Z_kNL
    __bf16 *Hl(int i) { return H.as<__bf16>() + (size_t)i * cap * kNH; }     // H1..H_kNL
};

namespace {

atmo_status ensure_cap(atmo_surrogate *s, int rows, bool train) {
    if (rows > s->cap) {
        const int c = rows;
        HIPCHK(s->X.ensure((size_t)c * kXS * 4));
        HIPCHK(s->T.ensure((size_t)c * kNOUT * 4));
        HIPCHK(s->H.ensure((size_t)kNL * c * kNH * 2));
        HIPCHK(s->Y.ensure((size_t)c * kNOUT * 4));
        s->cap = c;
    }
    if (train && rows > s->zcap) {
        const int c = rows;
        HIPCHK(s->Z.ensure((size_t)kNL * c * kNH * 4));
        HIPCHK(s->dZa.ensure((size_t)c * kNH * 2));
        HIPCHK(s->dZb.ensure((size_t)c * kNH * 2));
        HIPCHK(s->dZ1.ensure((size_t)c * kNH * 4));
        HIPCHK(s->dY.ensure((size_t)c * kNOUT * 4));
        s->zcap = c;
    }
    return ATMO_OK;
}

// fused-inference weight buffer: a1h, a1l [256][16] (bf16), a5h, a5l [4][256] (FP16), hidden
// biases [kNG][256] (FP32), hidden weights [kNG][256][256] as [out][in] (FP16)
constexpr size_t FZ_A1H = 0, FZ_A1L = FZ_A1H + kNH * 16 * 2, FZ_A5H = FZ_A1L + kNH * 16 * 2,
                 FZ_A5L = FZ_A5H + kNOUT * kNH * 2, FZ_BH = FZ_A5L + kNOUT * kNH * 2,
                 FZ_W16 = FZ_BH + kNG * kNH * 4, FZ_BYTES = FZ_W16 + (size_t)kNG * kNH * kNH * 2;

FusedWeights fused_weights(atmo_surrogate *s) {
    char *z = (char *)s->fz.p;
    const float *P = s->p.as<float>();
    return FusedWeights{(const __bf16 *)(z + FZ_A1H), (const __bf16 *)(z + FZ_A1L), (const _Float16 *)(z + FZ_W16),
                        (const float *)(z + FZ_BH), (const _Float16 *)(z + FZ_A5H),
                        (const _Float16 *)(z + FZ_A5L), P + OFF_B5};
}

atmo_status refresh_bf16(atmo_surrogate *s) {
    for (int L = 2; L <= kNL; ++L)
        HIPCHK(to_bf16(s->p.as<float>() + offW(L), s->Wb(L), s->Wt(L), s->st));
    const float *P = s->p.as<float>();
    char *z = (char *)s->fz.p;
    HIPCHK(prep_fused(P + OFF_W1, P + OFF_B1, P + OFF_W5, P + offW(2), P + offB(2), LAYER_SZ,
                      (__bf16 *)(z + FZ_A1H), (__bf16 *)(z + FZ_A1L), (_Float16 *)(z + FZ_W16),
                      (_Float16 *)(z + FZ_A5H), (_Float16 *)(z + FZ_A5L), (float *)(z + FZ_BH), s->st));
    return ATMO_OK;
}

// Forward over `rows` rows of s->X (rows % 128 == 0).  train: keeps Z, and
// computes the loss against s->T into loss_slot and dY.
atmo_status forward(atmo_surrogate *s, int rows, bool train, float *loss_slot, float *Yout) {
    const float *P = s->p.as<float>();
    HIPCHK(layer1_fwd(s->X.as<float>(), P + OFF_W1, P + OFF_B1, rows, train ? s->Zl(0) : nullptr,
                      s->Hl(0), s->st));
    for (int L = 2; L <= kNL; ++L) {
        GemmEpi ep{P + offB(L), nullptr, train ? s->Zl(L - 1) : nullptr, s->Hl(L - 1)};
        HIPCHK(gemm(false, true, train ? EPI_FWD : EPI_FWD_INF, s->Hl(L - 2), s->Wt(L), rows, kNH, kNH,
                    1, ep, s->st));
    }
    HIPCHK(layer5(s->Hl(kNL - 1), P + OFF_W5, P + OFF_B5, s->T.as<float>(), rows, train ? 0 : 1,
                  Yout, s->dY.as<float>(), loss_slot, s->st));
    return ATMO_OK;
}

atmo_status backward(atmo_surrogate *s, int rows) {
    const float *P = s->p.as<float>();
    float *G = s->g.as<float>();
    HIPCHK(hipMemsetAsync(G, 0, NPARAM * 4, s->st));
    __bf16 *cur = s->dZa.as<__bf16>(), *nxt = s->dZb.as<__bf16>();
    HIPCHK(layer5_bwd(s->Hl(kNL - 1), s->Zl(kNL - 1), P + OFF_W5, s->dY.as<float>(), rows, G + OFF_W5, G + OFF_B5,
                      cur, s->st));
    const int nsplit = std::max(1, std::min(64, rows / 1024));
    const int ns = (rows / nsplit) % 32 == 0 ? nsplit : 1;
    HIPCHK(s->part.ensure((size_t)ns * kNH * kNH * 4));
    for (int L = kNL; L >= 2; --L) {
        GemmEpi ew{nullptr, nullptr, s->part.as<float>(), nullptr};
        HIPCHK(gemm(true, false, EPI_PARTIAL, s->Hl(L - 2), cur, kNH, kNH, rows, ns, ew, s->st));
        HIPCHK(reduce_partials(s->part.as<float>(), ns, kNH * kNH, G + offW(L), s->st));
        HIPCHK(colsum(cur, rows, G + offB(L), s->st));
        GemmEpi ex{nullptr, s->Zl(L - 2), L == 2 ? s->dZ1.as<float>() : nullptr, nxt};
        HIPCHK(gemm(false, true, EPI_DX, cur, s->Wb(L), rows, kNH, kNH, 1, ex, s->st));
        std::swap(cur, nxt);
    }
    HIPCHK(layer1_bwd(s->X.as<float>(), s->dZ1.as<float>(), rows, G + OFF_W1, G + OFF_B1, s->st));
    return ATMO_OK;
}

atmo_status adam_step(atmo_surrogate *s, float lr) {
    s->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)s->adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)s->adam_t));
    HIPCHK(adam(s->p.as<float>(), s->g.as<float>(), s->m.as<float>(), s->v.as<float>(), (int)NPARAM, lr,
                b1, b2, 1e-8f, c1, c2, s->st));
    return refresh_bf16(s);
}

size_t set_bytes(const atmo_params &p) {
    return ((size_t)p.TRANSMITTANCE_W * p.TRANSMITTANCE_H * 3 + 64) * 4 + ((size_t)p.SKY_W * p.SKY_H * 3 + 64) * 4 +
           (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU * 16 + 256;
}

// Carves one set's .raw tables out of a buffer (16-byte aligned inscatter).
Tabs carve(char *&q, const atmo_params &p) {
    Tabs t;
    auto al = [](char *x) { return (char *)(((uintptr_t)x + 255) & ~(uintptr_t)255); };
    q = al(q);
    t.S = (float *)q; q += (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU * 16;
    q = al(q);
    t.T = (float *)q; q += (size_t)p.TRANSMITTANCE_W * p.TRANSMITTANCE_H * 12;
    q = al(q);
    t.E = (float *)q; q += (size_t)p.SKY_W * p.SKY_H * 12;
    return t;
}

}  // namespace

extern "C" {

size_t atmo_surrogate_param_count(void) { return NPARAM; }

atmo_status atmo_surrogate_create(atmo_surrogate **out, int device, uint64_t seed) {
    if (!out) return fail(ATMO_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    std::unique_ptr<atmo_surrogate> s(new atmo_surrogate);
    s->device = device;
    STCHK(atmo_engine_create(&s->e1, device));
    STCHK(atmo_engine_create(&s->ej, device));
    STCHK(atmo_engine_create(&s->ef, device));
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    HIPCHK(s->p.ensure(NPARAM * 4));
    HIPCHK(s->g.ensure(NPARAM * 4));
    HIPCHK(s->m.ensure(NPARAM * 4));
    HIPCHK(s->v.ensure(NPARAM * 4));
    HIPCHK(s->wb.ensure((size_t)kNG * kNH * kNH * 2));
    HIPCHK(s->wt.ensure((size_t)kNG * kNH * kNH * 2));
    HIPCHK(s->fz.ensure(FZ_BYTES));
    HIPCHK(s->loss.ensure(4));
    // Xavier-uniform weights, zero biases
    std::vector<float> h(NPARAM, 0.0f);
    uint64_t rs = seed;
    auto fill = [&](size_t off, int fin, int fout) {
        const double a = std::sqrt(6.0 / (fin + fout));
        for (size_t i = 0; i < (size_t)fin * fout; ++i)
            h[off + i] = (float)(a * (2.0 * ((splitmix(rs) >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
    };
    fill(OFF_W1, kNIN, kNH);
    for (int L = 2; L <= kNL; ++L) fill(offW(L), kNH, kNH);
    fill(OFF_W5, kNH, kNOUT);
    HIPCHK(hipMemcpy(s->p.p, h.data(), NPARAM * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->m.p, 0, NPARAM * 4));
    HIPCHK(hipMemset(s->v.p, 0, NPARAM * 4));
    STCHK(refresh_bf16(s.get()));
    HIPCHK(hipStreamSynchronize(s->st));
    *out = s.release();
    return ATMO_OK;
}

atmo_status atmo_surrogate_destroy(atmo_surrogate *s) {
    if (s) {
        (void)hipSetDevice(s->device);
        if (s->st) (void)hipStreamSynchronize(s->st);
        delete s;
    }
    return ATMO_OK;
}

atmo_status atmo_surrogate_get_params(atmo_surrogate *s, float *host, size_t n) {
    if (!s || !host || n < NPARAM) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(host, s->p.p, NPARAM * 4, hipMemcpyDeviceToHost));
    return ATMO_OK;
}

atmo_status atmo_surrogate_set_params(atmo_surrogate *s, const float *host, size_t n) {
    if (!s || !host || n < NPARAM) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(s->p.p, host, NPARAM * 4, hipMemcpyHostToDevice));
    STCHK(refresh_bf16(s));
    HIPCHK(hipStreamSynchronize(s->st));
    return ATMO_OK;
}

atmo_status atmo_surrogate_get_norm(atmo_surrogate *s, float mu[3], float sd[3], float *eps) {
    if (!s || !mu || !sd || !eps) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    for (int c = 0; c < 3; ++c) {
        mu[c] = s->tn.mu[c];
        sd[c] = s->tn.sd[c];
    }
    *eps = s->tn.eps;
    return ATMO_OK;
}

atmo_status atmo_surrogate_set_norm(atmo_surrogate *s, const float mu[3], const float sd[3], float eps) {
    if (!s || !mu || !sd || !(eps > 0.f)) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    for (int c = 0; c < 3; ++c) {
        if (!(sd[c] > 0.f)) return fail(ATMO_ERR_INVALID_ARGUMENT, "sd must be > 0");
        s->tn.mu[c] = mu[c];
        s->tn.sd[c] = sd[c];
    }
    s->tn.eps = eps;
    return ATMO_OK;
}

atmo_status atmo_surrogate_set_exact_orders(atmo_surrogate *s, int orders) {
    if (!s || orders < 1 || orders > 2) return fail(ATMO_ERR_INVALID_ARGUMENT, "exact orders must be 1 or 2");
    s->exact_orders = orders;
    return ATMO_OK;
}

atmo_status atmo_surrogate_features(const atmo_params *p, const float *S_exact, const float *E_exact,
                                    const uint32_t *codes, int n, float *X) {
    if (!p || !S_exact || !E_exact || !codes || !X || n <= 0) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    HIPCHK(code_rows(feat_set(*p), codes, n, (const float4 *)S_exact, E_exact, X, nullptr));
    HIPCHK(hipDeviceSynchronize());
    return ATMO_OK;
}

atmo_status atmo_surrogate_build_dataset(atmo_surrogate *s, const atmo_params *sets, int nsets,
                                         const atmo_options *opts, int samples_per_set, uint64_t seed,
                                         size_t *rows_out) {
    if (!s || !sets || nsets <= 0 || !opts || samples_per_set <= 0)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    if (opts->orders <= s->exact_orders)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "training data needs more orders than the %d computed exactly",
                    s->exact_orders);
    if (opts->flags & (ATMO_FLAG_HALF_STORAGE | ATMO_FLAG_LINEAR_TRANSMITTANCE | ATMO_FLAG_LINEAR_INSCATTER))
        return fail(ATMO_ERR_INVALID_ARGUMENT,
                    "the surrogate works on FP32 tables of the non-linear parameterisation "
                    "(no ATMO_FLAG_HALF_STORAGE, ATMO_FLAG_LINEAR_*)");
    HIPCHK(hipSetDevice(s->device));
    size_t rows = 0;
    for (int i = 0; i < nsets; ++i) rows += (size_t)samples_per_set + (size_t)sets[i].SKY_W * sets[i].SKY_H;
    HIPCHK(s->Xa.ensure(rows * kXS * 4));
    HIPCHK(s->Ta.ensure(rows * kNOUT * 4));
    atmo_options on = *opts, o1 = *opts;
    // every chunk of seeded sets is a new batch shape: no graph capture, no
    // per-shape autotuning (the default kernel variants, whose tables equal the
    // tuned variants' bit for bit)
    on.flags |= ATMO_FLAG_NO_GRAPH | ATMO_FLAG_NO_AUTOTUNE;
    o1.orders = s->exact_orders;
    o1.flags |= ATMO_FLAG_NO_GRAPH | ATMO_FLAG_NO_AUTOTUNE;
    const int CH = 16;
    TargetNorm raw{{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}, s->tn.eps};
    size_t row = 0;
    Buf bn, b1;
    for (int c0 = 0; c0 < nsets; c0 += CH) {
        const int nc = std::min(CH, nsets - c0);
        size_t bytes = 0;
        for (int i = 0; i < nc; ++i) bytes += set_bytes(sets[c0 + i]) + 1024;
        HIPCHK(bn.ensure(bytes));
        HIPCHK(b1.ensure(bytes));
        std::vector<atmo_tables> tn(nc), t1(nc);
        char *qn = bn.as<char>(), *q1 = b1.as<char>();
        for (int i = 0; i < nc; ++i) {
            Tabs a = carve(qn, sets[c0 + i]), b = carve(q1, sets[c0 + i]);
            tn[i] = {a.T, a.E, a.S};
            t1[i] = {b.T, b.E, b.S};
        }
        STCHK(atmo_engine_precompute(s->ej, sets + c0, nc, &on, tn.data(), s->st));
        STCHK(atmo_engine_precompute(s->e1, sets + c0, nc, &o1, t1.data(), s->st));
        for (int i = 0; i < nc; ++i) {
            const atmo_params &p = sets[c0 + i];
            HIPCHK(make_rows(feat_set(p), (const float4 *)tn[i].inscatter, (const float4 *)t1[i].inscatter,
                             tn[i].irradiance, t1[i].irradiance, samples_per_set, seed + (uint64_t)(c0 + i) * 7919u,
                             s->Xa.as<float>() + row * kXS, s->Ta.as<float>() + row * kNOUT, raw, s->st));
            row += (size_t)samples_per_set + (size_t)p.SKY_W * p.SKY_H;
        }
        HIPCHK(hipStreamSynchronize(s->st));
    }
    s->nrows = row;
    // standardise: mean and deviation of every target row (GPU reduction in FP64)
    Buf stats;
    HIPCHK(stats.ensure(6 * sizeof(double)));
    HIPCHK(hipMemsetAsync(stats.p, 0, 6 * sizeof(double), s->st));
    HIPCHK(target_stats(s->Ta.as<float>(), row, (double *)stats.p, s->st));
    double hs[6];
    HIPCHK(hipMemcpyAsync(hs, stats.p, sizeof hs, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (int c = 0; c < 3; ++c) {
        const double mu = hs[c] / (double)row, var = std::max(hs[3 + c] / (double)row - mu * mu, 1e-6);
        s->tn.mu[c] = (float)mu;
        s->tn.sd[c] = (float)std::sqrt(var);
    }
    HIPCHK(norm_targets(s->Ta.as<float>(), row, s->tn, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (rows_out) *rows_out = row;
    return ATMO_OK;
}

atmo_status atmo_surrogate_train(atmo_surrogate *s, int steps, int batch, float lr, float lr_end,
                                 uint64_t seed, float *losses) {
    if (!s || steps <= 0 || batch <= 0 || batch % 128) return fail(ATMO_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 128");
    if (!s->nrows) return fail(ATMO_ERR_INVALID_ARGUMENT, "no dataset (atmo_surrogate_build_dataset)");
    HIPCHK(hipSetDevice(s->device));
    STCHK(ensure_cap(s, batch, true));
    Buf lbuf;
    HIPCHK(lbuf.ensure((size_t)steps * 4));
    HIPCHK(hipMemsetAsync(lbuf.p, 0, (size_t)steps * 4, s->st));
    for (int k = 0; k < steps; ++k) {
        const double c = 0.5 * (1.0 + std::cos(M_PI * k / std::max(1, steps - 1)));
        const float lrk = (float)(lr_end + (lr - lr_end) * c);
        HIPCHK(gather(s->Xa.as<float>(), s->Ta.as<float>(), s->nrows, batch, seed * 1000003ull + k,
                      s->X.as<float>(), s->T.as<float>(), s->st));
        STCHK(forward(s, batch, true, lbuf.as<float>() + k, nullptr));
        STCHK(backward(s, batch));
        STCHK(adam_step(s, lrk));
    }
    HIPCHK(hipStreamSynchronize(s->st));
    if (losses) HIPCHK(hipMemcpy(losses, lbuf.p, (size_t)steps * 4, hipMemcpyDeviceToHost));
    return ATMO_OK;
}

atmo_status atmo_surrogate_loss_grad(atmo_surrogate *s, const float *X, const float *T, int B, float *loss,
                                     float *grad) {
    if (!s || !X || !T || B <= 0 || B % 128 || !loss || !grad)
        return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments (B must be a multiple of 128)");
    HIPCHK(hipSetDevice(s->device));
    STCHK(ensure_cap(s, B, true));
    HIPCHK(hipMemcpyAsync(s->X.p, X, (size_t)B * kXS * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->T.p, T, (size_t)B * kNOUT * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemsetAsync(s->loss.p, 0, 4, s->st));
    STCHK(forward(s, B, true, s->loss.as<float>(), nullptr));
    STCHK(backward(s, B));
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(loss, s->loss.p, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(grad, s->g.p, NPARAM * 4, hipMemcpyDeviceToHost));
    return ATMO_OK;
}

atmo_status atmo_surrogate_predict(atmo_surrogate *s, const float *X, int B, float *Y) {
    if (!s || !X || !Y || B <= 0) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    HIPCHK(hipSetDevice(s->device));
    const int Bp = (B + 127) / 128 * 128;
    STCHK(ensure_cap(s, Bp, false));
    HIPCHK(hipMemsetAsync(s->X.p, 0, (size_t)Bp * kXS * 4, s->st));
    HIPCHK(hipMemcpyAsync(s->X.p, X, (size_t)B * kXS * 4, hipMemcpyDeviceToDevice, s->st));
    STCHK(forward(s, Bp, false, nullptr, s->Y.as<float>()));
    HIPCHK(hipMemcpyAsync(Y, s->Y.p, (size_t)B * kNOUT * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return ATMO_OK;
}

atmo_status atmo_surrogate_precompute(atmo_surrogate *s, const atmo_params *sets, int nsets,
                                      const atmo_options *opts, const atmo_tables *out, float judge_fraction,
                                      float tolerance, uint64_t seed, atmo_surrogate_report *rep,
                                      void *stream) {
    if (!s || !sets || nsets <= 0 || !out) return fail(ATMO_ERR_INVALID_ARGUMENT, "bad arguments");
    if (!(judge_fraction >= 0.f && judge_fraction <= 1.f) || !(tolerance > 0.f))
        return fail(ATMO_ERR_INVALID_ARGUMENT, "judge_fraction in [0, 1], tolerance > 0");
    HIPCHK(hipSetDevice(s->device));
    atmo_options o = opts ? *opts : atmo_options{};
    if (!opts) atmo_options_default(&o);
    if (o.flags & (ATMO_FLAG_HALF_STORAGE | ATMO_FLAG_LINEAR_TRANSMITTANCE | ATMO_FLAG_LINEAR_INSCATTER))
        return fail(ATMO_ERR_INVALID_ARGUMENT,
                    "the surrogate works on FP32 tables of the non-linear parameterisation "
                    "(no ATMO_FLAG_HALF_STORAGE, ATMO_FLAG_LINEAR_*)");
    atmo_options o1 = o;
    o1.orders = std::min(s->exact_orders, o.orders);
    hipStream_t hs = (hipStream_t)stream;
    atmo_surrogate_report r;
    memset(&r, 0, sizeof r);
    r.nsets = nsets;
    hipEvent_t ev[5];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 5; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    // (1) exact transmittance and single scattering straight into the outputs
    HIPCHK(hipStreamSynchronize(hs));
    HIPCHK(hipEventRecord(ev[0], s->st));
    STCHK(atmo_engine_precompute(s->e1, sets, nsets, &o1, out, s->st));
    HIPCHK(hipEventRecord(ev[1], s->st));
    // (2) orders >= 2 from the network: one fused launch for the batch
    // (features, five layers and the combination with the exact part)
    {
        std::vector<InferSet> hs(nsets);
        const int rpb = fused_rows_per_block();
        int tiles = 0;
        for (int i = 0; i < nsets; ++i) {
            hs[i].fs = feat_set(sets[i]);
            hs[i].tile0 = tiles;
            hs[i].S = (float4 *)out[i].inscatter;
            hs[i].E = out[i].irradiance;
            tiles += (hs[i].fs.n4 + hs[i].fs.SW * hs[i].fs.SH + rpb - 1) / rpb;
        }
        HIPCHK(s->iset.ensure(hs.size() * sizeof(InferSet)));
        HIPCHK(hipMemcpyAsync(s->iset.p, hs.data(), hs.size() * sizeof(InferSet), hipMemcpyHostToDevice, s->st));
        HIPCHK(mlp_infer((const InferSet *)s->iset.p, nsets, tiles, fused_weights(s), s->tn, s->st));
        HIPCHK(hipStreamSynchronize(s->st));  // hs must outlive the copy
    }
    HIPCHK(hipEventRecord(ev[2], s->st));
    // (3) judge: plain recompute of ceil(fraction * nsets) seeded sets
    int nj = judge_fraction > 0.f ? std::max(1, (int)std::ceil(judge_fraction * nsets)) : 0;
    nj = std::min(nj, nsets);
    std::vector<int> idx(nsets);
    for (int i = 0; i < nsets; ++i) idx[i] = i;
    uint64_t rs = seed;
    for (int i = 0; i < nj; ++i) std::swap(idx[i], idx[i + (int)(splitmix(rs) % (uint64_t)(nsets - i))]);
    idx.resize(nj);
    std::sort(idx.begin(), idx.end());
    float worstS = 0.f, worstE = 0.f;
    Buf jb, flags;
    std::vector<atmo_params> js;
    std::vector<atmo_tables> jt;
    if (nj > 0) {
        size_t bytes = 0;
        for (int i : idx) bytes += set_bytes(sets[i]) + 1024;
        HIPCHK(jb.ensure(bytes));
        HIPCHK(flags.ensure((size_t)nj * 4 * 4));
        HIPCHK(hipMemsetAsync(flags.p, 0, (size_t)nj * 16, s->st));
        char *q = jb.as<char>();
        for (int i : idx) {
            Tabs t = carve(q, sets[i]);
            js.push_back(sets[i]);
            jt.push_back({t.T, t.E, t.S});
        }
        STCHK(atmo_engine_precompute(s->ej, js.data(), nj, &o, jt.data(), s->st));
        unsigned *fl = flags.as<unsigned>();
        for (int k = 0; k < nj; ++k) {
            const atmo_params &p = js[k];
            const size_t n4 = (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU, nE = (size_t)p.SKY_W * p.SKY_H;
            HIPCHK(absmax(jt[k].inscatter, n4, 4, 4, fl + 4 * k + 0, s->st));
            HIPCHK(absmax(jt[k].irradiance, nE, 3, 3, fl + 4 * k + 1, s->st));
        }
        std::vector<unsigned> hf((size_t)nj * 4);
        HIPCHK(hipMemcpyAsync(hf.data(), flags.p, hf.size() * 4, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        for (int k = 0; k < nj; ++k) {
            const atmo_params &p = js[k];
            const int i = idx[k];
            const size_t n4 = (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU, nE = (size_t)p.SKY_W * p.SKY_H;
            float mS, mE;
            memcpy(&mS, &hf[4 * k], 4);
            memcpy(&mE, &hf[4 * k + 1], 4);
            HIPCHK(relerr(out[i].inscatter, jt[k].inscatter, n4, 4, 4, 1e-4f * mS, fl + 4 * k + 2, s->st));
            HIPCHK(relerr(out[i].irradiance, jt[k].irradiance, nE, 3, 3, std::max(1e-4f * mE, 1e-30f),
                          fl + 4 * k + 3, s->st));
        }
        HIPCHK(hipMemcpyAsync(hf.data(), flags.p, hf.size() * 4, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        for (int k = 0; k < nj; ++k) {
            float eS, eE;
            memcpy(&eS, &hf[4 * k + 2], 4);
            memcpy(&eE, &hf[4 * k + 3], 4);
            worstS = std::max(worstS, eS);
            worstE = std::max(worstE, eE);
        }
    }
    HIPCHK(hipEventRecord(ev[3], s->st));
    r.judged = nj;
    r.max_rel_err_S = worstS;
    r.max_rel_err_E = worstE;
    // no judged set (judge_fraction = 0): the network is trusted as is
    r.accepted = nj == 0 || (worstS <= tolerance && worstE <= tolerance);
    if (r.accepted) {
        // judged sets keep their exact plain tables
        for (int k = 0; k < nj; ++k) {
            const atmo_params &p = js[k];
            const int i = idx[k];
            HIPCHK(hipMemcpyAsync(out[i].inscatter, jt[k].inscatter,
                                  (size_t)p.RES_R * p.RES_MU * p.RES_MU_S * p.RES_NU * 16, hipMemcpyDeviceToDevice, s->st));
            HIPCHK(hipMemcpyAsync(out[i].irradiance, jt[k].irradiance, (size_t)p.SKY_W * p.SKY_H * 12,
                                  hipMemcpyDeviceToDevice, s->st));
        }
        r.fallback_sets = nj;
    } else {
        STCHK(atmo_engine_precompute(s->ef, sets, nsets, &o, out, s->st));
        r.fallback_sets = nsets;
    }
    HIPCHK(hipEventRecord(ev[4], s->st));
    HIPCHK(hipEventSynchronize(ev[4]));
    HIPCHK(hipEventElapsedTime(&r.ms_exact, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&r.ms_mlp, ev[1], ev[2]));
    HIPCHK(hipEventElapsedTime(&r.ms_judge, ev[2], ev[3]));
    HIPCHK(hipEventElapsedTime(&r.ms_fallback, ev[3], ev[4]));
    if (rep) *rep = r;
    return ATMO_OK;
}

}  // extern "C"
