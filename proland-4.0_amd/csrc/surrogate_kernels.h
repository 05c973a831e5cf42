// surrogate_kernels.h — launchers of the surrogate MLP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace atmo {
namespace sg {

constexpr int kNIN = 14;   // 4 texel coordinates + 7 atmosphere features + 3 exact-part values
constexpr int kXS = 16;    // row stride of the input features (floats)
constexpr int kNH = 256;   // hidden width
constexpr int kNL = 3;     // hidden layers (H1 .. H_kNL)
constexpr int kNG = kNL - 1;   // hidden 256x256 layers (layers 2 .. kNL, on the matrix cores)
constexpr int kNOUT = 4;   // outputs (rgb + an always-zero alpha)

enum { EPI_FWD = 0, EPI_FWD_INF = 1, EPI_DX = 2, EPI_PARTIAL = 3 };

struct GemmEpi {
    const float *bias;   // EPI_FWD*: bias[N]
    const float *zin;    // EPI_DX: pre-activations of the layer below
    float *zout;         // EPI_FWD: pre-activations; EPI_DX: optional FP32 copy; EPI_PARTIAL: partials
    __bf16 *hout;        // EPI_FWD*: activations; EPI_DX: gradients
};

struct FeatSet {
    int RR, RMU, RMUS, RNU, SW, SH, n4, pad;
    float f[7];          // normalised atmosphere features
    float Rg, Rt;        // planet and atmosphere radii (the mu-row coordinate's path lengths)
};

struct TargetNorm {
    float mu[3], sd[3], eps;
};

// Weights of the fused inference kernel (device pointers), from prep_fused.
struct FusedWeights {
    const __bf16 *a1h, *a1l;     // [256][16] layer 1 (column 14: bias), bf16 hi / lo
    const _Float16 *wt;          // [kNG][256][256] layers 2..kNL as [out][in], FP16
    const float *bh;             // [kNG][256] biases of layers 2..kNL
    const _Float16 *a5h, *a5l;   // [4][256] layer 5 as [out][in], FP16 hi / lo
    const float *b5;             // [4]
};

// One set of a fused inference launch: its features, its first workgroup
// (tiles of fused_rows_per_block() rows: S texels, then E texels) and outputs.
struct InferSet {
    FeatSet fs;
    int tile0, pad[3];
    float4 *S;
    float *E;
};

hipError_t gemm(bool ta, bool tb, int epi, const __bf16 *A, const __bf16 *B, int M, int N, int K,
                int nsplit, const GemmEpi &ep, hipStream_t st);
hipError_t layer1_fwd(const float *X, const float *W1, const float *b1, int rows, float *Z1,
                      __bf16 *H1, hipStream_t st);
hipError_t layer5(const __bf16 *H4, const float *W5, const float *b5, const float *Tg, int rows,
                  int mode, float *Y, float *dY, float *loss, hipStream_t st);
hipError_t layer5_bwd(const __bf16 *H4, const float *Z4, const float *W5, const float *dY, int rows,
                      float *gW5, float *gb5, __bf16 *dZ4, hipStream_t st);
hipError_t layer1_bwd(const float *X, const float *dZ1, int rows, float *gW1, float *gb1,
                      hipStream_t st);
hipError_t colsum(const __bf16 *D, int rows, float *g, hipStream_t st);
hipError_t reduce_partials(const float *P, int nsplit, int n, float *g, hipStream_t st);
hipError_t adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                float eps, float c1, float c2, hipStream_t st);
hipError_t to_bf16(const float *W, __bf16 *Wb, __bf16 *Wt, hipStream_t st);
// S1, E1: the tables of the exactly computed orders (the targets are the rest)
hipError_t make_rows(const FeatSet &fs, const float4 *S4, const float4 *S1, const float *E, const float *E1, int ns,
                     unsigned long long seed, float *X, float *Tg, const TargetNorm &tn,
                     hipStream_t st);
// S1, E1: the exact tables of the set (the exact-part features)
hipError_t code_rows(const FeatSet &fs, const unsigned *codes, int n, const float4 *S1, const float *E1, float *X,
                     hipStream_t st);
// out[6] (device, zeroed by the caller): sums and sums of squares of the rgb targets
hipError_t target_stats(const float *Tg, size_t n, double *out, hipStream_t st);
hipError_t norm_targets(float *Tg, size_t n, const TargetNorm &tn, hipStream_t st);
// Wh0 / bh0: weight ([in][out]) and bias of layer 2, those of layers 3..kNL follow
// at stride bstride floats
hipError_t prep_fused(const float *W1, const float *b1, const float *W5, const float *Wh0, const float *bh0,
                      size_t bstride, __bf16 *a1h, __bf16 *a1l, _Float16 *wt, _Float16 *a5h, _Float16 *a5l,
                      float *bh, hipStream_t st);
// Fused inference of every S and E texel of a batch of sets (device array
// `sets`, tile0 ascending, ntiles workgroups in all): adds the predicted
// higher orders to S.rgb and E.
int fused_rows_per_block();
hipError_t mlp_infer(const InferSet *sets, int nsets, int ntiles, const FusedWeights &fw, const TargetNorm &tn,
                     hipStream_t st);
hipError_t gather(const float *Xa, const float *Ta, size_t n, int B, unsigned long long seed,
                  float *X, float *Tg, hipStream_t st);
hipError_t relerr(const float *a, const float *b, size_t n, int stride, int ncomp, float floorv,
                  unsigned *out, hipStream_t st);
hipError_t absmax(const float *a, size_t n, int stride, int ncomp, unsigned *out, hipStream_t st);

}  // namespace sg
}  // namespace atmo
