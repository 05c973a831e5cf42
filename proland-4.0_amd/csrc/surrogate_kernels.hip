// surrogate_kernels.hip — kernels of the learned multiple-scattering surrogate.
//
// MLP 14 -> 256 -> 256 -> 256 -> 4 (kNL = 3 hidden layers), act() on the hidden
// layers.  Layers are numbered as for four hidden layers: the output layer is
// "layer 5" and the last hidden layer's activations "H4" whatever kNL is.  The
// kNG = kNL - 1 256x256 layers (96 % of the FLOPs) run on the BF16 matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation) in one LDS-tiled GEMM kernel
// with transposable operands and fused epilogues; the thin first and last
// layers, the loss, the reductions and Adam are FP32 vector code.
#include "surrogate_kernels.h"

#include <mutex>

namespace atmo {
namespace sg {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Hidden activation: z gated by a smoothstep, act(z) = z s(t), t = clamp(z / 6
// + 1/2, 0, 1), s = t^2 (3 - 2 t).  Like SiLU (z sigmoid(z)) it is C1 with
// gate 1/2 and gate slope 1/4 at 0, but it is linear (z or 0) beyond |z| = 3 and
// takes 6 full-rate instructions where SiLU needs two quarter-rate
// transcendentals (exp, rcp) and three more: the activations of the fused
// inference kernel are a bound of its speed.  act'(z) = s + z t (1 - t).
__device__ __forceinline__ float gate_t(float z) {
    return __builtin_amdgcn_fmed3f(fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ float act(float z) {
    const float t = gate_t(z);
    return z * (t * t * fmaf(-2.0f, t, 3.0f));
}
__device__ __forceinline__ float dact(float z) {
    const float t = gate_t(z);
    return fmaf(z, t * (1.0f - t), t * t * fmaf(-2.0f, t, 3.0f));
}

// ---------------------------------------------------------------------------
// C[M x N] = op(A)[M x K] . op(B)[K x N], bf16 in, FP32 accumulate.
//   TA = false: A stored [M][K] (lda = K);  TA = true: A stored [K][M] (lda = M)
//   TB = true : B stored [N][K] (ldb = K);  TB = false: B stored [K][N] (ldb = N)
// Tile 128 x 128 x 32, 4 waves of 64 x 64 (2 x 2 MFMA 32x32 tiles).  blockIdx.z
// splits K into chunks of kchunk (EPI_PARTIAL writes one partial per chunk).
// M, N multiples of 128, K and kchunk multiples of 32 (checked by the host).
// ---------------------------------------------------------------------------
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const __bf16 *__restrict__ A,
                                              const __bf16 *__restrict__ Bm, int M, int N, int K,
                                              int kchunk, GemmEpi ep) {
    __shared__ __attribute__((aligned(16))) __bf16 As[128][40];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[128][40];
    const int n0 = blockIdx.x * 128, m0 = blockIdx.y * 128;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    for (int k0 = kb; k0 < ke; k0 += 32) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int c = tid + i * 256;
            if (!TA) {
                const int row = c >> 2, kc = (c & 3) * 8;
                const uint4 v = *(const uint4 *)(A + (size_t)(m0 + row) * K + k0 + kc);
                *(uint4 *)&As[row][kc] = v;
            } else {
                const int kr = c >> 4, mc = (c & 15) * 8;
                const uint4 v = *(const uint4 *)(A + (size_t)(k0 + kr) * M + m0 + mc);
                const __bf16 *pv = (const __bf16 *)&v;
#pragma unroll
                for (int j = 0; j < 8; ++j) As[mc + j][kr] = pv[j];
            }
            if (TB) {
                const int n = c >> 2, kc = (c & 3) * 8;
                const uint4 v = *(const uint4 *)(Bm + (size_t)(n0 + n) * K + k0 + kc);
                *(uint4 *)&Bs[n][kc] = v;
            } else {
                const int kr = c >> 4, nc = (c & 15) * 8;
                const uint4 v = *(const uint4 *)(Bm + (size_t)(k0 + kr) * N + n0 + nc);
                const __bf16 *pv = (const __bf16 *)&v;
#pragma unroll
                for (int j = 0; j < 8; ++j) Bs[nc + j][kr] = pv[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 16) {
            bf16x8 a[2], b[2];
            const int kk = ks + 8 * (lane >> 5);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = *(const bf16x8 *)&As[wm * 64 + t * 32 + (lane & 31)][kk];
                b[t] = *(const bf16x8 *)&Bs[wn * 64 + t * 32 + (lane & 31)][kk];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    // epilogue: lane holds rows 8(r/4) + 4(lane/32) + r%4 of column lane%32 of each 32x32 tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + wn * 64 + j * 32 + (lane & 31);
                const size_t o = (size_t)row * N + col;
                const float v = acc[i][j][r];
                if (EPI == EPI_FWD) {
                    const float z = v + ep.bias[col];
                    ep.zout[o] = z;
                    ep.hout[o] = (__bf16)act(z);
                } else if (EPI == EPI_FWD_INF) {
                    ep.hout[o] = (__bf16)act(v + ep.bias[col]);
                } else if (EPI == EPI_DX) {
                    const float g = v * dact(ep.zin[o]);
                    ep.hout[o] = (__bf16)g;
                    if (ep.zout) ep.zout[o] = g;
                } else {  // EPI_PARTIAL
                    ep.zout[(size_t)blockIdx.z * M * N + o] = v;
                }
            }
}

// ---------------------------------------------------------------------------
// Layer 1 (14 -> 256), FP32: Z1 = X W1 + b1, H1 = act(Z1).  X rows are 16
// floats (14 used).  One thread per (row, column).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_layer1_fwd(const float *__restrict__ X,
                                                    const float *__restrict__ W1,
                                                    const float *__restrict__ b1, int rows,
                                                    float *__restrict__ Z1, __bf16 *__restrict__ H1) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)rows * kNH) return;
    const int n = (int)(t % kNH);
    const size_t b = t / kNH;
    float z = b1[n];
#pragma unroll
    for (int i = 0; i < kNIN; ++i) z = fmaf(X[b * kXS + i], W1[i * kNH + n], z);
    if (Z1) Z1[t] = z;
    H1[t] = (__bf16)act(z);
}

// ---------------------------------------------------------------------------
// Layer 5 (256 -> 4), FP32 from bf16 H4, fused with the loss: a wave takes a
// row at a time (16 rows), each lane 4 of the 256 inputs, shuffles for the
// dot products.
//   mode 0 (train):   dY = 2 (Y - T) / (rows * 4), loss += sum (Y - T)^2 / (rows * 4)
//   mode 1 (predict): Y written
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_layer5(const __bf16 *__restrict__ H4,
                                                const float *__restrict__ W5,
                                                const float *__restrict__ b5,
                                                const float *__restrict__ Tg, int rows, int mode,
                                                float *__restrict__ Y, float *__restrict__ dY,
                                                float *__restrict__ loss) {
    // a block takes 64 rows, 16 per wave; the loss of its rows is summed in the
    // block and added with one atomic (one per row serialised on the loss word)
    __shared__ float lsum[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const float inv = 1.0f / ((float)rows * kNOUT);
    float l = 0.f;
    for (int rr = 0; rr < 16; ++rr) {
        const size_t b = (size_t)blockIdx.x * 64 + wv * 16 + rr;
        if (b >= (size_t)rows) break;
        float acc[kNOUT] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = lane * 4 + q;
            const float h = (float)H4[b * kNH + k];
#pragma unroll
            for (int o = 0; o < kNOUT; ++o) acc[o] = fmaf(h, W5[k * kNOUT + o], acc[o]);
        }
#pragma unroll
        for (int o = 0; o < kNOUT; ++o)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[o] += __shfl_xor(acc[o], off, 64);
        if (lane == 0) {
#pragma unroll
            for (int o = 0; o < kNOUT; ++o) {
                const float y = acc[o] + b5[o];
                if (mode == 0) {
                    const float d = y - Tg[b * kNOUT + o];
                    dY[b * kNOUT + o] = 2.0f * d * inv;
                    l += d * d;
                } else {
                    Y[b * kNOUT + o] = y;
                }
            }
        }
    }
    if (mode != 0) return;
    if (lane == 0) lsum[wv] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (lsum[0] + lsum[1] + lsum[2] + lsum[3]) * inv);
}

// Layer 5 backward: dW5 += H4^T dY, db5 += sum dY, dZ4 = (dY W5^T) * act'(Z4).
// A wave per row at a time as in the forward.
__global__ __launch_bounds__(256) void k_layer5_bwd(const __bf16 *__restrict__ H4,
                                                    const float *__restrict__ Z4,
                                                    const float *__restrict__ W5,
                                                    const float *__restrict__ dY, int rows,
                                                    float *__restrict__ gW5, float *__restrict__ gb5,
                                                    __bf16 *__restrict__ dZ4) {
    __shared__ float part[4][kNH * kNOUT];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 4 * kNH * kNOUT; i += 256) (&part[0][0])[i] = 0.f;
    __syncthreads();
    // each block handles 64 rows (16 per wave)
    float dsum = 0.f;  // lane o < 4: sum of dY[., o] over this wave's rows
    for (int rr = 0; rr < 16; ++rr) {
        const size_t b = (size_t)blockIdx.x * 64 + wv * 16 + rr;
        if (b >= (size_t)rows) break;
        float d[kNOUT];
#pragma unroll
        for (int o = 0; o < kNOUT; ++o) d[o] = dY[b * kNOUT + o];
        dsum += lane == 0 ? d[0] : lane == 1 ? d[1] : lane == 2 ? d[2] : lane == 3 ? d[3] : 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = lane * 4 + q;
            const size_t idx = b * kNH + k;
            const float h = (float)H4[idx];
            float g = 0.f;
#pragma unroll
            for (int o = 0; o < kNOUT; ++o) {
                part[wv][k * kNOUT + o] += h * d[o];
                g = fmaf(d[o], W5[k * kNOUT + o], g);
            }
            dZ4[idx] = (__bf16)(g * dact(Z4[idx]));
        }
    }
    if (lane < kNOUT) atomicAdd(&gb5[lane], dsum);
    __syncthreads();
    for (int i = threadIdx.x; i < kNH * kNOUT; i += 256)
        atomicAdd(&gW5[i], part[0][i] + part[1][i] + part[2][i] + part[3][i]);
}

// Layer 1 backward: dW1 += X^T dZ1 (11 x 256), db1 += sum dZ1.  Block = 256
// columns x a chunk of rows.
__global__ __launch_bounds__(256) void k_layer1_bwd(const float *__restrict__ X,
                                                    const float *__restrict__ dZ1, int rows,
                                                    int chunk, float *__restrict__ gW1,
                                                    float *__restrict__ gb1) {
    const int n = threadIdx.x;
    const size_t r0 = (size_t)blockIdx.x * chunk;
    const size_t r1 = min((size_t)rows, r0 + chunk);
    float a[kNIN], s = 0.f;
#pragma unroll
    for (int i = 0; i < kNIN; ++i) a[i] = 0.f;
    for (size_t b = r0; b < r1; ++b) {
        const float g = dZ1[b * kNH + n];
        s += g;
#pragma unroll
        for (int i = 0; i < kNIN; ++i) a[i] = fmaf(X[b * kXS + i], g, a[i]);
    }
#pragma unroll
    for (int i = 0; i < kNIN; ++i) atomicAdd(&gW1[i * kNH + n], a[i]);
    atomicAdd(&gb1[n], s);
}

// Column sums of a bf16 [rows x 256] matrix into g (bias gradients).
__global__ __launch_bounds__(256) void k_colsum(const __bf16 *__restrict__ D, int rows, int chunk,
                                                float *__restrict__ g) {
    const int n = threadIdx.x;
    const size_t r0 = (size_t)blockIdx.x * chunk;
    const size_t r1 = min((size_t)rows, r0 + chunk);
    float s = 0.f;
    for (size_t b = r0; b < r1; ++b) s += (float)D[b * kNH + n];
    atomicAdd(&g[n], s);
}

// g[i] += sum over nsplit partials P[s][i].
__global__ __launch_bounds__(256) void k_reduce_partials(const float *__restrict__ P, int nsplit,
                                                         int n, float *__restrict__ g) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; ++k) s += P[(size_t)k * n + i];
    g[i] += s;
}

// Adam (bias-corrected) on the flat FP32 parameter vector.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int n,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// bf16 copies of a 256x256 weight: W (row-major [in][out]) and Wt ([out][in]).
__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ W, __bf16 *__restrict__ Wb,
                                                 __bf16 *__restrict__ Wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNH * kNH) return;
    const int r = i / kNH, c = i % kNH;
    const __bf16 v = (__bf16)W[i];
    Wb[i] = v;
    Wt[c * kNH + r] = v;
}

// ---------------------------------------------------------------------------
// Features of table texels and dataset assembly
// ---------------------------------------------------------------------------
constexpr float kHorizonGap = 0.25f;

// Resolution-independent texel coordinates in [-1, 1] (the arguments of
// getMuMuSNu, common.glsl:175-201), the 7 normalised atmosphere features, and
// the exactly computed orders' value at the texel (ex: S.rgb of an inscatter
// texel, E of an irradiance texel) on a log scale: 0.25 (log(v + 1e-6) + 5),
// about -2 .. 1 over the tables' range.  The higher orders follow the lower
// ones closely (where the view ray meets the ground, where the sun sets), so
// the network learns them relative to what the exact part already resolves.
// E texels use the irradiance coordinates in slots 0 and 2 and -2 in 1 and 3.
__device__ __forceinline__ void texel_features(const FeatSet &fs, unsigned code, const float *ex, float *x) {
    if (code & 0x80000000u) {
        const unsigned t = code & 0x7fffffffu;
        const int i = t % fs.SW, j = t / fs.SW;
        x[0] = 2.f * j / (fs.SH - 1) - 1.f;
        x[1] = -2.f;
        x[2] = 2.f * i / (fs.SW - 1) - 1.f;
        x[3] = -2.f;
    } else {
        const int W4 = fs.RMUS * fs.RNU;
        const int xx = code % W4, y = (code / W4) % fs.RMU, l = code / (W4 * fs.RMU);
        // mu rows: the row's ray length d (getMuMuSNu, common.glsl:175-190, with
        // the layer radius of setLayer, PreprocessAtmo.cpp:255-271) on one log
        // scale for the whole set, from 1 m (the shortest ray of the table: up
        // from the highest layer) to 2 sqrt(Rt^2 - Rg^2) (longer than any ray),
        // the ground-ray half in [-1, -kHorizonGap] and the sky half in
        // [kHorizonGap, 1].  The table rows are linear in d up to a clamp at the
        // shortest ray of a layer (10 m down from the lowest layer), where the
        // inscatter, proportional to d, drops by 10-100x from one row to the
        // next; on a log scale common to all layers that drop is the same slope
        // everywhere, and the gap separates the straight-down and straight-up
        // rows, which are neighbours in the table only.
        const int half = fs.RMU / 2;
        const double Rg = fs.Rg, Rt = fs.Rt, Rg2 = Rg * Rg;
        double rl = (double)l / (fs.RR - 1);
        rl = sqrt(Rg2 + rl * rl * (Rt * Rt - Rg2)) + (l == 0 ? 0.01 : (l == fs.RR - 1 ? -0.001 : 0.0));
        const double rho = sqrt(fmax(rl * rl - Rg2, 0.0)), H = sqrt(Rt * Rt - Rg2);
        double dmin, dmax, d;
        if (y < half) {
            dmin = rl - Rg;
            dmax = 0.999 * rho;
            d = (1.0 - (double)y / (half - 1)) * rho;
        } else {
            dmin = Rt - rl;
            dmax = 0.999 * (rho + H);
            d = (double)(y - half) / (half - 1) * (rho + H);
        }
        d = fmin(fmax(d, dmin), dmax);
        const double dlo = 1e-3, dhi = 2.0 * H;   // in the parameters' length unit (km)
        const float t = (float)fmin(fmax(log(d / dlo) / log(dhi / dlo), 0.0), 1.0);
        x[0] = 2.f * l / (fs.RR - 1) - 1.f;
        x[1] = y < half ? fmaf(-(1.f - kHorizonGap), t, -kHorizonGap) : fmaf(1.f - kHorizonGap, t, kHorizonGap);
        x[2] = 2.f * (xx % fs.RMUS) / (fs.RMUS - 1) - 1.f;
        x[3] = 2.f * (xx / fs.RMUS) / (fs.RNU - 1) - 1.f;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) x[4 + k] = fs.f[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) x[11 + c] = 0.25f * (__logf(fmaxf(ex[c], 0.f) + 1e-6f) + 5.0f);
}

// The exact part's value at a texel (S.rgb, or E of an irradiance texel).
__device__ __forceinline__ void exact_value(const FeatSet &fs, unsigned code, const float4 *S1, const float *E1,
                                            float *ex) {
    if (code & 0x80000000u) {
        const unsigned e = code & 0x7fffffffu;
        ex[0] = E1[3 * e];
        ex[1] = E1[3 * e + 1];
        ex[2] = E1[3 * e + 2];
    } else {
        const float4 v = S1[code];
        ex[0] = v.x;
        ex[1] = v.y;
        ex[2] = v.z;
    }
}

__device__ __forceinline__ unsigned hash32(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (unsigned)((z ^ (z >> 31)) >> 32);
}

// Training rows of one set: `ns` seeded S texels (a quarter of them near the
// horizon) then all E texels.
// target c = (log(max(v_c, 0) + eps) - mu_c) / sd_c, v = S(orders<=N) - S(orders<=K) for S
// texels (rgb) and E(orders<=N) - E(orders<=K) for E texels, K the exactly computed
// orders (S1, E1); the fourth target is 0 (the orders above 1 add 0 to alpha).
__global__ __launch_bounds__(256) void k_make_rows(FeatSet fs, const float4 *__restrict__ S4,
                                                   const float4 *__restrict__ S1,
                                                   const float *__restrict__ E,
                                                   const float *__restrict__ E1, int ns,
                                                   unsigned long long seed, float *__restrict__ X,
                                                   float *__restrict__ Tg, TargetNorm tn) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int nE = fs.SW * fs.SH;
    if (t >= ns + nE) return;
    unsigned code;
    float v[3], ex[3];
    if (t < ns) {
        const unsigned hh = hash32(seed * 0x100000001ull + t);
        if ((t & 3) == 0 && fs.RMU >= 8) {
            // a quarter of the samples from the 8 mu rows around the horizon,
            // where each set's worst texels sit (the last ground rows, the
            // first sky rows); the rest uniform over the table
            const unsigned W4 = (unsigned)(fs.RMUS * fs.RNU);
            const unsigned l = hh % (unsigned)fs.RR, y = (unsigned)(fs.RMU / 2 - 4) + (hh / (unsigned)fs.RR) % 8u;
            code = (l * (unsigned)fs.RMU + y) * W4 + (hh / ((unsigned)fs.RR * 8u)) % W4;
        } else {
            code = hh % (unsigned)fs.n4;
        }
        const float4 a = S4[code], b = S1[code];
        v[0] = a.x - b.x; v[1] = a.y - b.y; v[2] = a.z - b.z;
        ex[0] = b.x; ex[1] = b.y; ex[2] = b.z;
    } else {
        const int e = t - ns;
        code = 0x80000000u | (unsigned)e;
        v[0] = E[3 * e] - E1[3 * e]; v[1] = E[3 * e + 1] - E1[3 * e + 1]; v[2] = E[3 * e + 2] - E1[3 * e + 2];
        ex[0] = E1[3 * e]; ex[1] = E1[3 * e + 1]; ex[2] = E1[3 * e + 2];
    }
    float x[kNIN];
    texel_features(fs, code, ex, x);
#pragma unroll
    for (int i = 0; i < kNIN; ++i) X[(size_t)t * kXS + i] = x[i];
#pragma unroll
    for (int i = kNIN; i < kXS; ++i) X[(size_t)t * kXS + i] = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) Tg[(size_t)t * kNOUT + c] = (logf(fmaxf(v[c], 0.f) + tn.eps) - tn.mu[c]) / tn.sd[c];
    Tg[(size_t)t * kNOUT + 3] = 0.f;
}

// Rows for caller-given texel codes.
__global__ __launch_bounds__(256) void k_code_rows(FeatSet fs, const unsigned *__restrict__ codes,
                                                   int n, const float4 *__restrict__ S1,
                                                   const float *__restrict__ E1, float *__restrict__ X) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    float ex[3], x[kNIN];
    exact_value(fs, codes[t], S1, E1, ex);
    texel_features(fs, codes[t], ex, x);
#pragma unroll
    for (int i = 0; i < kNIN; ++i) X[(size_t)t * kXS + i] = x[i];
#pragma unroll
    for (int i = kNIN; i < kXS; ++i) X[(size_t)t * kXS + i] = 0.f;
}

// Sum and sum of squares of the three rgb targets over n rows (FP64, for the
// output standardisation): out[0..2] sums, out[3..5] sums of squares.
__global__ __launch_bounds__(256) void k_target_stats(const float *__restrict__ Tg, size_t n,
                                                      double *__restrict__ out) {
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = Tg[t * kNOUT + c];
            a[c] += v;
            a[3 + c] += v * v;
        }
    }
    __shared__ double red[6][256];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) atomicAdd(&out[threadIdx.x], red[threadIdx.x][0]);
}

// In-place standardisation of the three rgb targets.
__global__ __launch_bounds__(256) void k_norm_targets(float *__restrict__ Tg, size_t n, TargetNorm tn) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) Tg[t * kNOUT + c] = (Tg[t * kNOUT + c] - tn.mu[c]) / tn.sd[c];
}


// Random training batch: rows idx = hash(seed, b) mod n copied from the dataset.
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ Xa, const float *__restrict__ Ta,
                                                size_t n, int B, unsigned long long seed,
                                                float *__restrict__ X, float *__restrict__ Tg) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    const unsigned long long h = ((unsigned long long)hash32(seed * 0x9E3779B1ull + 2 * b) << 32) |
                                 hash32(seed * 0x9E3779B1ull + 2 * b + 1);
    const size_t i = (size_t)(h % n);
#pragma unroll
    for (int k = 0; k < kXS; ++k) X[(size_t)b * kXS + k] = Xa[i * kXS + k];
#pragma unroll
    for (int k = 0; k < kNOUT; ++k) Tg[(size_t)b * kNOUT + k] = Ta[i * kNOUT + k];
}

// Per-set max relative error of a surrogate table against the plain one,
// floored at 1e-4 of the plain table's largest magnitude (scale given).
__global__ __launch_bounds__(256) void k_relerr(const float *__restrict__ a, const float *__restrict__ b,
                                                size_t n, int stride, int ncomp, float floorv,
                                                unsigned int *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        for (int c = 0; c < ncomp; ++c) {
            const float x = a[i * stride + c], y = b[i * stride + c];
            const float den = fmaxf(fabsf(y), floorv);
            float r = fabsf(x - y) / den;
            if (!(r == r)) r = 3.0e38f;  // NaN counts as a failure
            e = fmaxf(e, r);
        }
    }
    for (int off = 32; off > 0; off >>= 1) e = fmaxf(e, __shfl_xor(e, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(e));
}

// max |v| over a table (for the relative-error floor)
__global__ __launch_bounds__(256) void k_absmax(const float *__restrict__ a, size_t n, int stride,
                                                int ncomp, unsigned int *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    float e = 0.f;
    if (i < n)
        for (int c = 0; c < ncomp; ++c) e = fmaxf(e, fabsf(a[i * stride + c]));
    for (int off = 32; off > 0; off >>= 1) e = fmaxf(e, __shfl_xor(e, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(e));
}

// ---------------------------------------------------------------------------
// Fused inference: texel features -> layer 1 -> kNG x (256x256 + bias + act)
// -> layer 5 -> orders >= 2 added to S / written to E, all in one kernel with
// the activations in registers.
//
// The products are computed transposed, D = W^T H^T (output features x rows),
// with v_mfma_f32_32x32x16_f16 (FP16 weights and activations, FP32
// accumulation; layer 1 with the _bf16 form): the A operand is a 32-feature x 16-k slice of
// the weight (W^T rows, staged in LDS), the B operand a 16-k x 32-row slice of
// the activations, held by each lane for its own row (lane l: row l%32,
// features 16kb + 8(l/32) .. +7 of every k block kb).  A D tile gives lane l
// the features 8q + 4(l/32) + t (q, t < 4) of row l%32; one swap of 4 values
// with lane l^32 turns it into the B layout of the next layer.  Each wave owns
// 32*RBW rows through all five layers; a workgroup stages one 256x256 weight
// (rows padded to 528 B, conflict-free 16-byte LDS reads) per layer.
// Layers 1 and 5 split their FP32 weights into hi + lo parts, bf16 for layer 1
// (x W ~ xh Wh + xh Wl + xl Wh, its bias riding in input column 14) and FP16
// for layer 5 (its input is FP16 already, so Wh + Wl only): relative error
// ~1e-5 and ~1e-6, inside the FP16 rounding of the activations.
// ---------------------------------------------------------------------------
constexpr int kRowPad = 264;   // 16-bit values per staged weight row (512 B + 16 B pad)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// act of a (low half) and b (high half) in packed FP16: one v_cvt_pk_f16_f32
// (round to nearest even) and five v_pk_*_f16, the clamp of t riding on the
// first FMA.  FP16 keeps 11 significant bits where the bf16 activations of the
// training path keep 8, and its range covers the pre-activations (|z| < 65504).
__device__ __forceinline__ unsigned act_pk(float a, float b) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
    const f16x2 z = __builtin_convertvector((f32x2){a, b}, f16x2);
    const _Float16 k6 = (_Float16)(1.0f / 6.0f);
    const f16x2 c6 = {k6, k6}, c05 = {(_Float16)0.5f, (_Float16)0.5f};
    const f16x2 zero = {(_Float16)0.0f, (_Float16)0.0f}, one = {(_Float16)1.0f, (_Float16)1.0f};
    const f16x2 m2 = {(_Float16)-2.0f, (_Float16)-2.0f}, three = {(_Float16)3.0f, (_Float16)3.0f};
    f16x2 t = __builtin_elementwise_fma(z, c6, c05);
    t = __builtin_elementwise_min(__builtin_elementwise_max(t, zero), one);
    const f16x2 u = __builtin_elementwise_fma(m2, t, three);
    return __builtin_bit_cast(unsigned, z * (t * t * u));
}

// D tile (features x 32 rows) -> act(D + bias) as the FP16 B operands of k
// blocks 2nt (lo) and 2nt + 1 (hi).  bias: the tile's 32 biases or nullptr.
__device__ __forceinline__ void tile_to_b(const f32x16 &acc, const float *bias, int h, f16x8 &lo,
                                          f16x8 &hi) {
    unsigned P[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float z[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)   // (no "+ 0.0f" without a bias: it would be kept, -0 + 0 = +0)
            z[t] = bias ? acc[4 * q + t] + bias[8 * q + 4 * h + t] : acc[4 * q + t];
        P[q][0] = act_pk(z[0], z[1]);
        P[q][1] = act_pk(z[2], z[3]);
    }
    // lane h = 0 keeps features 0-3, 16-19 and needs 4-7, 20-23 from lane l^32;
    // lane h = 1 keeps 12-15, 28-31 and needs 8-11, 24-27.  v_permlane32_swap
    // exchanges the upper 32 lanes of its first operand with the lower 32 lanes
    // of its second: swapping (P[0], P[1]) leaves lanes h = 0 with their own
    // P[0] and their partner's P[0], lanes h = 1 with their partner's P[1] and
    // their own P[1], i.e. features 8h .. 8h + 7 in order in both halves (and
    // (P[2], P[3]) the same for 16 + 8h ..), with no lane-dependent selects.
    unsigned X[2], Y[2], Z[2], W[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const auto a = __builtin_amdgcn_permlane32_swap(P[0][i], P[1][i], false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(P[2][i], P[3][i], false, false);
        X[i] = a[0]; Y[i] = a[1]; Z[i] = b[0]; W[i] = b[1];
    }
    const u32x4 L = u32x4{X[0], X[1], Y[0], Y[1]};
    const u32x4 Hh = u32x4{Z[0], Z[1], W[0], W[1]};
    lo = __builtin_bit_cast(f16x8, L);
    hi = __builtin_bit_cast(f16x8, Hh);
}

typedef __attribute__((address_space(3))) void lds_void;

// Asynchronous copy of rows [r0, r0 + 128) of a [256][256] FP16 weight into
// the padded LDS image (global_load_lds_dwordx4: each active lane moves 16 B
// straight into LDS; one row = 32 lanes).  Each of the NW waves moves 128/NW rows.
template <int NW>
__device__ __forceinline__ void stage_half_async(const _Float16 *W, _Float16 *wl, int r0, int wv, int lane) {
    constexpr int RPWV = 128 / NW;
    if (lane < 32) {
#pragma unroll 8
        for (int r = 0; r < RPWV; ++r) {
            const int row = r0 + RPWV * wv + r;
            __builtin_amdgcn_global_load_lds((const void *)(W + row * kNH + lane * 8),
                                             (lds_void *)(wl + row * kRowPad), 16, 0, 0);
        }
    }
}

// Output tiles nt0 .. nt0+3 of one hidden layer from the staged weight rows,
// software-pipelined: the matrix-core products of tile nt run while the
// vector units apply act and relayout tile nt - 1.  The accumulators start at
// the tile's biases (element 4q + t of lane (c, h) is feature 8q + 4h + t).
template <int RBW>
__device__ __forceinline__ void tile_mfma(const _Float16 *wl, const float *bias, int nt, int c, int h,
                                          const f16x8 (&H)[RBW][16], f32x16 (&acc)[RBW]) {
    const float4 *bq = (const float4 *)(bias + 32 * nt + 4 * h);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 b = bq[2 * q];
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb) {
            acc[rb][4 * q] = b.x; acc[rb][4 * q + 1] = b.y; acc[rb][4 * q + 2] = b.z; acc[rb][4 * q + 3] = b.w;
        }
    }
    const _Float16 *wrow = wl + (32 * nt + c) * kRowPad + 8 * h;
#pragma unroll
    for (int kb = 0; kb < 16; ++kb) {
        const f16x8 a = *(const f16x8 *)(wrow + 16 * kb);
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
            acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, H[rb][kb], acc[rb], 0, 0, 0);
    }
}

template <int RBW>
__device__ __forceinline__ void hidden_half(const _Float16 *wl, const float *bias, int nt0, int c, int h,
                                            const f16x8 (&H)[RBW][16], f16x8 (&Hn)[RBW][16]) {
    f32x16 acc[2][RBW];
    tile_mfma<RBW>(wl, bias, nt0, c, h, H, acc[0]);
#pragma unroll
    for (int i = 1; i <= 4; ++i) {
        if (i < 4) tile_mfma<RBW>(wl, bias, nt0 + i, c, h, H, acc[i & 1]);
        const int nt = nt0 + i - 1;
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
            tile_to_b(acc[(i - 1) & 1][rb], nullptr, h, Hn[rb][2 * nt], Hn[rb][2 * nt + 1]);
    }
}

__device__ __forceinline__ void wait_all_and_sync() {
    __builtin_amdgcn_s_waitcnt(0);  // the LDS-direct loads of this wave have landed
    __syncthreads();
}

// LDS image: rows 0..127 of the current weight ("A") and rows 128..255 ("B"),
// 528-byte rows.  While one half is read by the matrix cores the other is
// being refilled for the next step, so every weight load after the first
// overlaps compute.  Layer 1's weights (16 KB) sit in B before layer 2's
// second half arrives; layer 5's (4 KB) in A after layer 4's first half.
template <int RBW, int NW>
__global__ __launch_bounds__(64 * NW) void k_mlp_infer(const InferSet *__restrict__ sets, int nsets,
                                                       FusedWeights fw, TargetNorm tn) {
    extern __shared__ __attribute__((aligned(16))) _Float16 wl[];   // [256][kRowPad] (layer 1: bf16 bits)
    __shared__ float bs[kNG][kNH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, h = lane >> 5, c = lane & 31;
    // the set of this workgroup: last set whose first tile <= blockIdx.x
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sets[mid].tile0 <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    const InferSet &is = sets[lo];
    const FeatSet fs = is.fs;
    const int row0 = 0, n = fs.n4 + fs.SW * fs.SH;
    float4 *__restrict__ S = is.S;
    float *__restrict__ E = is.E;
    const int rw = ((int)blockIdx.x - is.tile0) * (32 * RBW * NW) + wv * (32 * RBW);    // first row of this wave
    _Float16 *wB = wl + 128 * kRowPad;

    // layer-1 weights (hi then lo, [256][16]) into B, hidden biases; layer 2's first half -> A
    for (int i = threadIdx.x; i < 1024; i += 64 * NW)
        *(uint4 *)(wB + i * 8) = *(const uint4 *)((i < 512 ? fw.a1h : fw.a1l) + (i & 511) * 8);
    for (int i = threadIdx.x; i < kNG * kNH; i += 64 * NW) (&bs[0][0])[i] = fw.bh[i];
    __syncthreads();
    stage_half_async<NW>(fw.wt, wl, 0, wv, lane);

    f16x8 H[RBW][16];
    // (1) features and layer 1
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb) {
        const int r = min(rw + 32 * rb + c, n - 1) + row0;
        const unsigned code = r < fs.n4 ? (unsigned)r : (0x80000000u | (unsigned)(r - fs.n4));
        float ex[3], x[kNIN];
        exact_value(fs, code, is.S, is.E, ex);   // the exact part, before this kernel adds to it
        texel_features(fs, code, ex, x);
        float xe[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) xe[k] = k < kNIN ? x[k] : (k == kNIN ? 1.0f : 0.0f);
        bf16x8 xh, xl;
        const unsigned hm = 0u - (unsigned)h;  // all ones on lanes 32..63
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            // lane half h takes input columns 8h .. 8h+7 (bit select, no indexed array)
            const float v = __uint_as_float((__float_as_uint(xe[i]) & ~hm) | (__float_as_uint(xe[8 + i]) & hm));
            const __bf16 vh = (__bf16)v;
            xh[i] = vh;
            xl[i] = (__bf16)(v - (float)vh);
        }
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const bf16x8 ah = *(const bf16x8 *)(wB + (32 * nt + c) * 16 + 8 * h);   // bf16 bits
            const bf16x8 al = *(const bf16x8 *)(wB + 4096 + (32 * nt + c) * 16 + 8 * h);
            f32x16 acc;
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, xh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xl, acc, 0, 0, 0);
            tile_to_b(acc, nullptr, h, H[rb][2 * nt], H[rb][2 * nt + 1]);
        }
    }
    // (2) hidden layers 2..kNL
#pragma unroll 1
    for (int L = 0; L < kNG; ++L) {
        const _Float16 *W = fw.wt + (size_t)L * kNH * kNH;
        wait_all_and_sync();                          // A holds rows 0..127; B is free
        stage_half_async<NW>(W, wl, 128, wv, lane);   // B <- rows 128..255
        f16x8 Hn[RBW][16];
        hidden_half<RBW>(wl, bs[L], 0, c, h, H, Hn);
        wait_all_and_sync();                          // B holds rows 128..255; A is free
        if (L < kNG - 1) {
            stage_half_async<NW>(W + (size_t)kNH * kNH, wl, 0, wv, lane);   // A <- next layer's rows 0..127
        } else if (threadIdx.x < 256) {               // A <- layer 5 (hi then lo, [4][256])
            *(uint4 *)(wl + threadIdx.x * 8) = *(const uint4 *)((threadIdx.x < 128 ? fw.a5h : fw.a5l) +
                                                               (threadIdx.x & 127) * 8);
        }
        hidden_half<RBW>(wl, bs[L], 4, c, h, H, Hn);
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int kb = 0; kb < 16; ++kb) H[rb][kb] = Hn[rb][kb];
    }
    // (3) layer 5 and the combination with the exact part
    wait_all_and_sync();
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb) {
        f32x16 acc;
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
            f16x8 ah, al;
            if (c < kNOUT) {
                ah = *(const f16x8 *)(wl + c * kNH + 16 * kb + 8 * h);
                al = *(const f16x8 *)(wl + 1024 + c * kNH + 16 * kb + 8 * h);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) ah[i] = al[i] = (_Float16)0.0f;
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, H[rb][kb], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, H[rb][kb], acc, 0, 0, 0);
        }
        // lanes h = 0 hold outputs 0..3 of row c in acc[0..3]
        const int rr = rw + 32 * rb + c;
        if (h == 0 && rr < n) {
            const int r = row0 + rr;
            float v[3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
                v[k] = fmaxf(__expf((acc[k] + fw.b5[k]) * tn.sd[k] + tn.mu[k]) - tn.eps, 0.f);
            if (r < fs.n4) {
                float4 sv = S[r];
                sv.x += v[0]; sv.y += v[1]; sv.z += v[2];
                S[r] = sv;
            } else {
                const int e = r - fs.n4;
                E[3 * e] += v[0]; E[3 * e + 1] += v[1]; E[3 * e + 2] += v[2];
            }
        }
    }
}

// Weights of the fused kernel from the FP32 parameters: layer 1 (+ bias in
// column 14) split into bf16 hi/lo, layer 5 into FP16 hi/lo, the hidden
// weights as FP16 [out][in] and their biases packed.
__global__ __launch_bounds__(256) void k_prep_fused(const float *__restrict__ W1, const float *__restrict__ b1,
                                                    const float *__restrict__ W5, const float *__restrict__ Wh0,
                                                    const float *__restrict__ bh0, size_t bstride, __bf16 *a1h,
                                                    __bf16 *a1l, _Float16 *wt, _Float16 *a5h, _Float16 *a5l,
                                                    float *bh) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kNH * 16) {
        const int nn = i / 16, k = i % 16;
        const float v = k < kNIN ? W1[k * kNH + nn] : (k == kNIN ? b1[nn] : 0.0f);
        const __bf16 vh = (__bf16)v;
        a1h[i] = vh;
        a1l[i] = (__bf16)(v - (float)vh);
    } else if (i < kNH * 16 + kNOUT * kNH) {
        const int j = i - kNH * 16, o = j / kNH, k = j % kNH;
        const float v = W5[k * kNOUT + o];
        const _Float16 vh = (_Float16)v;
        a5h[j] = vh;
        a5l[j] = (_Float16)(v - (float)vh);
    } else if (i < kNH * 16 + kNOUT * kNH + kNG * kNH) {
        const int j = i - kNH * 16 - kNOUT * kNH;
        bh[j] = bh0[(size_t)(j / kNH) * bstride + j % kNH];
    } else if (i < kNH * 16 + kNOUT * kNH + kNG * kNH + kNG * kNH * kNH) {
        const int j = i - kNH * 16 - kNOUT * kNH - kNG * kNH;   // layer L, output n, input k
        const int L = j / (kNH * kNH), n = (j / kNH) % kNH, k = j % kNH;
        wt[j] = (_Float16)Wh0[(size_t)L * bstride + (size_t)k * kNH + n];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t gemm(bool ta, bool tb, int epi, const __bf16 *A, const __bf16 *B, int M, int N, int K,
                int nsplit, const GemmEpi &ep, hipStream_t st) {
    if (M % 128 || N % 128 || K % 32 || nsplit < 1 || (K / nsplit) % 32) return hipErrorInvalidValue;
    dim3 grid(N / 128, M / 128, nsplit);
    const int kc = K / nsplit;
#define SG_L(TA_, TB_, E_) hipLaunchKernelGGL((k_gemm<TA_, TB_, E_>), grid, dim3(256), 0, st, A, B, M, N, K, kc, ep)
    if (!ta && tb && epi == EPI_FWD) SG_L(false, true, EPI_FWD);
    else if (!ta && tb && epi == EPI_FWD_INF) SG_L(false, true, EPI_FWD_INF);
    else if (!ta && tb && epi == EPI_DX) SG_L(false, true, EPI_DX);
    else if (ta && !tb && epi == EPI_PARTIAL) SG_L(true, false, EPI_PARTIAL);
    else return hipErrorInvalidValue;
#undef SG_L
    return hipGetLastError();
}

static unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

hipError_t layer1_fwd(const float *X, const float *W1, const float *b1, int rows, float *Z1, __bf16 *H1,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_layer1_fwd, dim3(blocks((size_t)rows * kNH)), dim3(256), 0, st, X, W1, b1, rows, Z1, H1);
    return hipGetLastError();
}
hipError_t layer5(const __bf16 *H4, const float *W5, const float *b5, const float *Tg, int rows, int mode,
                  float *Y, float *dY, float *loss, hipStream_t st) {
    hipLaunchKernelGGL(k_layer5, dim3((rows + 63) / 64), dim3(256), 0, st, H4, W5, b5, Tg, rows, mode, Y, dY, loss);
    return hipGetLastError();
}
hipError_t layer5_bwd(const __bf16 *H4, const float *Z4, const float *W5, const float *dY, int rows,
                      float *gW5, float *gb5, __bf16 *dZ4, hipStream_t st) {
    hipLaunchKernelGGL(k_layer5_bwd, dim3((rows + 63) / 64), dim3(256), 0, st, H4, Z4, W5, dY, rows, gW5, gb5, dZ4);
    return hipGetLastError();
}
hipError_t layer1_bwd(const float *X, const float *dZ1, int rows, float *gW1, float *gb1, hipStream_t st) {
    const int chunk = 64;   // 1024 workgroups at the training batch
    hipLaunchKernelGGL(k_layer1_bwd, dim3((rows + chunk - 1) / chunk), dim3(256), 0, st, X, dZ1, rows, chunk, gW1, gb1);
    return hipGetLastError();
}
hipError_t colsum(const __bf16 *D, int rows, float *g, hipStream_t st) {
    const int chunk = 64;   // 1024 workgroups at the training batch
    hipLaunchKernelGGL(k_colsum, dim3((rows + chunk - 1) / chunk), dim3(256), 0, st, D, rows, chunk, g);
    return hipGetLastError();
}
hipError_t reduce_partials(const float *P, int nsplit, int n, float *g, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks(n)), dim3(256), 0, st, P, nsplit, n, g);
    return hipGetLastError();
}
hipError_t adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                float eps, float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t to_bf16(const float *W, __bf16 *Wb, __bf16 *Wt, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(kNH * kNH)), dim3(256), 0, st, W, Wb, Wt);
    return hipGetLastError();
}
hipError_t make_rows(const FeatSet &fs, const float4 *S4, const float4 *S1, const float *E, const float *E1, int ns,
                     unsigned long long seed, float *X, float *Tg, const TargetNorm &tn, hipStream_t st) {
    const int n = ns + fs.SW * fs.SH;
    hipLaunchKernelGGL(k_make_rows, dim3(blocks(n)), dim3(256), 0, st, fs, S4, S1, E, E1, ns, seed, X, Tg, tn);
    return hipGetLastError();
}
hipError_t code_rows(const FeatSet &fs, const unsigned *codes, int n, const float4 *S1, const float *E1, float *X,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_code_rows, dim3(blocks(n)), dim3(256), 0, st, fs, codes, n, S1, E1, X);
    return hipGetLastError();
}
hipError_t target_stats(const float *Tg, size_t n, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_target_stats, dim3(1024), dim3(256), 0, st, Tg, n, out);
    return hipGetLastError();
}
hipError_t norm_targets(float *Tg, size_t n, const TargetNorm &tn, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_targets, dim3(blocks(n)), dim3(256), 0, st, Tg, n, tn);
    return hipGetLastError();
}
hipError_t gather(const float *Xa, const float *Ta, size_t n, int B, unsigned long long seed, float *X,
                  float *Tg, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(blocks(B)), dim3(256), 0, st, Xa, Ta, n, B, seed, X, Tg);
    return hipGetLastError();
}
hipError_t relerr(const float *a, const float *b, size_t n, int stride, int ncomp, float floorv,
                  unsigned *out, hipStream_t st) {
    hipLaunchKernelGGL(k_relerr, dim3(blocks(n)), dim3(256), 0, st, a, b, n, stride, ncomp, floorv, out);
    return hipGetLastError();
}
hipError_t absmax(const float *a, size_t n, int stride, int ncomp, unsigned *out, hipStream_t st) {
    hipLaunchKernelGGL(k_absmax, dim3(blocks(n)), dim3(256), 0, st, a, n, stride, ncomp, out);
    return hipGetLastError();
}

}  // namespace sg
}  // namespace atmo

namespace atmo {
namespace sg {
constexpr int kFusedRBW = 1, kFusedNW = 8;
size_t fused_lds() { return (size_t)kNH * kRowPad * sizeof(__bf16); }
int fused_rows_per_block() { return 32 * kFusedRBW * kFusedNW; }
hipError_t prep_fused(const float *W1, const float *b1, const float *W5, const float *Wh0, const float *bh0,
                      size_t bstride, __bf16 *a1h, __bf16 *a1l, _Float16 *wt, _Float16 *a5h, _Float16 *a5l,
                      float *bh, hipStream_t st) {
    const int n = kNH * 16 + kNOUT * kNH + kNG * kNH + kNG * kNH * kNH;
    hipLaunchKernelGGL(k_prep_fused, dim3(blocks(n)), dim3(256), 0, st, W1, b1, W5, Wh0, bh0, bstride, a1h, a1l,
                       wt, a5h, a5l, bh);
    return hipGetLastError();
}
hipError_t mlp_infer(const InferSet *sets, int nsets, int ntiles, const FusedWeights &fw, const TargetNorm &tn,
                     hipStream_t st) {
    if (nsets <= 0 || ntiles <= 0) return hipSuccess;
    // the kernel takes more dynamic LDS than the default limit: raise it once per device
    static std::mutex mu;
    static unsigned long long done = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!(done & (1ull << dev))) {
            e = hipFuncSetAttribute((const void *)k_mlp_infer<kFusedRBW, kFusedNW>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)fused_lds());
            if (e != hipSuccess) return e;
            done |= 1ull << dev;
        }
    }
    hipLaunchKernelGGL((k_mlp_infer<kFusedRBW, kFusedNW>), dim3(ntiles), dim3(64 * kFusedNW), fused_lds(), st, sets,
                       nsets, fw, tn);
    return hipGetLastError();
}
}  // namespace sg
}  // namespace atmo
