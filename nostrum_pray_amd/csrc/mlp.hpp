// Radiance-cache MLP on gfx950 matrix cores.
//
// Network: 10 features -> 192 encoded inputs (171 used) -> 5 hidden layers of
// width 64 (ReLU) -> 3 outputs, no biases (input channel 171 is a constant 1,
// which gives the first layer an offset).  Weights and activations are BF16,
// products accumulate in FP32 (v_mfma_f32_16x16x32_bf16), the optimiser keeps
// FP32 master weights and Adam moments.
//
// Features per sample: position in the scene box mapped to [0,1]^3, ray
// direction and surface normal as (theta/pi, phi/2pi + 1/2), albedo rgb.
// Encoding (192 input channels; see enc_next): with channel c held in slot
// s of lane group g of reduction block m (c = kchan(m, g, s)) and
// q = 8m + s, e = 4(q % 6) + g:
//   q < 42   sin (e even) or cos (e odd) of 2 pi 2^(e/2 - 1) x_f, f = q / 6
//            (the lowest frequency of a feature evaluated, the others by
//            angle doubling)
//   q == 42  albedo r, g, b for g = 0, 1, 2 and the constant 1 for g = 3
//   q > 42   0
// so the feature read by a slot does not depend on the lane.
//
// All layers are computed "transposed": Z = W * H with output channels as
// matrix rows and samples as columns, 16 samples per wave.  The accumulator
// layout of one 16x16 tile (lane l holds rows 4(l/16)+r, column l%16) is then,
// after ReLU and rounding, the B operand of the next layer for a 32-deep
// reduction block whose k order is permuted:
//   k slot 8g+s  <->  channel 32m + 4g + s        (s < 4)
//                     channel 32m + 16 + 4g + s-4  (s >= 4)
// The weights are stored pre-permuted in the same order ("fragments": per
// 16x32 A block, 64 lanes x 8 bf16), so activations never leave registers.
// The backward pass uses the same trick with W transposed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"

namespace prm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;
constexpr int N_FEAT = 10;
constexpr int IN_PAD = 192;
constexpr int WIDTH = 64;
constexpr int N_LAYERS = 6;
constexpr int N_PARAMS = 64 * 192 + 4 * 64 * 64 + 32 * 64;  // 30720
constexpr int N_FFRAG = 24 + 4 * 8 + 4;                      // forward A blocks
constexpr int N_BFRAG = 4 * 8 + 4;                           // backward A blocks (layers 1..5)
constexpr int N_WTILES = 48 + 4 * 16 + 4;                    // weight-gradient 16x16 tiles
constexpr float LOSS_EPS = 0.01f;

struct Layer {
    int out, in, pofs, ffrag, bfrag, wtile;
};
__host__ __device__ constexpr Layer layer(int L) {
    return L == 0   ? Layer{64, 192, 0, 0, -1, 0}
           : L < 5  ? Layer{64, 64, 12288 + (L - 1) * 4096, 24 + (L - 1) * 8, (L - 1) * 8, 48 + (L - 1) * 16}
                    : Layer{32, 64, 28672, 56, 32, 112};
}

// Position of k (0..31) of a 32-deep reduction block in a fragment.
__host__ __device__ inline void kslot(int k, int &g, int &s) {
    const int kk = k & 15;
    g = kk >> 2;
    s = (kk & 3) + (k >= 16 ? 4 : 0);
}
// Channel held in slot s of lane group g for reduction block m.
__host__ __device__ inline int kchan(int m, int g, int s) { return 32 * m + (s < 4 ? 4 * g + s : 16 + 4 * g + s - 4); }

// Whether encoded input channel c can be nonzero (172 of the 192 can).
__host__ __device__ inline bool input_used(int c) {
    int g, s;
    kslot(c & 31, g, s);
    return 8 * (c >> 5) + s <= 42;
}

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// ReLU with +0 for every non-positive input (and NaN), so that a nonzero bf16
// bit pattern means an active unit.
__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

union Frag {
    bf16x8 v;
    uint4 u;
    uint16_t h[8];
};

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Encoded inputs in slot order.  Slot q = 8m + s of lane group g holds, for
// q < 42, sin (g even) or cos (g odd) of 2 pi 2^(2j + h - 1) x_f with
// f = q / 6, j = q % 6, h = g / 2; q == 42 the albedo (g < 3) or the constant
// 1; q > 42 zero.  The six frequencies of a feature in one lane grow by 4x
// from slot to slot, so only the first is evaluated (sincos2pi of the
// fraction of 2^(h-1) x_f); each next one comes from the previous by two
// angle doublings, sin 2a = 2 sin a cos a and cos 2a = (cos a - sin a)(cos a
// + sin a), in FP32 (within 1e-4 of the exact value at the highest
// frequency).  Calls must come in increasing q (m outer, s inner, both
// compile-time constants where this is inlined); `sn`, `cs` carry the
// current feature's sine and cosine from one call to the next.
__device__ __forceinline__ void enc_double(float &sn, float &cs) {
    const float s2 = 2.0f * (sn * cs);
    const float c2 = (cs - sn) * (cs + sn);
    sn = s2, cs = c2;
}

__device__ __forceinline__ float enc_next(const float *f, int m, int g, int s, float &sn, float &cs) {
    const int q = 8 * m + s;
    if (q < 42) {
        if (q % 6 == 0) {
            const float x = ldexpf(f[q / 6], (g >> 1) - 1);
            prd::sincos2pi(x - floorf(x), sn, cs);
        } else {
            enc_double(sn, cs);
            enc_double(sn, cs);
        }
        return (g & 1) ? cs : sn;
    }
    if (q == 42) return g == 0 ? f[7] : (g == 1 ? f[8] : (g == 2 ? f[9] : 1.f));
    return 0.f;
}

// Fragment positions of master parameter p (forward and, for layers >= 1,
// backward / transposed).  Returns the layer.
__device__ __forceinline__ int param_slots(int p, int &fpos, int &bpos) {
    int L = 0;
    while (L < N_LAYERS - 1 && p >= layer(L + 1).pofs) ++L;
    const Layer Y = layer(L);
    const int q = p - Y.pofs, o = q / Y.in, c = q % Y.in;
    int g, s;
    kslot(c & 31, g, s);
    fpos = ((Y.ffrag + (o >> 4) * (Y.in / 32) + (c >> 5)) * WAVE + g * 16 + (o & 15)) * 8 + s;
    bpos = -1;
    if (L >= 1) {
        kslot(o & 31, g, s);
        bpos = ((Y.bfrag + (c >> 4) * (Y.out / 32) + (o >> 5)) * WAVE + g * 16 + (c & 15)) * 8 + s;
    }
    return L;
}

static __global__ void k_mlp_pack(const float *master, uint16_t *ffrag, uint16_t *bfrag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N_PARAMS) return;
    int fpos, bpos;
    param_slots(p, fpos, bpos);
    const uint16_t b = bf16_bits(master[p]);
    ffrag[fpos] = b;
    if (bpos >= 0) bfrag[bpos] = b;
}

// Forward pass of one 16-sample block held by a wave.  X: encoded inputs as
// B fragments (6 reduction blocks), H: hidden activations (5 layers x 2
// blocks), returns the output tile (lanes 0..15 hold rgb in elements 0..2).
__device__ __forceinline__ f32x4 forward16(const bf16x8 *__restrict__ Wf, const Frag (&X)[6], Frag (&H)[5][2],
                                           uint32_t lane) {
    f32x4 acc[4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 6; ++m) acc[ob] = mfma(Wf[(ob * 6 + m) * WAVE + lane], X[m].v, acc[ob]);
    }
#pragma unroll
    for (int L = 0; L < 5; ++L) {
        if (L > 0) {
            const Layer Y = layer(L);
#pragma unroll
            for (int ob = 0; ob < 4; ++ob) {
                acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int m = 0; m < 2; ++m) acc[ob] = mfma(Wf[(Y.ffrag + ob * 2 + m) * WAVE + lane], H[L - 1][m].v, acc[ob]);
            }
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                H[L][m].h[r] = bf16_bits(relu(acc[2 * m][r]));
                H[L][m].h[4 + r] = bf16_bits(relu(acc[2 * m + 1][r]));
            }
    }
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
    const Layer Y5 = layer(5);
#pragma unroll
    for (int m = 0; m < 2; ++m) y = mfma(Wf[(Y5.ffrag + m) * WAVE + lane], H[4][m].v, y);
    return y;
}

__device__ __forceinline__ void encode_frags(const float *f, uint32_t g, Frag (&X)[6]) {
    float sn = 0.f, cs = 0.f;
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int s = 0; s < 8; ++s) X[m].h[s] = bf16_bits(enc_next(f, m, (int)g, s, sn, cs));
}

// Queries for inference: 16 floats each {features[10], weight rgb, slot bits,
// unused x2}.  Output: out[3*i] (when out is set), or weight * max(y, 0) into
// contrib[3*slot] (a pixel of the renderer) or targets[slot] (a training
// sample, marked valid).
struct InferArgs {
    const float *queries;
    uint32_t n;
    const uint32_t *n_dev;
    float *out;
    float *contrib;
    float4 *targets;
};

// Inference: the forward fragments are staged in LDS once per block (60 KB);
// each wave evaluates INF_NB 16-sample blocks per iteration so that every
// weight fragment read from LDS feeds INF_NB matrix operations.
constexpr int INF_NB = 4;

static __global__ __launch_bounds__(256) void k_mlp_infer(const bf16x8 *__restrict__ Wg, InferArgs A) {
    __shared__ bf16x8 Wf[N_FFRAG * WAVE];
    for (int k = threadIdx.x; k < N_FFRAG * WAVE; k += blockDim.x) Wf[k] = Wg[k];
    __syncthreads();
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t j = lane & 15, g = lane >> 4;
    const uint32_t n = A.n_dev ? *A.n_dev : A.n;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t n_waves = gridDim.x * blockDim.x / WAVE;
    constexpr uint32_t PER = 16 * INF_NB;
    for (uint32_t b = wave; b * PER < n; b += n_waves) {
        float f[INF_NB][N_FEAT];
#pragma unroll
        for (int q = 0; q < INF_NB; ++q) {
            const uint32_t i = b * PER + q * 16 + j;
#pragma unroll
            for (int k = 0; k < N_FEAT; ++k) f[q][k] = i < n ? A.queries[16 * (size_t)i + k] : 0.f;
        }
        f32x4 acc[4][INF_NB];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
            for (int q = 0; q < INF_NB; ++q) acc[ob][q] = f32x4{0.f, 0.f, 0.f, 0.f};
        float sn[INF_NB] = {}, cs[INF_NB] = {};
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            Frag X[INF_NB];
#pragma unroll
            for (int q = 0; q < INF_NB; ++q)
#pragma unroll
                for (int s = 0; s < 8; ++s) X[q].h[s] = bf16_bits(enc_next(f[q], m, (int)g, s, sn[q], cs[q]));
#pragma unroll
            for (int ob = 0; ob < 4; ++ob) {
                const bf16x8 w = Wf[(ob * 6 + m) * WAVE + lane];
#pragma unroll
                for (int q = 0; q < INF_NB; ++q) acc[ob][q] = mfma(w, X[q].v, acc[ob][q]);
            }
        }
        Frag H[INF_NB][2];
#pragma unroll
        for (int L = 0; L < 5; ++L) {
            if (L > 0) {
                const Layer Y = layer(L);
#pragma unroll
                for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
                    for (int q = 0; q < INF_NB; ++q) acc[ob][q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const bf16x8 w = Wf[(Y.ffrag + ob * 2 + m) * WAVE + lane];
#pragma unroll
                        for (int q = 0; q < INF_NB; ++q) acc[ob][q] = mfma(w, H[q][m].v, acc[ob][q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < INF_NB; ++q)
#pragma unroll
                for (int m = 0; m < 2; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        H[q][m].h[r] = bf16_bits(relu(acc[2 * m][q][r]));
                        H[q][m].h[4 + r] = bf16_bits(relu(acc[2 * m + 1][q][r]));
                    }
        }
        f32x4 y[INF_NB];
#pragma unroll
        for (int q = 0; q < INF_NB; ++q) y[q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const bf16x8 w = Wf[(layer(5).ffrag + m) * WAVE + lane];
#pragma unroll
            for (int q = 0; q < INF_NB; ++q) y[q] = mfma(w, H[q][m].v, y[q]);
        }
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < INF_NB; ++q) {
                const uint32_t i = b * PER + q * 16 + j;
                if (i >= n) continue;
                if (A.out) {
                    A.out[3 * (size_t)i] = y[q][0], A.out[3 * (size_t)i + 1] = y[q][1], A.out[3 * (size_t)i + 2] = y[q][2];
                    continue;
                }
                const float *qq = A.queries + 16 * (size_t)i;
                const uint32_t slot = __float_as_uint(qq[13]);
                const float r = qq[10] * fmaxf(y[q][0], 0.f), gg = qq[11] * fmaxf(y[q][1], 0.f),
                            bb = qq[12] * fmaxf(y[q][2], 0.f);
                if (A.targets) {
                    A.targets[slot] = make_float4(r, gg, bb, 1.f);
                } else {
                    float *c = A.contrib + 3 * (size_t)slot;
                    c[0] = r, c[1] = gg, c[2] = bb;
                }
            }
        }
    }
}

// Training pass (forward + relative-L2 loss + backward) over n_pad samples
// (a multiple of 64; samples >= n or with targets[i].w == 0 carry no
// gradient).  Writes the encoded inputs xin[c][i] (192 channels), the
// activations act[L][c][i] (L = 0..4) and the output gradients dz[L][c][i]
// (L = 0..4 with 64 channels, L = 5 with 16), all bf16, for the
// weight-gradient kernel, and one loss partial per wave.
struct TrainArgs {
    const float *feats;      // n x 10
    const float4 *targets;   // n x {rgb, valid}
    uint32_t n, n_pad;
    const uint32_t *count;   // valid samples (device)
    uint16_t *xin;           // 192 x n_pad
    uint16_t *act;           // 5 x 64 x n_pad
    uint16_t *dz;            // (5 x 64 + 16) x n_pad
    float *loss_part;        // n_pad / 16
};

static __global__ __launch_bounds__(256) void k_mlp_train(const bf16x8 *__restrict__ Wf, const bf16x8 *__restrict__ Wb,
                                                   TrainArgs A) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t j = lane & 15, g = lane >> 4;
    const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (b * 16 >= A.n_pad) return;
    const uint32_t i = b * 16 + j;
    // A row without a target (w == 0: a training pixel whose path did not
    // reach the cache this frame) holds whatever an earlier frame or call
    // left in the feature buffer.  It is trained on zero features: its
    // output gradient is zero, but stale activations that are not finite
    // would still turn its zero terms of the weight gradient into NaN.
    const bool ok = i < A.n && A.targets[i].w != 0.f;
    float f[N_FEAT];
#pragma unroll
    for (int k = 0; k < N_FEAT; ++k) f[k] = ok ? A.feats[N_FEAT * (size_t)i + k] : 0.f;
    Frag X[6], H[5][2];
    encode_frags(f, g, X);
    const f32x4 y = forward16(Wf, X, H, lane);
    const size_t NP = A.n_pad;
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int s = 0; s < 8; ++s) A.xin[(size_t)kchan(m, (int)g, s) * NP + i] = X[m].h[s];
#pragma unroll
    for (int L = 0; L < 5; ++L)
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int s = 0; s < 8; ++s) A.act[((size_t)L * 64 + kchan(m, (int)g, s)) * NP + i] = H[L][m].h[s];

    // loss and output gradient (lanes 0..15 hold the outputs)
    const float cnt = (float)max(*A.count, 1u);
    float dy[3] = {0.f, 0.f, 0.f}, loss = 0.f;
    if (g == 0 && ok) {
        const float4 t = A.targets[i];
        if (t.w != 0.f) {
            const float tt[3] = {t.x, t.y, t.z};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const float d = y[r] - tt[r];
                const float den = y[r] * y[r] + LOSS_EPS;
                loss += d * d / den;
                dy[r] = 2.f * d / den / (3.f * cnt);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) A.loss_part[b] = loss;

    // dz of the output layer: rows 0..15 (rows 3.. are zero)
    Frag D;
    D.u = make_uint4(0u, 0u, 0u, 0u);
    if (g == 0) {
        D.h[0] = bf16_bits(dy[0]), D.h[1] = bf16_bits(dy[1]), D.h[2] = bf16_bits(dy[2]);
    }
    uint16_t *dz5 = A.dz + (size_t)5 * 64 * NP;
#pragma unroll
    for (int r = 0; r < 4; ++r) dz5[(size_t)(4 * g + r) * NP + i] = D.h[r];

    // layer 5 backward: dH4 = W5^T dz5 (one reduction block of 32 outputs)
    Frag G[2];
    {
        f32x4 acc[4];
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) acc[ib] = mfma(Wb[(layer(5).bfrag + ib) * WAVE + lane], D.v, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                G[m].h[r] = H[4][m].h[r] ? bf16_bits(acc[2 * m][r]) : (uint16_t)0;
                G[m].h[4 + r] = H[4][m].h[4 + r] ? bf16_bits(acc[2 * m + 1][r]) : (uint16_t)0;
            }
    }
#pragma unroll
    for (int L = 4; L >= 0; --L) {
        // G holds dz of layer L (mask of layer L's ReLU applied)
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int s = 0; s < 8; ++s) A.dz[((size_t)L * 64 + kchan(m, (int)g, s)) * NP + i] = G[m].h[s];
        if (L == 0) break;
        f32x4 acc[4];
        const Layer Y = layer(L);
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            acc[ib] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < 2; ++m) acc[ib] = mfma(Wb[(Y.bfrag + ib * 2 + m) * WAVE + lane], G[m].v, acc[ib]);
        }
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                G[m].h[r] = H[L - 1][m].h[r] ? bf16_bits(acc[2 * m][r]) : (uint16_t)0;
                G[m].h[4 + r] = H[L - 1][m].h[4 + r] ? bf16_bits(acc[2 * m + 1][r]) : (uint16_t)0;
            }
    }
}

// Weight gradients: one wave per (16x16 tile, sample chunk); the reduction
// over the chunk's samples runs on the matrix cores with samples as the k
// dimension.  partial[chunk][p] holds the chunk's sum for parameter p.
struct WgradArgs {
    uint32_t n_pad, chunk;  // chunk: samples per chunk, a multiple of 32
    const uint16_t *xin, *act, *dz;
    float *partial;
};

static __global__ __launch_bounds__(64) void k_mlp_wgrad(WgradArgs A) {
    const uint32_t lane = threadIdx.x;
    const uint32_t tile = blockIdx.x, ch = blockIdx.y;
    int L = 0;
    while (L < 5 && (int)tile >= layer(L + 1).wtile) ++L;
    const Layer Y = layer(L);
    const int t = (int)tile - Y.wtile, cb_n = Y.in / 16;
    const int ob = t / cb_n, cb = t % cb_n;
    const uint32_t j = lane & 15, g = lane >> 4;
    const size_t NP = A.n_pad;
    const uint16_t *dzL = A.dz + (size_t)L * 64 * NP;
    const uint16_t *hin = L > 0 ? A.act + (size_t)(L - 1) * 64 * NP : A.xin;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint32_t s0 = ch * A.chunk, s1 = min(s0 + A.chunk, A.n_pad);
    for (uint32_t q = s0; q < s1; q += 32) {
        const uint32_t sb = q + 8 * g;  // this lane's 8 samples
        Frag a, bb;
        a.u = *reinterpret_cast<const uint4 *>(dzL + (size_t)(ob * 16 + j) * NP + sb);
        bb.u = *reinterpret_cast<const uint4 *>(hin + (size_t)(cb * 16 + j) * NP + sb);
        acc = mfma(a.v, bb.v, acc);
    }
    float *out = A.partial + (size_t)ch * N_PARAMS + Y.pofs;
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(size_t)(ob * 16 + 4 * g + r) * Y.in + cb * 16 + j] = acc[r];
}

// Sum of the chunk partials: the mean-loss gradient of the batch.
static __global__ void k_mlp_reduce_grad(const float *partial, uint32_t n_chunks, float *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N_PARAMS) return;
    float s = 0.f;
    for (uint32_t c = 0; c < n_chunks; ++c) s += partial[(size_t)c * N_PARAMS + p];
    grad[p] = s;
}

struct AdamArgs {
    float *master, *m, *v;
    const float *partial;
    uint32_t n_chunks;
    const uint32_t *step;  // steps taken so far (device)
    float lr, beta1, beta2, eps;
    uint16_t *ffrag, *bfrag;
};

static __global__ void k_mlp_adam(AdamArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N_PARAMS) return;
    float gsum = 0.f;
    for (uint32_t c = 0; c < A.n_chunks; ++c) gsum += A.partial[(size_t)c * N_PARAMS + p];
    const float t = (float)(*A.step + 1u);
    const float m = A.beta1 * A.m[p] + (1.f - A.beta1) * gsum;
    const float v = A.beta2 * A.v[p] + (1.f - A.beta2) * gsum * gsum;
    A.m[p] = m;
    A.v[p] = v;
    const float mh = m / (1.f - powf(A.beta1, t));
    const float vh = v / (1.f - powf(A.beta2, t));
    const float w = A.master[p] - A.lr * mh / (sqrtf(vh) + A.eps);
    A.master[p] = w;
    int fpos, bpos;
    param_slots(p, fpos, bpos);
    const uint16_t b = bf16_bits(w);
    A.ffrag[fpos] = b;
    if (bpos >= 0) A.bfrag[bpos] = b;
}

// After a step: sum the loss partials in a fixed order, count the step and
// update the running loss that gates the cache.
struct FinishArgs {
    const float *loss_part;
    uint32_t n_part;
    const uint32_t *count;
    uint32_t *step;
    float *loss_last, *loss_ema;
    uint32_t *cache_on;
    float threshold;
    uint32_t min_steps;
    float *hist;              // optional: {loss, cache_on} per step from hist_base on
    uint32_t hist_base, hist_cap;
};

static __global__ __launch_bounds__(256) void k_mlp_finish(FinishArgs A) {
    __shared__ float red[256];
    float s = 0.f;
    for (uint32_t k = threadIdx.x; k < A.n_part; k += 256) s += A.loss_part[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float cnt = (float)max(*A.count, 1u);
        const float l = red[0] / (3.f * cnt);
        const uint32_t st = *A.step + 1u;
        *A.step = st;
        *A.loss_last = l;
        const float e = st == 1u ? l : 0.9f * *A.loss_ema + 0.1f * l;
        *A.loss_ema = e;
        const uint32_t on = (st >= A.min_steps && e < A.threshold) ? 1u : 0u;
        if (A.cache_on) *A.cache_on = on;
        const uint32_t k = st - 1u - A.hist_base;
        if (A.hist && k < A.hist_cap) A.hist[2 * k] = l, A.hist[2 * k + 1] = (float)on;
    }
}

}  // namespace prm
