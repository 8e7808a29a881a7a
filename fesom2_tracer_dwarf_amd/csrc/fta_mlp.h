// fta_mlp.h -- the learned limiter's network on gfx950 matrix cores.
//
// MLP 10 -> 64 -> 64 -> 1 (ReLU, ReLU, sigmoid), BF16 weights and
// activations into v_mfma_f32_16x16x16_bf16 with FP32 accumulation.
//
// Activations are kept transposed (hidden x rows): a 16-row tile of inputs is
// the B operand and the weights are the A operand, so the accumulator layout
// of one layer (lane l holds D[4*(l/16)+e][l%16], e = 0..3) is exactly the B
// operand layout of the next (lane l holds B[4*(l/16)+e][l%16]) and the whole
// forward pass chains 4 + 16 + 4 MFMAs per 16 rows without moving data
// between lanes. Operand layouts of v_mfma_f32_16x16x16_bf16 (wave64):
//   A[i][k]: lane i + 16*(k/4), element k%4
//   B[k][j]: lane j + 16*(k/4), element k%4
//   D[i][j]: lane j + 16*(i/4), element i%4
#pragma once

#include <hip/hip_runtime.h>

namespace fta {

constexpr int kMlpIn = 10;     // features per edge level
constexpr int kMlpInPad = 16;  // padded to one MFMA K step
constexpr int kMlpHid = 64;
// FP32 master parameters, flattened: W1[f][h] (10x64), b1 (64), W2[h1][h2]
// (64x64), b2 (64), w3 (64), b3 (1)
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHid;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHid;
constexpr int kMlpParams = kMlpOffB3 + 1;

typedef short mlp_v4s __attribute__((ext_vector_type(4)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));

// device copies of the weights in the operand orders the kernels read
struct MlpDev {
    const unsigned short *w1t;  // [64][16]: W1^T, features 10..15 zero (bf16 bits)
    const unsigned short *w2t;  // [64][64]: W2^T (h2 x h1)
    const unsigned short *w2;   // [64][64]: W2 (h1 x h2), for the backward pass
    const unsigned short *w3;   // [64]
    const float *b1, *b2;       // [64]
    const float *b3;            // [1]
};

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(unsigned short h)
{
    const unsigned u = (unsigned)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ mlp_v4s ld_bf4(const unsigned short *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    mlp_v4s v;
    v[0] = (short)(u.x & 0xffffu); v[1] = (short)(u.x >> 16);
    v[2] = (short)(u.y & 0xffffu); v[3] = (short)(u.y >> 16);
    return v;
}

__device__ __forceinline__ mlp_v4s pack_bf4(mlp_v4f x)
{
    mlp_v4s v;
    v[0] = (short)f2bf(x[0]); v[1] = (short)f2bf(x[1]);
    v[2] = (short)f2bf(x[2]); v[3] = (short)f2bf(x[3]);
    return v;
}

__device__ __forceinline__ mlp_v4f mfma16(mlp_v4s a, mlp_v4s b, mlp_v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// weights of one wave in A-operand registers
struct MlpRegs {
    mlp_v4s w1[4];       // W1^T tile t: rows 16t..16t+15, K = features
    mlp_v4s w2[4][4];    // W2^T tile (t2, c): rows 16t2.., K = 16c..16c+15
    mlp_v4s w3[4];       // w3 padded to a 16-row A tile (row 0), K chunk c
    float b1[4][4], b2[4][4];  // bias of D rows 16t + 4*(l/16) + e
    float b3;
};

__device__ __forceinline__ void mlp_load(const MlpDev &w, MlpRegs &r)
{
    const int l = (int)(threadIdx.x & 63), i = l & 15, g = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        r.w1[t] = ld_bf4(w.w1t + (16 * t + i) * kMlpInPad + 4 * g);
#pragma unroll
        for (int c = 0; c < 4; ++c) r.w2[t][c] = ld_bf4(w.w2t + (16 * t + i) * kMlpHid + 16 * c + 4 * g);
        const mlp_v4s z = {0, 0, 0, 0};
        r.w3[t] = i == 0 ? ld_bf4(w.w3 + 16 * t + 4 * g) : z;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            r.b1[t][e] = w.b1[16 * t + 4 * g + e];
            r.b2[t][e] = w.b2[16 * t + 4 * g + e];
        }
    }
    r.b3 = w.b3[0];
}

// forward pass of one 16-row tile whose inputs are the B operand x (lane l:
// row l%16, features 4*(l/16)..+3). Returns z3 (pre-sigmoid) in lanes 0..15
// (row = lane); z1/z2 pre-activations and a1/a2 bf16 activations are kept
// for the backward pass.
struct MlpTile {
    mlp_v4f z1[4], z2[4];
    mlp_v4s a1[4], a2[4];
    float z3;
};

__device__ __forceinline__ void mlp_forward(const MlpRegs &r, mlp_v4s x, MlpTile &T)
{
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mlp_v4f acc = {r.b1[t][0], r.b1[t][1], r.b1[t][2], r.b1[t][3]};
        acc = mfma16(r.w1[t], x, acc);
        T.z1[t] = acc;
        mlp_v4f a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = acc[e] > 0.0f ? acc[e] : 0.0f;
        T.a1[t] = pack_bf4(a);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mlp_v4f acc = {r.b2[t][0], r.b2[t][1], r.b2[t][2], r.b2[t][3]};
#pragma unroll
        for (int c = 0; c < 4; ++c) acc = mfma16(r.w2[t][c], T.a1[c], acc);
        T.z2[t] = acc;
        mlp_v4f a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = acc[e] > 0.0f ? acc[e] : 0.0f;
        T.a2[t] = pack_bf4(a);
    }
    mlp_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = mfma16(r.w3[c], T.a2[c], acc);
    T.z3 = acc[0] + r.b3;  // D row 0: lanes 0..15, element 0
}

__device__ __forceinline__ float mlp_sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

// sum over the 16 lanes that share l/16 (the rows of a tile)
__device__ __forceinline__ float sum16(float v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// One epoch-slice of training: every wave takes 16-row tiles (rows with a
// negative target are padding and count zero), runs the forward pass, the
// backward pass of the loss sum((sigmoid(z3) - y)^2) * scale, and accumulates
// parameter gradients in registers: dA1 = W2 . dZ2 and both weight gradients
// on the matrix cores (the weight gradients need the activation tiles
// transposed, which goes through LDS). The block then adds its gradients and
// loss into grad[kMlpParams] / loss[0] with float atomics (training is not
// bitwise reproducible across runs; inference is).
// Loop trip counts are block-uniform (__syncthreads inside).
__global__ __launch_bounds__(256) void k_mlp_train(MlpDev w, const unsigned short *__restrict__ X,
                                                   const float *__restrict__ y, long long rows, int iters,
                                                   float scale, float *__restrict__ grad, float *__restrict__ loss)
{
    __shared__ float sA[4][kMlpHid][17];
    __shared__ float sB[4][kMlpHid][17];
    __shared__ float sX[4][16][17];
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63), i = l & 15, g = l >> 4;
    MlpRegs r;
    mlp_load(w, r);
    mlp_v4s w2n[4][4];
    float w3f[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int c = 0; c < 4; ++c) w2n[t][c] = ld_bf4(w.w2 + (16 * t + i) * kMlpHid + 16 * c + 4 * g);
#pragma unroll
        for (int e = 0; e < 4; ++e) w3f[t][e] = bf2f(w.w3[16 * t + 4 * g + e]);
    }
    mlp_v4f gW2[4][4], gW1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.0f, lsum = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        gW1[t] = mlp_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < 4; ++c) gW2[t][c] = mlp_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int e = 0; e < 4; ++e) { gb1[t][e] = 0.0f; gb2[t][e] = 0.0f; gw3[t][e] = 0.0f; }
    }
    const mlp_v4s zero4 = {0, 0, 0, 0};
    for (int it = 0; it < iters; ++it) {
        const long long tile = ((long long)it * gridDim.x + blockIdx.x) * 4 + wv;
        const long long row = tile * 16 + i;
        const float tgt = row < rows ? y[row] : -1.0f;
        const bool valid = tgt >= 0.0f;
        const mlp_v4s x = valid ? ld_bf4(X + row * kMlpInPad + 4 * g) : zero4;
        MlpTile T;
        mlp_forward(r, x, T);
        float dz3 = 0.0f;
        if (g == 0 && valid) {
            const float pr = mlp_sigmoid(T.z3), d = pr - tgt;
            lsum += d * d;
            dz3 = scale * 2.0f * d * pr * (1.0f - pr);
        }
        gb3 += dz3;
        const float dzb = __shfl(dz3, i, 64);
        mlp_v4f dz2[4];
        mlp_v4s dz2b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                dz2[c][e] = T.z2[c][e] > 0.0f ? w3f[c][e] * dzb : 0.0f;
                gw3[c][e] += dzb * bf2f((unsigned short)T.a2[c][e]);
            }
            dz2b[c] = pack_bf4(dz2[c]);
#pragma unroll
            for (int e = 0; e < 4; ++e) gb2[c][e] += dz2[c][e];
        }
        mlp_v4s dz1b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            mlp_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < 4; ++c) acc = mfma16(w2n[t][c], dz2b[c], acc);
            mlp_v4f d1;
#pragma unroll
            for (int e = 0; e < 4; ++e) d1[e] = T.z1[t][e] > 0.0f ? acc[e] : 0.0f;
            dz1b[t] = pack_bf4(d1);
#pragma unroll
            for (int e = 0; e < 4; ++e) gb1[t][e] += d1[e];
        }
        // transposes for the weight gradients
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                sA[wv][16 * t + 4 * g + e][i] = bf2f((unsigned short)T.a1[t][e]);
                sB[wv][16 * t + 4 * g + e][i] = bf2f((unsigned short)dz2b[t][e]);
            }
#pragma unroll
        for (int e = 0; e < 4; ++e) sX[wv][i][4 * g + e] = bf2f((unsigned short)x[e]);
        __syncthreads();
        // dW2[h1][h2] += sum_r a1[h1][r] dz2[h2][r]
#pragma unroll
        for (int t1 = 0; t1 < 4; ++t1) {
            mlp_v4f av;
#pragma unroll
            for (int e = 0; e < 4; ++e) av[e] = sA[wv][16 * t1 + i][4 * g + e];
            const mlp_v4s a = pack_bf4(av);
#pragma unroll
            for (int t2 = 0; t2 < 4; ++t2) {
                mlp_v4f bv;
#pragma unroll
                for (int e = 0; e < 4; ++e) bv[e] = sB[wv][16 * t2 + i][4 * g + e];
                gW2[t1][t2] = mfma16(a, pack_bf4(bv), gW2[t1][t2]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) sB[wv][16 * t + 4 * g + e][i] = bf2f((unsigned short)dz1b[t][e]);
        __syncthreads();
        // dW1[f][h1] += sum_r x[r][f] dz1[h1][r]
        {
            mlp_v4f xv;
#pragma unroll
            for (int e = 0; e < 4; ++e) xv[e] = sX[wv][4 * g + e][i];
            const mlp_v4s a = pack_bf4(xv);
#pragma unroll
            for (int t1 = 0; t1 < 4; ++t1) {
                mlp_v4f bv;
#pragma unroll
                for (int e = 0; e < 4; ++e) bv[e] = sB[wv][16 * t1 + i][4 * g + e];
                gW1[t1] = mfma16(a, pack_bf4(bv), gW1[t1]);
            }
        }
        __syncthreads();
    }
    // per-lane partial sums over this lane's rows -> over the 16 rows of the group
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float s1 = sum16(gb1[t][e]), s2 = sum16(gb2[t][e]), s3 = sum16(gw3[t][e]);
            if (i == 0) {
                atomicAdd(grad + kMlpOffB1 + 16 * t + 4 * g + e, s1);
                atomicAdd(grad + kMlpOffB2 + 16 * t + 4 * g + e, s2);
                atomicAdd(grad + kMlpOffW3 + 16 * t + 4 * g + e, s3);
            }
        }
    const float sb3 = sum16(gb3), sl = sum16(lsum);
    if (l == 0) {
        atomicAdd(grad + kMlpOffB3, sb3);
        atomicAdd(loss, sl);
    }
#pragma unroll
    for (int t1 = 0; t1 < 4; ++t1) {
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                atomicAdd(grad + kMlpOffW2 + (16 * t1 + 4 * g + e) * kMlpHid + 16 * t2 + i, gW2[t1][t2][e]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int f = 4 * g + e;
            if (f < kMlpIn) atomicAdd(grad + kMlpOffW1 + f * kMlpHid + 16 * t1 + i, gW1[t1][e]);
        }
    }
}

// Adam on the FP32 master parameters, then their bf16 operand copies
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ m1,
                                                  float *__restrict__ m2, const float *__restrict__ g, float lr,
                                                  float c1, float c2)
{
    const int k = (int)(blockIdx.x * 256 + threadIdx.x);
    if (k >= kMlpParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * m1[k] + (1.0f - b1) * g[k];
    const float v = b2 * m2[k] + (1.0f - b2) * g[k] * g[k];
    m1[k] = m;
    m2[k] = v;
    p[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

struct MlpMut {
    unsigned short *w1t, *w2t, *w2, *w3;
    float *b1, *b2, *b3;
};

__global__ __launch_bounds__(256) void k_mlp_pack(const float *__restrict__ p, MlpMut w)
{
    const int k = (int)(blockIdx.x * 256 + threadIdx.x);
    if (k < kMlpHid * kMlpInPad) {
        const int h = k / kMlpInPad, f = k % kMlpInPad;
        w.w1t[k] = f < kMlpIn ? f2bf(p[kMlpOffW1 + f * kMlpHid + h]) : (unsigned short)0;
    }
    if (k < kMlpHid * kMlpHid) {
        const int a = k / kMlpHid, b = k % kMlpHid;
        w.w2t[k] = f2bf(p[kMlpOffW2 + b * kMlpHid + a]);  // W2^T[h2][h1]
        w.w2[k] = f2bf(p[kMlpOffW2 + k]);                 // W2[h1][h2]
    }
    if (k < kMlpHid) {
        w.w3[k] = f2bf(p[kMlpOffW3 + k]);
        w.b1[k] = p[kMlpOffB1 + k];
        w.b2[k] = p[kMlpOffB2 + k];
    }
    if (k == 0) w.b3[0] = p[kMlpOffB3];
}

__global__ __launch_bounds__(256) void k_fill_f32(float *__restrict__ a, long long n, float v)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) a[k] = v;
}

// inference of arbitrary rows (bf16 [rows][16]) -> sigmoid outputs; one
// 16-row tile per wave
// ---------------------------------------------------------------------------
// Inference-only forward pass (the step's limiter and k_mlp_infer).
// Layer 1 as above (K = 16 features). Layer 2 on v_mfma_f32_16x16x32_bf16:
// the order of K inside one instruction is free as long as A and B agree, so
// the B operand of K block c (hidden 32c..32c+31) is simply the two layer-1
// accumulator tiles 2c and 2c+1 of the lane (lane l, element e: hidden
// 32c + 16*(e/4) + 4*(l/16) + e%4), and the weights are loaded in that same
// order. Layer 3 (64 -> 1) takes two more of these instructions with w3 as
// row 0 of the A operand (rows 1-15 zero) and the layer-2 activations as B,
// permuted the same way: its output is D row 0, lanes 0..15, with no
// cross-lane sum. Roundings are those of the training pass: bf16 inputs,
// weights and activations, FP32 sums.
// ---------------------------------------------------------------------------
typedef __bf16 mlp_v8b __attribute__((ext_vector_type(8)));
typedef short mlp_v8s __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_v2b __attribute__((ext_vector_type(2)));
typedef float mlp_v2f __attribute__((ext_vector_type(2)));

__device__ __forceinline__ mlp_v8b cat8(mlp_v4s a, mlp_v4s b)
{
    const mlp_v8s s = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
    return __builtin_bit_cast(mlp_v8b, s);
}

// round to bf16 (v_cvt_pk_bf16_f32: nearest even) and apply the ReLU on the
// bf16 bits (v_pk_max_i16 with 0: a set sign bit is a negative int16), two
// values per instruction each; equal to bf16(max(x, 0)) for finite x
typedef short mlp_v2s __attribute__((ext_vector_type(2)));
__device__ __forceinline__ mlp_v2b relu_bf2(float a, float b)
{
    const mlp_v2f v = {a, b};
    const mlp_v2s h = __builtin_bit_cast(mlp_v2s, __builtin_convertvector(v, mlp_v2b));
    const mlp_v2s z = {0, 0};
    return __builtin_bit_cast(mlp_v2b, __builtin_elementwise_max(h, z));
}
__device__ __forceinline__ mlp_v4s relu_bf4(mlp_v4f acc)
{
    const mlp_v2b lo = relu_bf2(acc[0], acc[1]), hi = relu_bf2(acc[2], acc[3]);
    const unsigned ul = __builtin_bit_cast(unsigned, lo), uh = __builtin_bit_cast(unsigned, hi);
    const mlp_v4s v = {(short)(ul & 0xffffu), (short)(ul >> 16), (short)(uh & 0xffffu), (short)(uh >> 16)};
    return v;
}

struct MlpInfer {
    mlp_v4s w1[4];        // W1^T tile t (A operand, K = features)
    mlp_v8b w2[4][2];     // W2^T row tile t2, K block c (A operand, permuted K)
    float b1[4][4], b2[4][4];
    mlp_v8b w3[2];        // w3 as row 0 of the A operand, K block c (permuted K), zero in rows 1-15
    float b3;
};

__device__ __forceinline__ void mlp_infer_load(const MlpDev &w, MlpInfer &r)
{
    const int l = (int)(threadIdx.x & 63), i = l & 15, g = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        r.w1[t] = ld_bf4(w.w1t + (16 * t + i) * kMlpInPad + 4 * g);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned short *row = w.w2t + (16 * t + i) * kMlpHid + 32 * c + 4 * g;
            r.w2[t][c] = cat8(ld_bf4(row), ld_bf4(row + 16));
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            r.b1[t][e] = w.b1[16 * t + 4 * g + e];
            r.b2[t][e] = w.b2[16 * t + 4 * g + e];
        }
    }
    const mlp_v4s z4 = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 2; ++c)
        r.w3[c] = i == 0 ? cat8(ld_bf4(w.w3 + 32 * c + 4 * g), ld_bf4(w.w3 + 32 * c + 16 + 4 * g)) : cat8(z4, z4);
    r.b3 = w.b3[0];
}

// z3 (pre-sigmoid) of tile row l%16, in lanes 0..15, for NTILE tiles at once:
// each layer's instructions of the tiles are interleaved, so that one tile's
// matrix instructions overlap another's epilogue
template <int NTILE>
__device__ __forceinline__ void mlp_infer_tiles(const MlpInfer &r, const mlp_v4s (&x)[NTILE], float (&z3)[NTILE])
{
    mlp_v4s a1[NTILE][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const mlp_v4f bias = {r.b1[t][0], r.b1[t][1], r.b1[t][2], r.b1[t][3]};
#pragma unroll
        for (int u = 0; u < NTILE; ++u) a1[u][t] = relu_bf4(mfma16(r.w1[t], x[u], bias));
    }
    mlp_v8b b0[NTILE], b1[NTILE];
#pragma unroll
    for (int u = 0; u < NTILE; ++u) {
        b0[u] = cat8(a1[u][0], a1[u][1]);
        b1[u] = cat8(a1[u][2], a1[u][3]);
    }
    mlp_v4s a2[NTILE][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mlp_v4f acc[NTILE];
#pragma unroll
        for (int u = 0; u < NTILE; ++u) {
            acc[u] = mlp_v4f{r.b2[t][0], r.b2[t][1], r.b2[t][2], r.b2[t][3]};
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.w2[t][0], b0[u], acc[u], 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < NTILE; ++u) {
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.w2[t][1], b1[u], acc[u], 0, 0, 0);
            a2[u][t] = relu_bf4(acc[u]);
        }
    }
    mlp_v4f z[NTILE];
#pragma unroll
    for (int u = 0; u < NTILE; ++u) {
        z[u] = mlp_v4f{0.0f, 0.0f, 0.0f, 0.0f};
        z[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.w3[0], cat8(a2[u][0], a2[u][1]), z[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < NTILE; ++u) {
        z[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(r.w3[1], cat8(a2[u][2], a2[u][3]), z[u], 0, 0, 0);
        z3[u] = z[u][0] + r.b3;
    }
}

__device__ __forceinline__ float mlp_infer(const MlpInfer &r, mlp_v4s x)
{
    const mlp_v4s xs[1] = {x};
    float z[1];
    mlp_infer_tiles<1>(r, xs, z);
    return z[0];
}

__device__ __forceinline__ float mlp_sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__global__ __launch_bounds__(256) void k_mlp_infer(MlpDev w, const unsigned short *__restrict__ X, long long rows,
                                                   float *__restrict__ out)
{
    const int l = (int)(threadIdx.x & 63), i = l & 15, g = l >> 4;
    const long long tile = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 16 >= rows) return;
    MlpInfer r;
    mlp_infer_load(w, r);
    const long long row = tile * 16 + i;
    const mlp_v4s zero4 = {0, 0, 0, 0};
    const mlp_v4s x = row < rows ? ld_bf4(X + row * kMlpInPad + 4 * g) : zero4;
    const float z = mlp_infer(r, x);
    if (g == 0 && row < rows) out[row] = mlp_sigmoid_fast(z);
}

}  // namespace fta
