// Learned rate majorant: an 8 -> 64 -> 64 -> 1 perceptron (SiLU hidden layers, softplus output) in BF16
// with FP32 accumulation on the gfx950 matrix cores. No counterpart in the reference: it bounds an
// owner's total bath rate so that event times can be drawn by Poisson thinning, with every accepted
// event checked against the exact FP64 rates (col_bath_rate.cu:13-129 forms those rates).
//
// Layout of one wavefront's work: 16 samples per tile, every layer in transposed form so that the
// output fragment of one product is the B operand of the next one without lane moves
// (v_mfma_f32_16x16x16_bf16: lane (g, i) holds A[i][4g..4g+3] and B[4g..4g+3][i], and gets
// D[4g..4g+3][i], with g = lane / 16 and i = lane % 16):
//   layer 1   A1^T = W1^T X^T + b1       A fragment W1[4g+v][16c+i] (zero for inputs 8..15)
//   layer 2   A2^T = W2^T H1^T + b2      A fragment W2[16c+4g+v][16m+i], B = SiLU(A1^T) fragment c
//   layer 3   z = w3 . SiLU(A2) + b3     per lane over its 16 hidden values, then across g
#ifndef GDX_MLP_H
#define GDX_MLP_H

#include <hip/hip_runtime.h>

namespace gdx {
namespace mlp {

constexpr int F = 8;                  // input features
constexpr int H = 64;                 // hidden width
// parameter vector (FP32 master copy): W1[F][H] (input-major), b1[H], W2[H][H], b2[H], w3[H], b3
constexpr int P_W1 = 0, P_B1 = P_W1 + F*H, P_W2 = P_B1 + H, P_B2 = P_W2 + H*H, P_W3 = P_B2 + H,
              P_B3 = P_W3 + H, P_COUNT = P_B3 + 1;
constexpr int P_LOSS = P_COUNT;       // loss sum slot of a gradient partial
constexpr int P_STRIDE = P_COUNT + 3; // partial row stride (floats)

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// BF16 fragments and FP32 vectors of one parameter set, rebuilt after every parameter change
struct pack
{
    v4s w1a[4][64];        // [c] lane (g, i): W1[4g+v][16c+i]
    v4s w2a[16][64];       // [4m+c] lane (g, i): W2[16c+4g+v][16m+i]
    v4s w2b[16][64];       // [4a+b] lane (g, i): W2[16a+i][16b+4g+v]   (backward: dH1^T = W2 dA2^T)
    float b1[H], b2[H], w3[H], b3;
    float mu[F], inv_sigma[F];
};

// round to nearest even (finite inputs): the hardware conversion (v_cvt_pk_bf16_f32)
__device__ __forceinline__ short bf16_rn(float f) { return __builtin_bit_cast(short, static_cast<__bf16>(f)); }
__device__ __forceinline__ float bf16_f(short b) { return __uint_as_float(static_cast<unsigned int>(static_cast<unsigned short>(b)) << 16); }

// the hardware exponential and reciprocal (about 1 ulp each): far below the BF16 rounding that
// follows every hidden activation
__device__ __forceinline__ float sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
__device__ __forceinline__ float silu(float x) { return x*sigmoid(x); }
__device__ __forceinline__ float silu_grad(float x)
{
    float s = sigmoid(x);
    return s*(1.0f + x*(1.0f - s));
}
__device__ __forceinline__ float softplus(float z) { return fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// B operand of layer 1 for sample row j of this lane: standardized inputs 4g..4g+3, zero past F
__device__ __forceinline__ v4s input_fragment(const pack *P, const float *x, bool valid)
{
    const int g = (threadIdx.x & 63) >> 4;
    v4s r = {0, 0, 0, 0};
    if (valid && g < F/4)
    {
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = bf16_rn((x[4*g + v] - P->mu[4*g + v])*P->inv_sigma[4*g + v]);
    }
    return r;
}

struct forward_state
{
    v4f a1[4], a2[4];      // pre-activations, C layout: lane (g, j) -> [j][16c + 4g + v]
    v4s h1[4];             // BF16 SiLU(a1): the layer-2 B fragments
};

// one 16-sample tile; every lane of the wavefront must be active. Returns z for sample lane % 16.
__device__ __forceinline__ float tile_forward(const pack *P, v4s x, forward_state &f)
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        v4f acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = P->b1[16*c + 4*g + v];
        acc = mfma(P->w1a[c][lane], x, acc);
        f.a1[c] = acc;
        v4s h;
#pragma unroll
        for (int v = 0; v < 4; ++v) h[v] = bf16_rn(silu(acc[v]));
        f.h1[c] = h;
    }
    float part = 0.0f;
#pragma unroll
    for (int m = 0; m < 4; ++m)
    {
        v4f acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = P->b2[16*m + 4*g + v];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc = mfma(P->w2a[4*m + c][lane], f.h1[c], acc);
        f.a2[m] = acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) part += P->w3[16*m + 4*g + v]*silu(acc[v]);
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + P->b3;
}

}  // namespace mlp
}  // namespace gdx

#endif  // GDX_MLP_H
