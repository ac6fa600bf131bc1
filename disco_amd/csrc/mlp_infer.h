// The learned cons2prim guess network (MLP 8 -> 64x4 -> 2, SiLU): parameter
// layout, input features, the BF16 MFMA inference of one wavefront and the
// host-side packing of its weights.  Shared by the standalone cons2prim
// (c2p.hip) and the GRMHD step's in-loop learned guess (hydro_step.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "phys_grmhd.h"

namespace disco {
namespace mlp {

#define MLPDEV __device__ static inline __attribute__((always_inline))

// ------------------------------------------------------------------ MLP layout
constexpr int NIN = 8, NH = 64, NOUT = 2;
constexpr int OFF_W1 = 0;                      // [NIN][NH]
constexpr int OFF_B1 = OFF_W1 + NIN * NH;      // [NH]
constexpr int OFF_W2 = OFF_B1 + NH;            // [NH][NH]
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NH * NH;
constexpr int OFF_W5 = OFF_B4 + NH;            // [NH][NOUT]
constexpr int OFF_B5 = OFF_W5 + NH * NOUT;
constexpr int NPARAM = OFF_B5 + NOUT;          // 13186
static_assert(NPARAM == 13186, "MLP parameter count");
__host__ __device__ constexpr int w_off(int l) { return l == 0 ? OFF_W1 : l == 1 ? OFF_W2 : l == 2 ? OFF_W3 : l == 3 ? OFF_W4 : OFF_W5; }
__host__ __device__ constexpr int b_off(int l) { return l == 0 ? OFF_B1 : l == 1 ? OFF_B2 : l == 2 ? OFF_B3 : l == 3 ? OFF_B4 : OFF_B5; }
__host__ __device__ constexpr int l_in(int l) { return l == 0 ? NIN : NH; }
__host__ __device__ constexpr int l_out(int l) { return l == 4 ? NOUT : NH; }

// Feature / target standardisation of a trained network.
struct MlpNorm {
    float fmean[NIN], fstd[NIN];
    float tmean[NOUT], tstd[NOUT];
};

// Scale-free features of the Noble 2D system (it depends on s2, e, Q, psi
// and Gamma only).  psi <= Q s2 by Cauchy-Schwarz, so f6 is a cos^2.
MLPDEV void c2p_features(const grmhd::C2PInvariants& v, float* f) {
    // the network is BF16: features in FP32 arithmetic from the FP64 invariants
    const float e = (float)v.e, s2 = (float)v.s2, Q = (float)v.Q, psi = (float)v.psi;
    const float eq = e + Q;
    f[0] = __logf(e);
    f[1] = log1pf(s2);
    f[2] = s2 / (eq * eq);
    f[3] = log1pf(Q);
    f[4] = Q / eq;
    f[5] = log1pf(psi);
    const float qs = Q * s2;
    f[6] = qs > 0.0f ? fminf(psi / qs, 1.0f) : 0.0f;
    f[7] = __logf(fmaxf((float)(v.e - 1.0), 1e-12f));
}

// ------------------------------------------------------------------ MFMA inference
// BF16 weights, transposed per layer and zero-padded to 16-multiples:
// Wt_l[j][k] (out j, in k), K = 16 for layer 0 (8 inputs), N = 16 for the
// output layer (2 outputs).  Activations live in LDS as bf16 rows [s][k].
constexpr int K0P = 16, NOP = 16;
constexpr int WT_L0 = 0;                   // [NH][K0P]
constexpr int WT_L1 = WT_L0 + NH * K0P;    // [NH][NH]
constexpr int WT_L2 = WT_L1 + NH * NH;
constexpr int WT_L3 = WT_L2 + NH * NH;
constexpr int WT_L4 = WT_L3 + NH * NH;     // [NOP][NH]
constexpr int WT_TOTAL = WT_L4 + NOP * NH; // bf16 elements
constexpr int NBIAS = 4 * NH + NOP;

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));


typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

// Two 4-element bf16 fragments as the 8-element operand of
// v_mfma_f32_16x16x32_bf16 (slots 0-3 from lo, 4-7 from hi).
MLPDEV bf16x8_t cat8(short4_t lo, short4_t hi) {
    typedef short short8_t __attribute__((ext_vector_type(8)));
    short8_t v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_bit_cast(bf16x8_t, v);
}

// float4 -> 4 x bf16 with the hardware's round-to-nearest-even convert
// (v_cvt_pk_bf16_f32)
MLPDEV short4_t pack4(float4_t v) {
    bf16x4_t b = __builtin_convertvector(v, bf16x4_t);
    return __builtin_bit_cast(short4_t, b);
}

// SiLU for inference: v_exp_f32 + v_rcp_f32 (the network is BF16)
MLPDEV float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

MLPDEV float4_t silu4(float4_t a) {
    float4_t h = {silu_fast(a[0]), silu_fast(a[1]), silu_fast(a[2]), silu_fast(a[3])};
    return h;
}

// Block-shared network parameters in LDS: transposed bf16 weights and fp32
// biases (zero-padded), loaded once per block.
struct MlpLds {
    unsigned short W[WT_TOTAL];
    float b[NBIAS];
};

MLPDEV void mlp_load(MlpLds& L, const unsigned short* __restrict__ Wt_g, const float* __restrict__ W32) {
    for (int i = threadIdx.x; i < WT_TOTAL / 4; i += blockDim.x)
        reinterpret_cast<short4_t*>(L.W)[i] = reinterpret_cast<const short4_t*>(Wt_g)[i];
    for (int i = threadIdx.x; i < NBIAS; i += blockDim.x) {
        float b = 0.0f;
        if (i < 4 * NH) b = W32[b_off(i / NH) + (i % NH)];
        else if (i - 4 * NH < NOUT) b = W32[OFF_B5 + (i - 4 * NH)];
        L.b[i] = b;
    }
}

// The network on the 64 states of one wavefront, transposed: each layer
// computes Y^T = W^T X^T on the matrix cores (FP32 accumulate).  Output
// layout of the MFMA: lane l receives D[4(l/16)+t][l%16], t = 0..3; with A a
// tile of W^T (output features x input features, from LDS) and B a tile of
// X^T (input features x states), a lane receives output features 4(l/16)+t
// of state l%16, and act[ts][kt] holds features 16kt + 4(l/16) + t of state
// 16ts + l%16.  The input layer (8 features, padded to 16) runs
// v_mfma_f32_16x16x16_bf16 with that fragment as its B operand.  The hidden
// and output layers run gfx950's v_mfma_f32_16x16x32_bf16, whose A and B
// operands carry 8 bf16 values per lane for the same K positions: the B
// operand of K-tile kt is (act[ts][2kt], act[ts][2kt+1]) as the lane already
// holds it, i.e. K positions paired with features 32kt + 4(l/16) + t and
// 32kt + 16 + 4(l/16) + t, and the A operand reads the weights of exactly
// those features (two 4-element LDS reads), so the contraction runs over the
// right pairs whatever order the instruction assigns to K positions and no
// activation crosses lanes.  Only the input (one row of features per lane)
// and the output (two numbers per state) pass through LDS.
MLPDEV void mlp_wave(const MlpLds& L, unsigned short* sF, float* sY, const float* fstd_in, float* out2) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    // lane -> row: this lane's standardised features (bf16), zero padded
    {
        short4_t lo = pack4((float4_t){fstd_in[0], fstd_in[1], fstd_in[2], fstd_in[3]});
        short4_t hi = pack4((float4_t){fstd_in[4], fstd_in[5], fstd_in[6], fstd_in[7]});
        short4_t z = {0, 0, 0, 0};
        short4_t* row = reinterpret_cast<short4_t*>(sF + lane * K0P);
        row[0] = lo;
        row[1] = hi;
        row[2] = z;
        row[3] = z;
    }
    __syncthreads();
    short4_t act[4][4], nxt[4][4];
#pragma unroll
    for (int ts = 0; ts < 4; ++ts) act[ts][0] = *reinterpret_cast<const short4_t*>(sF + (16 * ts + r) * K0P + 4 * g);
    // layer 0: K = 16 (one k tile)
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        short4_t a = *reinterpret_cast<const short4_t*>(L.W + WT_L0 + (16 * to + r) * K0P + 4 * g);
        const float4_t bias = {L.b[16 * to + 4 * g], L.b[16 * to + 4 * g + 1], L.b[16 * to + 4 * g + 2],
                               L.b[16 * to + 4 * g + 3]};
#pragma unroll
        for (int ts = 0; ts < 4; ++ts) {
            float4_t acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[ts][0], bias, 0, 0, 0);
            nxt[ts][to] = pack4(silu4(acc));
        }
    }
    // hidden layers 1..3: K = N = 64, two K-tiles of 32
    const int woff[3] = {WT_L1, WT_L2, WT_L3};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
#pragma unroll
        for (int ts = 0; ts < 4; ++ts)
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) act[ts][kt] = nxt[ts][kt];
#pragma unroll
        for (int to = 0; to < 4; ++to) {
            const float* bb = L.b + (l + 1) * NH + 16 * to + 4 * g;
            float4_t acc[4];
#pragma unroll
            for (int ts = 0; ts < 4; ++ts) acc[ts] = (float4_t){bb[0], bb[1], bb[2], bb[3]};
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) {
                const unsigned short* wr = L.W + woff[l] + (16 * to + r) * NH + 32 * kt + 4 * g;
                const bf16x8_t a = cat8(*reinterpret_cast<const short4_t*>(wr),
                                        *reinterpret_cast<const short4_t*>(wr + 16));
#pragma unroll
                for (int ts = 0; ts < 4; ++ts)
                    acc[ts] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, cat8(act[ts][2 * kt], act[ts][2 * kt + 1]),
                                                                      acc[ts], 0, 0, 0);
            }
#pragma unroll
            for (int ts = 0; ts < 4; ++ts) nxt[ts][to] = pack4(silu4(acc[ts]));
        }
    }
    // output layer: N = 16 (2 used); outputs 0, 1 land in lanes with g = 0
    {
        const float* bb = L.b + 4 * NH + 4 * g;
        float4_t acc[4];
#pragma unroll
        for (int ts = 0; ts < 4; ++ts) acc[ts] = (float4_t){bb[0], bb[1], bb[2], bb[3]};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
            const unsigned short* wr = L.W + WT_L4 + r * NH + 32 * kt + 4 * g;
            const bf16x8_t a = cat8(*reinterpret_cast<const short4_t*>(wr), *reinterpret_cast<const short4_t*>(wr + 16));
#pragma unroll
            for (int ts = 0; ts < 4; ++ts)
                acc[ts] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, cat8(nxt[ts][2 * kt], nxt[ts][2 * kt + 1]),
                                                                  acc[ts], 0, 0, 0);
        }
        if (g == 0) {
#pragma unroll
            for (int ts = 0; ts < 4; ++ts) {
                sY[(16 * ts + r) * 2] = acc[ts][0];
                sY[(16 * ts + r) * 2 + 1] = acc[ts][1];
            }
        }
    }
    __syncthreads();
    out2[0] = sY[lane * 2];
    out2[1] = sY[lane * 2 + 1];
}

// Host: transposed, zero-padded BF16 copy (round to nearest even) of the
// FP32 master weights, the layout mlp_wave reads.
static inline void mlp_pack_bf16(const float* W, unsigned short* T) {
    auto bf = [](float f) {
        unsigned u;
        std::memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    for (int i = 0; i < WT_TOTAL; ++i) T[i] = 0;
    for (int j = 0; j < NH; ++j)
        for (int kk = 0; kk < NIN; ++kk) T[WT_L0 + j * K0P + kk] = bf(W[OFF_W1 + kk * NH + j]);
    const int offs[3] = {WT_L1, WT_L2, WT_L3};
    const int woffs[3] = {OFF_W2, OFF_W3, OFF_W4};
    for (int l = 0; l < 3; ++l)
        for (int j = 0; j < NH; ++j)
            for (int kk = 0; kk < NH; ++kk) T[offs[l] + j * NH + kk] = bf(W[woffs[l] + kk * NH + j]);
    for (int o = 0; o < NOUT; ++o)
        for (int kk = 0; kk < NH; ++kk) T[WT_L4 + o * NH + kk] = bf(W[OFF_W5 + kk * NOUT + o]);
}

}  // namespace mlp
}  // namespace disco
