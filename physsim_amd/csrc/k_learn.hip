// k_learn.hip — learned far-field correction of the Barnes-Hut walk (not in the
// reference; the reference's accepted node is a monopole, forces.js:680-685).
//
// For every (observer i, accepted divided node n) pair the walk lists, a
// 2 x 64 hard-swish MLP predicts the monopole-truncation residual from the node's
// quadrupole moments and the offset normalised by the node size.  The inputs
// and outputs are expressed in the frame of the offset direction u = d/|d|,
// which makes the map rotation-equivariant by construction:
//   f0 = s/r, f1 = eps^2/r^2,
//   f2..f4 = (Q_uu, Q_uv, Q_vv) / (M s^2)        mass quadrupole about the COM
//   f5, f6 = (D_u, D_v) / (A s)                  charge dipole about the COM
//   f7..f9 = (C_uu, C_uv, C_vv) / (A s^2)        charge quadrupole about the COM
// with r^2 = |d|^2 + eps^2, s the node width, M its mass, A its total |charge|,
// v = (-u_y, u_x).  Outputs o = (g_u, g_v, c_u, c_v, p_g, p_c) (times oscale) give
//   F_i  += m_i M (s/r)^2 / r^2 (g_u u + g_v v) + q_i A (s/r) / r^2 (c_u u + c_v v),
//   PE_i += m_i M (s/r)^2 / r p_g + q_i A (s/r) / r p_c
// (the potential residual keeps the energy readout consistent with the forces).
// Network: x = ((f - mean) * istd, 1, 0...) in 16 inputs (input 10 carries
// the first layer's bias), z1 = W1^T x, a1 = act(z1), z2 = W2^T a1 + b2,
// a2 = act(z2), o = W3^T a2 + b3, with the hard-swish act(z) = z clamp((z+3)/6,
// 0, 1).  BF16 operands, FP32 accumulation, on v_mfma_f32_32x32x16_bf16.
//
// MFMA 32x32x16 lane layout (gfx950, lane l, h = l / 32, c = l % 32):
//   A[m][k]: lane holds A[c][8h + j], j = 0..7
//   B[k][n]: lane holds B[8h + j][c]
//   C[m][n]: lane holds C[8(r/4) + 4h + r%4][c] in accumulator r = 0..15.
// All activations are kept transposed (hidden x rows, rows = pairs in the
// N dimension), so a C tile of one layer is the B operand of the next once
// the K order of the next layer's weights is permuted to the C layout:
// K-step t takes accumulators r = 8(t%2) + j of M-block t/2, i.e. hidden unit
//   kperm(t, h, j) = 32(t/2) + 16(t%2) + 8(j/4) + 4h + (j%4).
#include "geon_kernels.h"

namespace geon {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ floatx16 mfma32(const bf16x8& a, const bf16x8& b, const floatx16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx16 mfma32(const halfx8& a, const halfx8& b, const floatx16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__host__ __device__ __forceinline__ int kperm(int t, int h, int j) {
    return 32 * (t >> 1) + 16 * (t & 1) + 8 * (j >> 2) + 4 * h + (j & 3);
}

// Hard-swish, act(z) = z clamp(z/6 + 1/2, 0, 1): three VALU operations per
// hidden value (fma, med3, mul), where a polynomial tanh-GELU took ten; fitted
// to the far-field residuals it reaches the same loss (2 x 64 network, same
// training budget: relative test loss 1.79e-3 against GELU's 1.97e-3).  In the
// inference pass (packed FP16, act8h) it is about half of the vector work.
// Cheaper piecewise forms fit worse: clamp(z, 0, 1) 3.07e-3, its square 3.44e-3,
// ReLU 2.45e-3.
__device__ __forceinline__ float act(float z) { return z * __builtin_amdgcn_fmed3f(fmaf(z, 1.f / 6.f, 0.5f), 0.f, 1.f); }
__device__ __forceinline__ float act_grad(float z) { return z < -3.f ? 0.f : (z > 3.f ? 1.f : fmaf(z, 1.f / 3.f, 0.5f)); }

__device__ __forceinline__ bf16x8 pack8(const float* v) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (__bf16)v[j];
    return r;
}
__device__ __forceinline__ halfx8 pack8h(const float* v) {
    halfx8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (_Float16)v[j];
    return r;
}

// The activation of 8 accumulators in packed FP16 for an FP16 MFMA operand:
// per two values one conversion, one fma whose clamp modifier does the
// [0, 1] clamp, and one multiply (the FP32 form takes three per value).
__device__ __forceinline__ halfx8 act8h(const float* z) {
    const halfx2 k = {(_Float16)(1.f / 6.f), (_Float16)(1.f / 6.f)}, c = {(_Float16)0.5f, (_Float16)0.5f};
    const halfx2 lo = {(_Float16)0.f, (_Float16)0.f}, hi = {(_Float16)1.f, (_Float16)1.f};
    halfx8 r;
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        const halfx2 zz = {(_Float16)z[j], (_Float16)z[j + 1]};
        const halfx2 t = __builtin_elementwise_min(__builtin_elementwise_max(__builtin_elementwise_fma(zz, k, c), lo), hi);
        const halfx2 a = zz * t;
        r[j] = a.x;
        r[j + 1] = a.y;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Features of one (observer, node) pair
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }
__device__ __forceinline__ float rsq_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq_t(double x) { return 1.0 / sqrt(x); }

// sg0, sc0: force scales, sp0, cp0: potential scales (without m_i, q_i).
template <typename T>
__device__ __forceinline__ void pair_features(T dx, T dy, T eps2, T s, T M, const NodeMom& c, T* f, T& ux, T& uy,
                                              T& sg0, T& sc0, T& sp0, T& cp0) {
    const T rho2 = dx * dx + dy * dy;
    const T r2 = rho2 + eps2;
    const T ir = rsq_t(r2);
    const T ir2 = ir * ir;
    const T irho = rsq_t(rho2);
    ux = dx * irho; uy = dy * irho;
    const T sr = s * ir;
    const T is = rcp_t(s);
    const T iM = M > T(0) ? rcp_t(M) * is * is : T(0);
    const T A = (T)c.aq;
    const T iA = A > T(0) ? rcp_t(A) * is : T(0);
    const T uu = ux * ux, vv = uy * uy, uv = ux * uy;
    f[0] = sr;
    f[1] = eps2 * ir2;
    f[2] = (uu * c.qxx + T(2) * uv * c.qxy + vv * c.qyy) * iM;
    f[3] = (-uv * c.qxx + (uu - vv) * c.qxy + uv * c.qyy) * iM;
    f[4] = (vv * c.qxx - T(2) * uv * c.qxy + uu * c.qyy) * iM;
    f[5] = (ux * c.dx + uy * c.dy) * iA;
    f[6] = (-uy * c.dx + ux * c.dy) * iA;
    f[7] = (uu * c.cxx + T(2) * uv * c.cxy + vv * c.cyy) * iA * is;
    f[8] = (-uv * c.cxx + (uu - vv) * c.cxy + uv * c.cyy) * iA * is;
    f[9] = (vv * c.cxx - T(2) * uv * c.cxy + uu * c.cyy) * iA * is;
    sg0 = M * sr * sr * ir2;
    sc0 = A * sr * ir2;
    sp0 = M * sr * sr * ir;
    cp0 = A * sr * ir;
}

// ---------------------------------------------------------------------------
// Model fragments: the network's weights arranged per lane for the MFMAs
// ---------------------------------------------------------------------------
struct FarFrags {
    bf16x8 w1[2][kWave];            // layer 1, A = W1^T, M-block mb
    bf16x8 w2[2][4][kWave];         // layer 2, A = W2^T, K permuted
    bf16x8 w3[4][kWave];            // layer 3, A = W3^T padded to 32 rows, K permuted
    halfx8 w2h[2][4][kWave];        // inference: layers 2 and 3 in FP16 (as w2, w3)
    halfx8 w3h[4][kWave];
    halfx8 b2h[2][kWave];           // inference: A[m][0], A[m][1] = b2 split into FP16 hi + lo, K 2..15 zero
    bf16x8 w3t[2][kWave];           // backward: A = W3 (hidden x outputs; K = 8h + j <-> output 4h + j)
    bf16x8 w2t[2][4][kWave];        // backward: A = W2 (in x out), K (out) permuted
    float b2[kFarHid];
    float b3[kFarOut];
    float mean[kFarIn], istd[kFarIn], oscale[kFarOut];
};

// Layer-3 outputs live in accumulator r of lane half h as output 4h + r (r < 4).
__host__ __device__ __forceinline__ int out_of(int h, int r) { return 4 * h + r; }

__global__ void k_far_pack(const float* __restrict__ P, const float* __restrict__ stats, FarFrags* __restrict__ fr) {
    const int l = threadIdx.x;          // one wave
    const int h = l >> 5, c = l & 31;
    const float* W1 = P + kFarOffW1;
    const float* W2 = P + kFarOffW2;
    const float* W3 = P + kFarOffW3;
    for (int mb = 0; mb < 2; mb++) {
        float v[8];
        for (int j = 0; j < 8; j++) v[j] = W1[(8 * h + j) * kFarHid + 32 * mb + c];
        fr->w1[mb][l] = pack8(v);
        for (int t = 0; t < 4; t++) {
            for (int j = 0; j < 8; j++) v[j] = W2[kperm(t, h, j) * kFarHid + 32 * mb + c];
            fr->w2[mb][t][l] = pack8(v);
            fr->w2h[mb][t][l] = pack8h(v);
            for (int j = 0; j < 8; j++) v[j] = W2[(32 * mb + c) * kFarHid + kperm(t, h, j)];
            fr->w2t[mb][t][l] = pack8(v);
        }
        for (int j = 0; j < 8; j++)
            v[j] = (j < 4 && out_of(h, j) < kFarOut) ? W3[(32 * mb + c) * kFarOut + out_of(h, j)] : 0.f;
        fr->w3t[mb][l] = pack8(v);
    }
    for (int t = 0; t < 4; t++) {
        float v[8];
        for (int j = 0; j < 8; j++) v[j] = c < kFarOut ? W3[kperm(t, h, j) * kFarOut + c] : 0.f;
        fr->w3[t][l] = pack8(v);
        fr->w3h[t][l] = pack8h(v);
    }
    fr->b2[l] = P[kFarOffB2 + l];
    for (int mb = 0; mb < 2; mb++) {
        float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (h == 0) {
            const float b = P[kFarOffB2 + 32 * mb + c];
            const _Float16 hi = (_Float16)b;
            v[0] = (float)hi;
            v[1] = b - (float)hi;
        }
        fr->b2h[mb][l] = pack8h(v);
    }
    if (l < kFarOut) fr->b3[l] = P[kFarOffB3 + l];
    if (l < kFarIn) {
        fr->mean[l] = stats[l];
        fr->istd[l] = stats[kFarIn + l];
    }
    if (l < kFarOut) fr->oscale[l] = stats[2 * kFarIn + l];
}

// Standardised inputs of one row as the 16 BF16 values of the layer-1 operand.
template <typename FR>
__device__ __forceinline__ void std_inputs(const FR& fr, const float* f, bool valid, float* x) {
#pragma unroll
    for (int k = 0; k < kFarIn; k++) x[k] = valid ? (f[k] - fr.mean[k]) * fr.istd[k] : 0.f;
    x[kFarIn] = valid ? 1.f : 0.f;
#pragma unroll
    for (int k = kFarIn + 1; k < 16; k++) x[k] = 0.f;
}

// Accumulator r of M-block mb holds hidden unit 32mb + 8(r/4) + 4h + r%4.
__device__ __forceinline__ int c_row(int mb, int r, int h) { return 32 * mb + 8 * (r >> 2) + 4 * h + (r & 3); }

// Forward pass of one 64-row tile whose standardised inputs are in LDS
// (xs[row][16] BF16).  Returns the layer-3 tiles: lane (h, c) holds outputs
// 4h + r of row 32 nb + c in acc[r] of o[nb] (r < 4, 4h + r < kFarOut); the
// pre-activations z1, z2 are kept for training.
struct Fwd {
    floatx16 z1[2][2], z2[2][2];
    bf16x8 a1[4][2], a2[4][2];
    floatx16 o[2];
};

__device__ __forceinline__ void mlp_forward(const FarFrags& fr, const __bf16* xs, int l, Fwd& F) {
    const int h = l >> 5, c = l & 31;
    bf16x8 bx[2];
#pragma unroll
    for (int nb = 0; nb < 2; nb++) bx[nb] = *reinterpret_cast<const bf16x8*>(xs + (32 * nb + c) * 16 + 8 * h);
    const floatx16 zero = {};
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++) F.z1[mb][nb] = mfma32(fr.w1[mb][l], bx[nb], zero);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = act(F.z1[t >> 1][nb][8 * (t & 1) + j]);
            F.a1[t][nb] = pack8(v);
        }
#pragma unroll
    for (int mb = 0; mb < 2; mb++) {
        floatx16 binit;
#pragma unroll
        for (int r = 0; r < 16; r++) binit[r] = fr.b2[c_row(mb, r, h)];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            floatx16 acc = binit;
#pragma unroll
            for (int t = 0; t < 4; t++) acc = mfma32(fr.w2[mb][t][l], F.a1[t][nb], acc);
            F.z2[mb][nb] = acc;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = act(F.z2[t >> 1][nb][8 * (t & 1) + j]);
            F.a2[t][nb] = pack8(v);
        }
    floatx16 binit3 = {};
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (out_of(h, r) < kFarOut) binit3[r] = fr.b3[out_of(h, r)];
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
        floatx16 acc = binit3;
#pragma unroll
        for (int t = 0; t < 4; t++) acc = mfma32(fr.w3[t][l], F.a2[t][nb], acc);
        F.o[nb] = acc;
    }
}

// ---------------------------------------------------------------------------
// Inference: add the predicted residual of every listed pair to the forces
// ---------------------------------------------------------------------------
// The inference part of FarFrags, staged once per block in LDS so that no wave
// holds the weights in registers across its tiles.
struct FarInfer {
    bf16x8 w1[2][kWave];
    halfx8 w2[2][4][kWave];             // FP16: the packed-FP16 activations feed them
    halfx8 w3[4][kWave];
    halfx8 b2h[2][kWave];               // layer-2 bias as an MFMA operand (times a column of ones)
};

constexpr int kMlpWaves = 4;                   // waves per block, one 64-particle group each

// The gathered node of one listed pair: {COM x, y, mass, charge}, half width
// and moments.
struct FarNode {
    float4 c4;
    float half;
    NodeMom mo;
};

__device__ __forceinline__ void far_gather(const FarArgs& a, uint32_t node, FarNode& nd) {
    nd.c4 = *reinterpret_cast<const float4*>(&a.nodes[node].comx);
    nd.half = a.nodes[node].half;
    nd.mo = a.node_mom[node];
}

// __builtin_amdgcn_permlane32_swap(x, y) returns {x', y'}: lanes 32-63 of x
// trade places with lanes 0-31 of y (v_permlane32_swap).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Layers 1-3 for NB 32-row column blocks whose layer-1 B operands are bx[nb]
// (lane (h, c) holds inputs 8h..8h+7 of row 32 nb + c), all blocks per weight
// fragment read from LDS: lane (h, c) gets outputs 4h + r of row 32 nb + c,
// without the layer-3 bias, in accumulator r of o[nb] (r < 4, 4h + r <
// kFarOut).
//
// A wave issues independent vector instructions while its own MFMAs execute,
// but an activation waits for the product it reads.  The network is
// therefore laid out as stages (fenced with sched_barrier) in which the
// products of one step issue next to the activations of the step before:
//   S0-S3   layer-1 products (mb, nb) in turn, each with the activation of
//           the previous one;
//   S4, S5  layer-2 products of M-block 0 (bias, K-steps 0-1 need only the
//           layer-1 M-block 0), with the last layer-1 activation;
//   S6, S7  layer-2 products of M-block 1, with the activations of M-block 0;
//   S8, S9  layer-3 products (K-steps 0-1 read M-block 0), with the
//           activations of M-block 1.
// The layer-2 bias enters as the C operand of the chains (one product per
// M-block); the six layer-3 biases are cheaper to add to the outputs.
// Weight fragments are read from LDS one stage before their products.
#define GEON_STAGE() __builtin_amdgcn_sched_barrier(0)
template <int NB, typename Hook>
__device__ __forceinline__ void mlp_tile(const FarInfer& fr, const bf16x8* bx, int l, floatx16* o, Hook&& hook) {
    static_assert(NB % 2 == 0, "the M-block activations are split over two stages");
    const int h = l >> 5;
    const floatx16 zero = {};
    // B operand of the bias product: B[0][n] = B[1][n] = 1, the rest 0
    halfx8 ones = {};
    if (h == 0) { ones[0] = (_Float16)1.f; ones[1] = (_Float16)1.f; }
    auto act16 = [](const floatx16& z, halfx8& lo, halfx8& hi) {
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = z[r];
        lo = act8h(v);
        hi = act8h(v + 8);
    };
    halfx8 a1[4][NB], a2[4][NB];
    floatx16 z[2], acc0[NB], acc1[NB];
    // S0 .. S(2NB-1): layer-1 product k = NB mb + nb, activation of k - 1
    const bf16x8 w1a = fr.w1[0][l], w1b = fr.w1[1][l];
    halfx8 wb0 = fr.b2h[0][l], w0 = fr.w2[0][0][l], w1 = fr.w2[0][1][l];
#pragma unroll
    for (int k = 0; k < 2 * NB; k++) {
        z[k & 1] = mfma32(k < NB ? w1a : w1b, bx[k % NB], zero);
        if (k > 0) {
            const int p = k - 1;
            act16(z[p & 1], a1[2 * (p / NB)][p % NB], a1[2 * (p / NB) + 1][p % NB]);
        }
        hook(k);
        GEON_STAGE();
    }
    // S4: layer-2 M-block 0, bias and K-steps 0-1; the last layer-1 activation
    halfx8 w2 = fr.w2[0][2][l], w3 = fr.w2[0][3][l];
    {
        const floatx16 bias = mfma32(wb0, ones, zero);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc0[nb] = mfma32(w0, a1[0][nb], bias);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc0[nb] = mfma32(w1, a1[1][nb], acc0[nb]);
        const int p = 2 * NB - 1;
        act16(z[p & 1], a1[2][NB - 1], a1[3][NB - 1]);
    }
    hook(4);
    GEON_STAGE();
    // S5: layer-2 M-block 0, K-steps 2-3
    halfx8 wb1 = fr.b2h[1][l];
    w0 = fr.w2[1][0][l];
    w1 = fr.w2[1][1][l];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc0[nb] = mfma32(w2, a1[2][nb], acc0[nb]);
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc0[nb] = mfma32(w3, a1[3][nb], acc0[nb]);
    hook(5);
    GEON_STAGE();
    // S6: layer-2 M-block 1, bias and K-steps 0-1; half of M-block 0's activations
    w2 = fr.w2[1][2][l];
    w3 = fr.w2[1][3][l];
    {
        const floatx16 bias = mfma32(wb1, ones, zero);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc1[nb] = mfma32(w0, a1[0][nb], bias);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc1[nb] = mfma32(w1, a1[1][nb], acc1[nb]);
#pragma unroll
        for (int nb = 0; nb < NB / 2; nb++) act16(acc0[nb], a2[0][nb], a2[1][nb]);
    }
    hook(6);
    GEON_STAGE();
    // S7: layer-2 M-block 1, K-steps 2-3; the rest of M-block 0's activations
    w0 = fr.w3[0][l];
    w1 = fr.w3[1][l];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc1[nb] = mfma32(w2, a1[2][nb], acc1[nb]);
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc1[nb] = mfma32(w3, a1[3][nb], acc1[nb]);
#pragma unroll
    for (int nb = NB / 2; nb < NB; nb++) act16(acc0[nb], a2[0][nb], a2[1][nb]);
    hook(7);
    GEON_STAGE();
    // S8: layer 3, K-steps 0-1 (M-block 0); M-block 1's activations.  The
    //     caller adds b3 once the outputs sit in their rows' lanes.
    w2 = fr.w3[2][l];
    w3 = fr.w3[3][l];
    {
#pragma unroll
        for (int nb = 0; nb < NB; nb++) o[nb] = mfma32(w0, a2[0][nb], zero);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) o[nb] = mfma32(w1, a2[1][nb], o[nb]);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) act16(acc1[nb], a2[2][nb], a2[3][nb]);
    }
    hook(8);
    GEON_STAGE();
    // S9: layer 3, K-steps 2-3
#pragma unroll
    for (int nb = 0; nb < NB; nb++) o[nb] = mfma32(w2, a2[2][nb], o[nb]);
#pragma unroll
    for (int nb = 0; nb < NB; nb++) o[nb] = mfma32(w3, a2[3][nb], o[nb]);
    hook(9);
}
#undef GEON_STAGE

// Waves per SIMD of k_far_mlp: the staged network fits 170 registers.
constexpr int kFarMlpOccupancy = 3;

// The features of one listed pair, formed in parts so that the next tile's
// pair can be worked on inside this tile's network stages (pair_features in
// FP32, with the node's scales 1/(M s^2), 1/(A s), 1/(A s^2) precomputed at
// the tree build).  A row past the lane's list
// takes the offset (1, 0) and whatever node record the registers hold, all
// finite, and zero scales: its outputs add nothing.
struct PairFeat {
    float f[kFarIn];
    float ux, uy, uu, vv, uv;
    float sg, sc, spg, spc;              // force and potential scales, times m_i and q_i (0 past the list)
};

__device__ __forceinline__ void feat_geometry(const FarArgs& a, const FarNode& nd, const float4& s4, bool valid,
                                              PairFeat& P) {
    float dx = valid ? nd.c4.x - s4.x : 1.f, dy = valid ? nd.c4.y - s4.y : 0.f;
    if (a.periodic) {
        dx = torus_wrap(dx, a.dom_w, 0.5f * a.dom_w);
        dy = torus_wrap(dy, a.dom_h, 0.5f * a.dom_h);
    }
    const float rho2 = dx * dx + dy * dy;
    const float r2 = rho2 + a.soft2;
    const float ir = rsq_t(r2);
    const float ir2 = ir * ir;
    const float irho = rsq_t(rho2);
    P.ux = dx * irho; P.uy = dy * irho;
    const float sr = 2.f * nd.half * ir;
    P.uu = P.ux * P.ux; P.vv = P.uy * P.uy; P.uv = P.ux * P.uy;
    P.f[0] = sr;
    P.f[1] = a.soft2 * ir2;
    const float M = nd.c4.z, A = nd.mo.aq;
    const float mi = valid && a.grav ? s4.z : 0.f, qi = valid && a.coul ? s4.w : 0.f;
    P.sg = mi * (M * sr * sr * ir2);
    P.sc = qi * (A * sr * ir2);
    P.spg = mi * (M * sr * sr * ir);
    P.spc = qi * (A * sr * ir);
    if (!valid) P.ux = P.uy = 0.f;       // after the products above: the moments below stay finite
}

__device__ __forceinline__ void feat_moments(const NodeMom& c, PairFeat& P) {
    const float uu = P.uu, vv = P.vv, uv = P.uv, ux = P.ux, uy = P.uy;
    P.f[2] = (uu * c.qxx + 2.f * uv * c.qxy + vv * c.qyy) * c.ims2;
    P.f[3] = (-uv * c.qxx + (uu - vv) * c.qxy + uv * c.qyy) * c.ims2;
    P.f[4] = (vv * c.qxx - 2.f * uv * c.qxy + uu * c.qyy) * c.ims2;
    P.f[5] = (ux * c.dx + uy * c.dy) * c.ias;
    P.f[6] = (-uy * c.dx + ux * c.dy) * c.ias;
    P.f[7] = (uu * c.cxx + 2.f * uv * c.cxy + vv * c.cyy) * c.ias2;
    P.f[8] = (-uv * c.cxx + (uu - vv) * c.cxy + uv * c.cyy) * c.ias2;
    P.f[9] = (vv * c.cxx - 2.f * uv * c.cxy + uu * c.cyy) * c.ias2;
}

// Standardised inputs (f - mean) istd, the bias input, then the layer-1 B
// operands: lane (h, c) needs inputs 8h..8h+7 of rows c and 32 + c, i.e. its
// own half h and the half h of its partner lane l ^ 32.
__device__ __forceinline__ void feat_operands(const PairFeat& P, bool valid, const float* mean, const float* istd,
                                              bf16x8* bx) {
    float x[16];
#pragma unroll
    for (int k = 0; k < kFarIn; k++) x[k] = (P.f[k] - mean[k]) * istd[k];
    x[kFarIn] = valid ? 1.f : 0.f;
#pragma unroll
    for (int k = kFarIn + 1; k < 16; k++) x[k] = 0.f;
    u32x4 u0 = __builtin_bit_cast(u32x4, pack8(x)), u1 = __builtin_bit_cast(u32x4, pack8(x + 8));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const auto r = __builtin_amdgcn_permlane32_swap(u0[k], u1[k], false, false);
        u0[k] = r[0];
        u1[k] = r[1];
    }
    bx[0] = __builtin_bit_cast(bf16x8, u0);
    bx[1] = __builtin_bit_cast(bf16x8, u1);
}

// One wave per 64-particle group, lane l for particle 64 g + l: tile t holds
// every lane's t-th listed node (far_slot), row l of the tile is lane l's
// pair, so the predicted residual of a row lands in the lane that owns its
// observer and is summed there, in list order, without a cross-lane scatter.
// The outputs change lane halves with v_permlane32_swap: row l's force
// outputs (0..3) are in lane (0, l % 32), its potential outputs (4, 5) in
// lane (1, l % 32), both of tile half l / 32.
//
// Software pipeline over the tiles: tile t + 1's features are formed inside
// tile t's network stages that have room for vector work (geometry in S4,
// moments and operands in S5), tile t + 2's node is gathered in S5 once
// tile t + 1's features no longer need the registers, and its list entry is
// read a tile earlier still.
__global__ __launch_bounds__(kMlpWaves * kWave) __attribute__((amdgpu_waves_per_eu(kFarMlpOccupancy, kFarMlpOccupancy)))
void k_far_mlp(FarArgs a) {
    __shared__ FarInfer wfr;
    {
        // stage the weights: every field is a whole number of dwords
        const FarFrags& g = *a.frags;
        auto copy = [&](uint32_t* dst, const uint32_t* src, int ndw) {
            for (int k = threadIdx.x; k < ndw; k += kMlpWaves * kWave) dst[k] = src[k];
        };
        copy(reinterpret_cast<uint32_t*>(wfr.w1), reinterpret_cast<const uint32_t*>(g.w1), sizeof(g.w1) / 4);
        copy(reinterpret_cast<uint32_t*>(wfr.w2), reinterpret_cast<const uint32_t*>(g.w2h), sizeof(g.w2h) / 4);
        copy(reinterpret_cast<uint32_t*>(wfr.w3), reinterpret_cast<const uint32_t*>(g.w3h), sizeof(g.w3h) / 4);
        copy(reinterpret_cast<uint32_t*>(wfr.b2h), reinterpret_cast<const uint32_t*>(g.b2h), sizeof(g.b2h) / 4);
    }
    __syncthreads();
    // wave-uniform constants, read once into scalar registers (the inputs are
    // standardised as (f - mean) istd, one scalar operand per instruction)
    float mean[kFarIn], istd[kFarIn], osc[kFarOut], b3[kFarOut];
#pragma unroll
    for (int k = 0; k < kFarIn; k++) {
        mean[k] = a.frags->mean[k];
        istd[k] = a.frags->istd[k];
    }
#pragma unroll
    for (int k = 0; k < kFarOut; k++) {
        osc[k] = a.frags->oscale[k];
        b3[k] = a.frags->b3[k];
    }
    const int l = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int64_t g = (int64_t)blockIdx.x * kMlpWaves + wv;
    const int n = a.ctrl->n_alive;
    if (g * kWave >= n) return;                  // no barrier below: a whole wave leaves
    const FarInfer& fr = wfr;
    const int i = (int)(g * kWave) + l;
    const int cnt = i < n ? a.far_count[g * kWave + l] : 0;
    int tiles = cnt;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tiles = max(tiles, __shfl_xor(tiles, off, kWave));
    const float4 s4 = i < n ? a.src4[i] : make_float4(0.f, 0.f, 0.f, 0.f);   // observer x, y, m, q
    auto entry = [&](int t) { return t < cnt ? a.far_list[far_slot(g, a.far_cap, l, t)] : 0u; };
    float sx = 0.f, sy = 0.f, sp = 0.f;
    // prologue: tile 0's operands, tile 1's node, tile 2's list entry
    FarNode cur = {};
    PairFeat P;
    bf16x8 bx[2];
    if (0 < cnt) far_gather(a, entry(0), cur);
    uint32_t e_next = entry(1);
    feat_geometry(a, cur, s4, 0 < cnt, P);
    feat_moments(cur.mo, P);
    feat_operands(P, 0 < cnt, mean, istd, bx);
    if (1 < cnt) far_gather(a, e_next, cur);
    e_next = entry(2);
    for (int t = 0; t < tiles; t++) {
        // The loop stores nothing to LDS: without this compiler barrier the
        // weight fragments (about 90 registers) would be hoisted out of it.
        asm volatile("" ::: "memory");
        const float ux = P.ux, uy = P.uy, sg = P.sg, sc = P.sc, spg = P.spg, spc = P.spc;
        const bool vn = t + 1 < cnt;
        bf16x8 bn[2];
        floatx16 o[2];
        mlp_tile<2>(fr, bx, l, o, [&](int stage) {
            if (stage == 4) feat_geometry(a, cur, s4, vn, P);
            if (stage == 5) {
                feat_moments(cur.mo, P);
                feat_operands(P, vn, mean, istd, bn);
                if (t + 2 < cnt) far_gather(a, e_next, cur);
                e_next = entry(t + 3);
            }
        });
        // this lane's row: force outputs from half 0, potential outputs from
        // half 1 of its own tile half, then rotate, scale and add
        float q[4], p[2];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const auto w = __builtin_amdgcn_permlane32_swap(__float_as_uint(o[0][r]), __float_as_uint(o[1][r]), false,
                                                            false);
            q[r] = __uint_as_float(w[0]) + b3[r];
            if (r < 2) p[r] = __uint_as_float(w[1]) + b3[4 + r];
        }
        const float gu = q[0] * osc[0], gv = q[1] * osc[1];
        const float cu = q[2] * osc[2], cv = q[3] * osc[3];
        sx += sg * (ux * gu - uy * gv) + sc * (ux * cu - uy * cv);
        sy += sg * (uy * gu + ux * gv) + sc * (uy * cu + ux * cv);
        sp += spg * p[0] * osc[4] + spc * p[1] * osc[5];
        bx[0] = bn[0];
        bx[1] = bn[1];
    }
    if (i < n) {
        a.fx[i] += sx;
        a.fy[i] += sy;
        a.pe[i] += sp;
    }
}

// ---------------------------------------------------------------------------
// Training samples: features, exact residual targets and loss weights
// ---------------------------------------------------------------------------
// One wave per listed pair of the selected groups.  The residual is the FP64
// sum over the node's points (its Morton range) of the softened gravity and
// Coulomb kernels minus the node's monopole, each point at offset
// d + (x_j - COM) from the observer.
__global__ __launch_bounds__(256) void k_far_samples(FarSampleArgs a) {
    const int l = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
    if (row >= a.n_rows) return;
    // selected group owning this row
    int lo = 0, hi = a.n_groups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.row_off[mid] <= row) lo = mid; else hi = mid - 1;
    }
    const int64_t g = a.groups[lo];
    // the group's rows run through its lane queues in lane order
    const int r = (int)(row - a.row_off[lo]);
    const int c = a.far_count[g * kWave + l];
    int incl = c;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int v = __shfl_up(incl, off, kWave);
        if (l >= off) incl += v;
    }
    const int lane = (int)__builtin_ctzll(__ballot(r < incl));
    const int k = r - __shfl(incl - c, lane, kWave);
    const int node = (int)a.far_list[far_slot(g, a.far_cap, lane, k)];
    const int i = (int)(g * kWave) + lane;
    const double xi = a.x[i], yi = a.y[i];
    const double cx = a.comx64[node], cy = a.comy64[node];
    const double M = a.mass64[node], Q = a.qq64[node] * kChargeQuantum;
    double dx = cx - xi, dy = cy - yi;
    if (a.periodic) {
        const double W = a.dom_w, H = a.dom_h;
        dx = dx > 0.5 * W ? dx - W : (dx < -0.5 * W ? dx + W : dx);
        dy = dy > 0.5 * H ? dy - H : (dy < -0.5 * H ? dy + H : dy);
    }
    const int first = a.first[node], np = a.npts[node];
    double gx = 0, gy = 0, qx = 0, qy = 0, pg = 0, pq = 0;
    for (int j = first + l; j < first + np; j += kWave) {
        const double ex = dx + ((double)a.x[j] - cx), ey = dy + ((double)a.y[j] - cy);
        const double r2 = ex * ex + ey * ey + a.soft2;
        const double ir = 1.0 / sqrt(r2);
        const double hh = ir / r2;
        const double mj = a.m[j], qj = a.qq[j] * kChargeQuantum;
        gx += mj * ex * hh; gy += mj * ey * hh;
        qx += qj * ex * hh; qy += qj * ey * hh;
        pg -= mj * ir;                               // pair potentials (forces.js:199, 216)
        pq += qj * ir;
    }
    for (int off = 32; off > 0; off >>= 1) {
        gx += __shfl_xor(gx, off, 64); gy += __shfl_xor(gy, off, 64);
        qx += __shfl_xor(qx, off, 64); qy += __shfl_xor(qy, off, 64);
        pg += __shfl_xor(pg, off, 64); pq += __shfl_xor(pq, off, 64);
    }
    if (l != 0) return;
    const double r2 = dx * dx + dy * dy + a.soft2;
    const double ir = 1.0 / sqrt(r2);
    const double hh = ir / r2;
    // residuals per unit observer mass (gravity) and charge (Coulomb)
    const double rgx = gx - M * dx * hh, rgy = gy - M * dy * hh;
    const double rcx = -(qx - Q * dx * hh), rcy = -(qy - Q * dy * hh);
    const double rpg = pg + M * ir, rpq = pq - Q * ir;
    const NodeMom mo = a.node_mom[node];
    double f[kFarIn], ux, uy, sg0, sc0, sp0, cp0;
    pair_features<double>(dx, dy, (double)a.soft2, 2.0 * (double)a.nodes[node].half, M, mo, f, ux, uy, sg0, sc0,
                          sp0, cp0);
    float* out = a.out + (size_t)row * kFarSample;
    for (int k = 0; k < kFarIn; k++) out[k] = (float)f[k];
    const double ig = sg0 > 0 ? 1.0 / sg0 : 0.0, ic = sc0 > 0 ? 1.0 / sc0 : 0.0;
    const double igp = sp0 > 0 ? 1.0 / sp0 : 0.0, icp = cp0 > 0 ? 1.0 / cp0 : 0.0;
    float* t = out + kFarIn;
    t[0] = (float)((ux * rgx + uy * rgy) * ig);
    t[1] = (float)((-uy * rgx + ux * rgy) * ig);
    t[2] = (float)((ux * rcx + uy * rcy) * ic);
    t[3] = (float)((-uy * rcx + ux * rcy) * ic);
    t[4] = (float)(rpg * igp);
    t[5] = (float)(rpq * icp);
    if (a.obs) a.obs[row] = a.slot[i];
    const double mi = a.m[i], qi = a.qq[i] * kChargeQuantum;
    const double iF2 = 1.0 / (a.f2ref[0] > 0 ? a.f2ref[0] : 1.0);
    const double iP2 = 1.0 / (a.f2ref[1] > 0 ? a.f2ref[1] : 1.0);
    float* w = out + kFarIn + kFarOut;
    w[0] = a.grav && sg0 > 0 ? (float)((mi * sg0) * (mi * sg0) * iF2) : 0.f;
    w[1] = a.coul && sc0 > 0 ? (float)((qi * sc0) * (qi * sc0) * iF2) : 0.f;
    w[2] = a.grav && sp0 > 0 ? (float)((mi * sp0) * (mi * sp0) * iP2) : 0.f;
    w[3] = a.coul && cp0 > 0 ? (float)((qi * cp0) * (qi * cp0) * iP2) : 0.f;
}

// ---------------------------------------------------------------------------
// Training: forward + backward of the weighted squared error on a batch
// ---------------------------------------------------------------------------
// Loss L = wscale / B * sum_rows sum_j w_j (oscale_j o_j - t_j)^2 with the
// sample's weight of output j: gravity force (j < 2), Coulomb force (j < 4),
// gravity PE (j = 4), Coulomb PE (j = 5).  Each block (one wave) takes one
// tile of 64 rows and writes its partial gradient (kFarParams floats) and loss.
__host__ __device__ __forceinline__ int wsel(int j) { return j < 4 ? (j >> 1) : j - 2; }
constexpr int kHS = 72;      // LDS row stride (BF16) of the hidden-major tiles

__device__ __forceinline__ void store_hmajor(__bf16* dst, const floatx16& v, int mb, int nb, int h, int c) {
#pragma unroll
    for (int r = 0; r < 16; r++) dst[c_row(mb, r, h) * kHS + 32 * nb + c] = (__bf16)v[r];
}

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = p[j];
    return r;
}

__global__ __launch_bounds__(kWave) void k_far_train(FarTrainArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 xs[kWave * 16];     // [row][16]
    __shared__ __bf16 xf[16 * kHS];                                     // [input][row]
    __shared__ __bf16 a1f[kFarHid * kHS], a2f[kFarHid * kHS];           // [hidden][row]
    __shared__ __bf16 dz2f[kFarHid * kHS], dz1f[kFarHid * kHS];
    __shared__ __bf16 dof[kFarOut * kHS];                               // [output][row]
    __shared__ float tg[kWave][kFarOut], wt[kWave][kFarW];
    __shared__ float g1s[4 * 16 * kWave], g2s[4 * 16 * kWave];         // act'(z) in accumulator order
    const int l = threadIdx.x;
    const int h = l >> 5, c = l & 31;
    const FarFrags& fr = *a.frags;
    const floatx16 zero = {};
    floatx16 gw1[2], gw2[2][2], gw3[2];
    float gb2[2][16], gb3[4] = {0.f, 0.f, 0.f, 0.f};  // gb3[r]: output 4h + r
    gw1[0] = gw1[1] = zero;
    gw2[0][0] = gw2[0][1] = gw2[1][0] = gw2[1][1] = zero;
    gw3[0] = gw3[1] = zero;
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int r = 0; r < 16; r++) gb2[mb][r] = 0.f;
    float loss = 0.f;
    const float inv_b = 1.f / (float)a.batch;
    {
        const int64_t brow = (int64_t)blockIdx.x * kWave + l;
        bool valid = brow < a.batch;
        int64_t s = 0;
        if (valid) {
            if (a.rows) s = a.rows[brow];
            else {
                // counter-based pick of a sample (splitmix64 of seed, step, row)
                uint64_t z = a.seed + 0x9E3779B97F4A7C15ull * (uint64_t)(brow + 1) + 0xD1B54A32D192ED03ull * a.step;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                s = (int64_t)(z % (uint64_t)a.n_samples);
            }
        }
        float f[kFarIn];
        const float* smp = a.samples + (size_t)s * kFarSample;
#pragma unroll
        for (int k = 0; k < kFarIn; k++) f[k] = valid ? smp[k] : 0.f;
        float x[16];
        std_inputs(fr, f, valid, x);
        *reinterpret_cast<bf16x8*>(xs + l * 16) = pack8(x);
        *reinterpret_cast<bf16x8*>(xs + l * 16 + 8) = pack8(x + 8);
#pragma unroll
        for (int k = 0; k < 16; k++) xf[k * kHS + l] = (__bf16)x[k];
#pragma unroll
        for (int j = 0; j < kFarOut; j++) tg[l][j] = valid ? smp[kFarIn + j] : 0.f;
#pragma unroll
        for (int j = 0; j < kFarW; j++) wt[l][j] = valid ? smp[kFarIn + kFarOut + j] * a.wscale : 0.f;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        Fwd F;
        mlp_forward(fr, xs, l, F);
        // activations, hidden-major, for the weight gradients
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                floatx16 v1, v2;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float z1 = F.z1[mb][nb][r], z2 = F.z2[mb][nb][r];
                    v1[r] = act(z1); v2[r] = act(z2);
                    g1s[((mb * 2 + nb) * 16 + r) * kWave + l] = act_grad(z1);
                    g2s[((mb * 2 + nb) * 16 + r) * kWave + l] = act_grad(z2);
                }
                store_hmajor(a1f, v1, mb, nb, h, c);
                store_hmajor(a2f, v2, mb, nb, h, c);
            }
        // output error: lane (h, c) holds outputs 4h + r of row 32 nb + c; as the
        // B operand of dA2 its K index 8h + r carries output 4h + r
        bf16x8 bdo[2];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            const int row = 32 * nb + c;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = out_of(h, r);
                if (j < kFarOut) {
                    const float os = fr.oscale[j];
                    const float err = os * F.o[nb][r] - tg[row][j];
                    const float w = wt[row][wsel(j)];
                    loss += w * err * err * inv_b;
                    v[r] = 2.f * w * err * os * inv_b;
                    gb3[r] += v[r];
                    dof[j * kHS + row] = (__bf16)v[r];
                }
            }
            bdo[nb] = pack8(v);
        }
        // dA2 = W3 dO, dZ2 = dA2 * act'(z2)
        floatx16 dz2[2][2];
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                floatx16 d = mfma32(fr.w3t[mb][l], bdo[nb], zero);
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    d[r] *= g2s[((mb * 2 + nb) * 16 + r) * kWave + l];
                    gb2[mb][r] += d[r];
                }
                dz2[mb][nb] = d;
                store_hmajor(dz2f, d, mb, nb, h, c);
            }
        // dA1 = W2 dZ2 (K over layer-2 units, permuted), dZ1 = dA1 * act'(z1)
        bf16x8 bdz2[4][2];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                float v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) v[j] = dz2[t >> 1][nb][8 * (t & 1) + j];
                bdz2[t][nb] = pack8(v);
            }
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                floatx16 d = zero;
#pragma unroll
                for (int t = 0; t < 4; t++) d = mfma32(fr.w2t[mb][t][l], bdz2[t][nb], d);
#pragma unroll
                for (int r = 0; r < 16; r++) d[r] *= g1s[((mb * 2 + nb) * 16 + r) * kWave + l];
                store_hmajor(dz1f, d, mb, nb, h, c);
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // weight gradients: sums over the tile's rows (K = rows)
#pragma unroll 1
        for (int ks = 0; ks < 4; ks++) {
            const int k0 = 16 * ks + 8 * h;
            bf16x8 ax = c < 16 ? ld8(xf + c * kHS + k0) : bf16x8{};
            bf16x8 a1[2], a2[2], bz1[2], bz2[2];
#pragma unroll
            for (int mb = 0; mb < 2; mb++) {
                a1[mb] = ld8(a1f + (32 * mb + c) * kHS + k0);
                a2[mb] = ld8(a2f + (32 * mb + c) * kHS + k0);
                bz1[mb] = ld8(dz1f + (32 * mb + c) * kHS + k0);
                bz2[mb] = ld8(dz2f + (32 * mb + c) * kHS + k0);
            }
            bf16x8 bo = c < kFarOut ? ld8(dof + c * kHS + k0) : bf16x8{};
#pragma unroll
            for (int nb = 0; nb < 2; nb++) gw1[nb] = mfma32(ax, bz1[nb], gw1[nb]);
#pragma unroll
            for (int ma = 0; ma < 2; ma++) {
#pragma unroll
                for (int nb = 0; nb < 2; nb++) gw2[ma][nb] = mfma32(a1[ma], bz2[nb], gw2[ma][nb]);
                gw3[ma] = mfma32(a2[ma], bo, gw3[ma]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // bias sums over the 32 row lanes of each half
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            for (int off = 16; off > 0; off >>= 1) gb2[mb][r] += __shfl_xor(gb2[mb][r], off, 64);
#pragma unroll
    for (int r = 0; r < 4; r++)
        for (int off = 16; off > 0; off >>= 1) gb3[r] += __shfl_xor(gb3[r], off, 64);
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    float* G = a.grad_part + (size_t)blockIdx.x * kFarParams;
    // W1[k][m]: gw1[nb] has M = input k (rows c_row), N = hidden 32 nb + c
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int k = c_row(0, r, h);
            if (k < 16) G[kFarOffW1 + k * kFarHid + 32 * nb + c] = gw1[nb][r];
        }
#pragma unroll
    for (int ma = 0; ma < 2; ma++) {
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
#pragma unroll
            for (int r = 0; r < 16; r++) G[kFarOffW2 + c_row(ma, r, h) * kFarHid + 32 * nb + c] = gw2[ma][nb][r];
#pragma unroll
        for (int r = 0; r < 16; r++)
            if (c < kFarOut) G[kFarOffW3 + c_row(ma, r, h) * kFarOut + c] = gw3[ma][r];
    }
    if (c == 0) {
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int r = 0; r < 16; r++) G[kFarOffB2 + c_row(mb, r, h)] = gb2[mb][r];
    }
    if (c == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (out_of(h, r) < kFarOut) G[kFarOffB3 + out_of(h, r)] = gb3[r];
    }
    if (l == 0) a.loss_part[blockIdx.x] = loss;
}

// Sum of the block partials (fixed order: four interleaved partial sums per
// parameter, combined in LDS), then optionally one Adam step with bias
// correction on the FP32 master parameters.  Block: 64 parameters x 4 waves.
__global__ __launch_bounds__(256) void k_far_adam(FarTrainArgs a, int nblocks, int apply, float lr, int64_t t) {
    __shared__ float part[4][kWave];
    const int pl = threadIdx.x & (kWave - 1), q = threadIdx.x / kWave;
    const int p = blockIdx.x * kWave + pl;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float L = 0.f;
        for (int b = 0; b < nblocks; b++) L += a.loss_part[b];
        if (a.loss_out) a.loss_out[0] = L;
    }
    float g = 0.f;
    if (p < kFarParams)
        for (int b = q; b < nblocks; b += 4) g += a.grad_part[(size_t)b * kFarParams + p];
    part[q][pl] = g;
    __syncthreads();
    if (q != 0 || p >= kFarParams) return;
    g = (part[0][pl] + part[1][pl]) + (part[2][pl] + part[3][pl]);
    a.grad[p] = g;
    if (!apply) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * a.adam_m[p] + (1.f - b1) * g;
    const float v = b2 * a.adam_v[p] + (1.f - b2) * g * g;
    a.adam_m[p] = m;
    a.adam_v[p] = v;
    const float mh = m / (1.f - powf(b1, (float)t)), vh = v / (1.f - powf(b2, (float)t));
    a.params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Mean |F|^2 and mean PE^2 over the live particles (FP64), the loss-weight normalisers.
__global__ __launch_bounds__(1024) void k_far_f2(const float* __restrict__ fx, const float* __restrict__ fy,
                                                 const float* __restrict__ pe, const Ctrl* __restrict__ ctrl,
                                                 double* __restrict__ out) {
    __shared__ double sh[2][1024];
    const int n = ctrl->n_alive;
    double acc = 0.0, accp = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        acc += (double)fx[i] * fx[i] + (double)fy[i] * fy[i];
        accp += (double)pe[i] * pe[i];
    }
    sh[0][threadIdx.x] = acc;
    sh[1][threadIdx.x] = accp;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) { sh[0][threadIdx.x] += sh[0][threadIdx.x + off]; sh[1][threadIdx.x] += sh[1][threadIdx.x + off]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = n > 0 ? sh[0][0] / n : 0.0; out[1] = n > 0 ? sh[1][0] / n : 0.0; }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_far_f2(hipStream_t s, const float* fx, const float* fy, const float* pe, const Ctrl* ctrl, double* out) {
    hipLaunchKernelGGL(k_far_f2, dim3(1), dim3(1024), 0, s, fx, fy, pe, ctrl, out);
}

size_t far_frags_bytes() { return sizeof(FarFrags); }

void launch_far_pack(hipStream_t s, const float* params, const float* stats, void* frags) {
    hipLaunchKernelGGL(k_far_pack, dim3(1), dim3(kWave), 0, s, params, stats, (FarFrags*)frags);
}

void launch_far_mlp(hipStream_t s, const FarArgs& a, int n_upper) {
    const int groups = (n_upper + kWave - 1) / kWave;
    const int blocks = (groups + kMlpWaves - 1) / kMlpWaves;
    if (groups > 0) hipLaunchKernelGGL(k_far_mlp, dim3(blocks), dim3(kMlpWaves * kWave), 0, s, a);
}

void launch_far_samples(hipStream_t s, const FarSampleArgs& a) {
    if (a.n_rows <= 0) return;
    const int64_t blocks = (a.n_rows + 3) / 4;
    hipLaunchKernelGGL(k_far_samples, dim3((unsigned)blocks), dim3(256), 0, s, a);
}

void launch_far_train(hipStream_t s, const FarTrainArgs& a, int nblocks, int apply, float lr, int64_t t) {
    hipLaunchKernelGGL(k_far_train, dim3(nblocks), dim3(kWave), 0, s, a);
    hipLaunchKernelGGL(k_far_adam, dim3((kFarParams + kWave - 1) / kWave), dim3(256), 0, s, a, nblocks, apply, lr, t);
}

}  // namespace geon
