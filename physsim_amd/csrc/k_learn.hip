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

// The same features for the inference pass, in FP32 with the node's scales
// 1/(M s^2), 1/(A s), 1/(A s^2) precomputed at the tree build.
__device__ __forceinline__ void pair_features_fast(float dx, float dy, float eps2, float s, float M,
                                                   const NodeMom& c, float* f, float& ux, float& uy, float& sg0,
                                                   float& sc0, float& sp0, float& cp0) {
    const float rho2 = dx * dx + dy * dy;
    const float r2 = rho2 + eps2;
    const float ir = rsq_t(r2);
    const float ir2 = ir * ir;
    const float irho = rsq_t(rho2);
    ux = dx * irho; uy = dy * irho;
    const float sr = s * ir;
    const float uu = ux * ux, vv = uy * uy, uv = ux * uy;
    f[0] = sr;
    f[1] = eps2 * ir2;
    f[2] = (uu * c.qxx + 2.f * uv * c.qxy + vv * c.qyy) * c.ims2;
    f[3] = (-uv * c.qxx + (uu - vv) * c.qxy + uv * c.qyy) * c.ims2;
    f[4] = (vv * c.qxx - 2.f * uv * c.qxy + uu * c.qyy) * c.ims2;
    f[5] = (ux * c.dx + uy * c.dy) * c.ias;
    f[6] = (-uy * c.dx + ux * c.dy) * c.ias;
    f[7] = (uu * c.cxx + 2.f * uv * c.cxy + vv * c.cyy) * c.ias2;
    f[8] = (-uv * c.cxx + (uu - vv) * c.cxy + uv * c.cyy) * c.ias2;
    f[9] = (vv * c.cxx - 2.f * uv * c.cxy + uu * c.cyy) * c.ias2;
    const float A = c.aq;
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
    halfx8 b3h[kWave];              // inference: the same for b3 (rows m < kFarOut)
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
    {
        float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (h == 0 && c < kFarOut) {
            const float b = P[kFarOffB3 + c];
            const _Float16 hi = (_Float16)b;
            v[0] = (float)hi;
            v[1] = b - (float)hi;
        }
        fr->b3h[l] = pack8h(v);
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
struct FarRow {
    float ux, uy, sg, sc, sp, cp;
};

// The inference part of FarFrags, staged once per block in LDS so that no wave
// holds the weights in registers across its tiles.
struct FarInfer {
    bf16x8 w1[2][kWave];
    halfx8 w2[2][4][kWave];             // FP16: the packed-FP16 activations feed them
    halfx8 w3[4][kWave];
    halfx8 b2h[2][kWave];               // layer-2 bias as an MFMA operand (times a column of ones)
    halfx8 b3h[kWave];                  // layer-3 bias, the same way
    float istd[kFarIn], nmi[kFarIn], oscale[8];   // x = f istd + nmi, nmi = -mean istd
};

constexpr int kMlpWaves = 4;                   // waves per block, one 64-particle group each

// The gathered inputs of one listed pair: the observer's {x, y, m, q}, the
// node's {COM x, y, mass, charge} and half width, its moments.
struct FarIn {
    float4 s4, c4;
    float half;
    NodeMom mo;
    int lane;
};

__device__ __forceinline__ void far_gather(const FarArgs& a, int64_t g, uint32_t e, FarIn& in) {
    const int node = (int)(e & kFarNodeMask);
    in.lane = (int)(e >> kFarLaneShift);
    in.s4 = a.src4[(int)(g * kWave) + in.lane];
    in.c4 = *reinterpret_cast<const float4*>(&a.nodes[node].comx);
    in.half = a.nodes[node].half;
    in.mo = a.node_mom[node];
}

// Layers 1-3 for the 64 rows of the tile in xs, both 32-row halves per weight
// fragment read from LDS (the LDS is busy most of this kernel's time):
// lane (h, c) gets outputs 4h + r of row 32 nb + c in accumulator r of o[nb]
// (r < 4, 4h + r < kFarOut).
__device__ __forceinline__ void mlp_tile(const FarInfer& fr, const __bf16* xs, int l, floatx16* o) {
    const int h = l >> 5, c = l & 31;
    bf16x8 bx[2];
#pragma unroll
    for (int nb = 0; nb < 2; nb++) bx[nb] = *reinterpret_cast<const bf16x8*>(xs + (32 * nb + c) * 16 + 8 * h);
    const floatx16 zero = {};
    // B operand of the bias product: B[0][n] = B[1][n] = 1, the rest 0
    halfx8 ones = {};
    if (h == 0) { ones[0] = (_Float16)1.f; ones[1] = (_Float16)1.f; }
    halfx8 a1[4][2], a2[4][2];
#pragma unroll
    for (int mb = 0; mb < 2; mb++) {
        const bf16x8 w = fr.w1[mb][l];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            const floatx16 z1 = mfma32(w, bx[nb], zero);
            float v[16];
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = z1[r];
            a1[2 * mb][nb] = act8h(v);
            a1[2 * mb + 1][nb] = act8h(v + 8);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 2; mb++) {
        // b2 in every column (hi + lo in FP32) as the C operand of both chains
        const floatx16 bias = mfma32(fr.b2h[mb][l], ones, zero);
        floatx16 acc[2];
        {
            const halfx8 w = fr.w2[mb][0][l];
#pragma unroll
            for (int nb = 0; nb < 2; nb++) acc[nb] = mfma32(w, a1[0][nb], bias);
        }
#pragma unroll
        for (int t = 1; t < 4; t++) {
            const halfx8 w = fr.w2[mb][t][l];
#pragma unroll
            for (int nb = 0; nb < 2; nb++) acc[nb] = mfma32(w, a1[t][nb], acc[nb]);
        }
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            float v[16];
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = acc[nb][r];
            a2[2 * mb][nb] = act8h(v);
            a2[2 * mb + 1][nb] = act8h(v + 8);
        }
    }
    // b3 in every column: accumulator r < 4 of lane half h holds output 4h + r
    const floatx16 b3 = mfma32(fr.b3h[l], ones, zero);
    {
        const halfx8 w = fr.w3[0][l];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) o[nb] = mfma32(w, a2[0][nb], b3);
    }
#pragma unroll
    for (int t = 1; t < 4; t++) {
        const halfx8 w = fr.w3[t][l];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) o[nb] = mfma32(w, a2[t][nb], o[nb]);
    }
}

__global__ __launch_bounds__(kMlpWaves * kWave) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_far_mlp(FarArgs a) {
    __shared__ FarInfer wfr;
    __shared__ __attribute__((aligned(16))) __bf16 xs_all[kMlpWaves][kWave * 16];
    __shared__ FarRow rows_all[kMlpWaves][kWave];
    __shared__ int rlane_all[kMlpWaves][kWave];
    __shared__ float sacc_all[kMlpWaves][kWave][3];     // fx, fy, PE of each observer lane
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
        copy(reinterpret_cast<uint32_t*>(wfr.b3h), reinterpret_cast<const uint32_t*>(g.b3h), sizeof(g.b3h) / 4);
        copy(reinterpret_cast<uint32_t*>(wfr.istd), reinterpret_cast<const uint32_t*>(g.istd), kFarIn);
        if (threadIdx.x < kFarIn) wfr.nmi[threadIdx.x] = -g.mean[threadIdx.x] * g.istd[threadIdx.x];
        copy(reinterpret_cast<uint32_t*>(wfr.oscale), reinterpret_cast<const uint32_t*>(g.oscale), kFarOut);
    }
    __syncthreads();
    const int l = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int h = l >> 5, c = l & 31;
    const int64_t g = (int64_t)blockIdx.x * kMlpWaves + wv;
    const int n = a.ctrl->n_alive;
    if (g * kWave >= n) return;                  // no barrier below: a whole wave leaves
    __bf16* xs = xs_all[wv];
    FarRow* rows = rows_all[wv];
    int* rlane = rlane_all[wv];
    float (*sacc)[3] = sacc_all[wv];
    const FarInfer& fr = wfr;
    sacc[l][0] = 0.f; sacc[l][1] = 0.f; sacc[l][2] = 0.f;
    const int cnt = a.far_count[g];
    const uint32_t* list = a.far_list + g * a.far_cap;
    const float half_w = 0.5f * a.dom_w, half_h = 0.5f * a.dom_h;
    // software pipeline: once tile k's features are in LDS, the gathers of
    // tile k + 1 go into the same registers and stay in flight through tile
    // k's network pass
    FarIn cur = {};
    if (l < cnt) far_gather(a, g, list[l], cur);
    uint32_t e_next = kWave + l < cnt ? list[kWave + l] : 0u;     // tile 1's entries
    for (int base = 0; base < cnt; base += kWave) {
        // 1. features of row l
        const bool valid = base + l < cnt;
        float f[kFarIn];
#pragma unroll
        for (int k = 0; k < kFarIn; k++) f[k] = 0.f;
        FarRow rw = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        int lane_of = 0;
        if (valid) {
            lane_of = cur.lane;
            const float4 s4 = cur.s4;
            float dx = cur.c4.x - s4.x, dy = cur.c4.y - s4.y;
            if (a.periodic) { dx = torus_wrap(dx, a.dom_w, half_w); dy = torus_wrap(dy, a.dom_h, half_h); }
            float sg0, sc0, sp0, cp0;
            pair_features_fast(dx, dy, a.soft2, 2.f * cur.half, cur.c4.z, cur.mo, f, rw.ux, rw.uy, sg0, sc0, sp0,
                               cp0);
            rw.sg = a.grav ? s4.z * sg0 : 0.f;
            rw.sc = a.coul ? s4.w * sc0 : 0.f;
            rw.sp = a.grav ? s4.z * sp0 : 0.f;
            rw.cp = a.coul ? s4.w * cp0 : 0.f;
        }
        // Standardised inputs without branches: a row past the list holds
        // finite values whose outputs are discarded (MFMA output column n
        // depends on input column n only).
        float x[16];
#pragma unroll
        for (int k = 0; k < kFarIn; k++) x[k] = fmaf(f[k], fr.istd[k], fr.nmi[k]);
        x[kFarIn] = valid ? 1.f : 0.f;
#pragma unroll
        for (int k = kFarIn + 1; k < 16; k++) x[k] = 0.f;
        *reinterpret_cast<bf16x8*>(xs + l * 16) = pack8(x);
        *reinterpret_cast<bf16x8*>(xs + l * 16 + 8) = pack8(x + 8);
        rows[l] = rw;
        rlane[l] = lane_of;
        if (base + kWave + l < cnt) far_gather(a, g, e_next, cur);
        e_next = base + 2 * kWave + l < cnt ? list[base + 2 * kWave + l] : 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // 2. the network, then 3. rotate, scale and accumulate per observer
        //    lane: the force from outputs 0..3 (lanes h = 0), the potential
        //    from outputs 4, 5 (h = 1)
        floatx16 o[2];
        mlp_tile(fr, xs, l, o);
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            const int row = 32 * nb + c;
            if (base + row < cnt) {
                const FarRow q = rows[row];
                if (h == 0) {
                    const float gu = o[nb][0] * fr.oscale[0], gv = o[nb][1] * fr.oscale[1];
                    const float cu = o[nb][2] * fr.oscale[2], cv = o[nb][3] * fr.oscale[3];
                    const float fx = q.sg * (q.ux * gu - q.uy * gv) + q.sc * (q.ux * cu - q.uy * cv);
                    const float fy = q.sg * (q.uy * gu + q.ux * gv) + q.sc * (q.uy * cu + q.ux * cv);
                    atomicAdd(&sacc[rlane[row]][0], fx);
                    atomicAdd(&sacc[rlane[row]][1], fy);
                } else {
                    const float pe = q.sp * o[nb][0] * fr.oscale[4] + q.cp * o[nb][1] * fr.oscale[5];
                    atomicAdd(&sacc[rlane[row]][2], pe);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int i = (int)(g * kWave) + l;
    if (i < n) {
        a.fx[i] += sacc[l][0];
        a.fy[i] += sacc[l][1];
        a.pe[i] += sacc[l][2];
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
    const uint32_t e = a.far_list[g * a.far_cap + (row - a.row_off[lo])];
    const int node = (int)(e & kFarNodeMask);
    const int i = (int)(g * kWave) + (int)(e >> kFarLaneShift);
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
