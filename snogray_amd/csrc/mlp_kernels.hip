// MLP kernels of the learned proposal: input rows of the ray queue,
// forward pass (MFMA bf16 x bf16 -> fp32, 16 rows per wave), the
// training passes (per-row forward + backward, weight gradients as
// row-reduction GEMMs in fixed order, Adam) -- all deterministic.
#include <hip/hip_runtime.h>

#include "dev_common.h"
#include "dev_guide.h"
#include "guide.h"
#include "kernels.h"

namespace sgr {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Products accumulate over k in tiles of 16: lane l holds the tile's
// A[l%16][4(l/16)+i] and B[4(l/16)+i][l%16], i < 4 (the fragments of
// v_mfma_f32_16x16x16_bf16).  gfx950's v_mfma_f32_16x16x32_bf16 takes
// two such tiles in one instruction: each lane passes its fragment of
// tile 2j followed by its fragment of tile 2j+1 as A and as B.  Both
// operands then list k in the same order, so the instruction sums A.B
// over both tiles; D[4(l/16)+i][l%16] comes back in acc[i] either way.
__device__ __forceinline__ f4v mfma_2tiles(s4v a0, s4v a1, s4v b0, s4v b1, f4v acc) {
  const s8v a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
  const s8v b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a),
                                                 __builtin_bit_cast(bf8v, b), acc, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma_tile(s4v a, s4v b, f4v acc) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
}
// acc + sum over the KT k-tiles of A(kt).B(kt), two tiles per instruction
// (an odd last tile alone); FA / FB give a lane's fragments of tile kt.
template<int KT, typename FA, typename FB>
__device__ __forceinline__ f4v mfma_k(FA fa, FB fb, f4v acc) {
#pragma unroll
  for (int kt = 0; kt + 1 < KT; kt += 2) acc = mfma_2tiles(fa(kt), fa(kt + 1), fb(kt), fb(kt + 1), acc);
  if (KT & 1) acc = mfma_tile(fa(KT - 1), fb(KT - 1), acc);
  return acc;
}

#define GD_WAVES 4  // waves per workgroup

// Orders this wave's LDS writes before its later LDS reads (lanes of one
// wave exchange tiles through LDS; no other wave touches them).
#define WAVE_SYNC()                                        \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
  } while (0)

// Per-wave LDS of a 16-row tile.
struct TileLds {
  uint16_t a0[16 * GD_FEAT];
  uint16_t a1[16 * GD_H];
  uint16_t a2[16 * GD_H];
  uint16_t a3[16 * GD_H];
  float out[16 * GD_OUTP];
  uint16_t d[16 * GD_H];   // gradient of the current layer's output
  uint16_t e[16 * GD_H];   // gradient being formed
  float rl[16], rw[16];    // per-row loss and weight
};

// out[16][N] = act[16][K] . W[N][K]^T + b (optional ReLU), on MFMA
// (mfma_k above).
template<int K, int N, bool RELU, bool F32OUT>
__device__ __forceinline__ void mlp_layer(const uint16_t *act, const uint16_t *W, const float *bias,
                                          uint16_t *outb, float *outf, int lane) {
  const int r = lane & 15, g = lane >> 4;
  for (int nt = 0; nt < N / 16; nt++) {
    const f4v acc = mfma_k<K / 16>(
        [&](int kt) { return *(const s4v *)(act + r * K + kt * 16 + 4 * g); },
        [&](int kt) { return *(const s4v *)(W + (size_t)(nt * 16 + r) * K + kt * 16 + 4 * g); },
        f4v{0.0f, 0.0f, 0.0f, 0.0f});
    const int col = nt * 16 + r;
    const float bb = bias[col];
    for (int i = 0; i < 4; i++) {
      const int row = 4 * g + i;
      float z = acc[i] + bb;
      if (RELU) z = fmaxf(z, 0.0f);
      if (F32OUT) outf[row * N + col] = z;
      else outb[row * N + col] = f2bf(z);
    }
  }
}

// Backward through one layer: e[16][KIN] = (d[16][NOUT] . W[NOUT][KIN])
// masked by act > 0 (the ReLU of the layer below).
template<int NOUT, int KIN>
__device__ __forceinline__ void mlp_back(const uint16_t *d, const uint16_t *W, const uint16_t *act,
                                         uint16_t *e, int lane) {
  const int r = lane & 15, g = lane >> 4;
  for (int nt = 0; nt < KIN / 16; nt++) {
    const f4v acc = mfma_k<NOUT / 16>(
        [&](int kt) { return *(const s4v *)(d + r * NOUT + kt * 16 + 4 * g); },
        [&](int kt) {
          s4v b;
          for (int i = 0; i < 4; i++)
            b[i] = (short)W[(size_t)(kt * 16 + 4 * g + i) * KIN + nt * 16 + r];
          return b;
        },
        f4v{0.0f, 0.0f, 0.0f, 0.0f});
    const int col = nt * 16 + r;
    for (int i = 0; i < 4; i++) {
      const int row = 4 * g + i;
      const bool on = (int16_t)act[row * KIN + col] > 0;  // bf16 of a ReLU output > 0
      e[row * KIN + col] = on ? f2bf(acc[i]) : (uint16_t)0;
    }
  }
}

__device__ __forceinline__ void encode_tile(const float *in6, uint32_t row0, uint32_t n,
                                            uint16_t *a0, int lane) {
  for (int e = lane; e < 16 * GD_FEAT; e += 64) {
    const int row = e / GD_FEAT, f = e - row * GD_FEAT;
    float v = 0.0f;
    if (row0 + row < n) v = gd_feature(in6 + (size_t)(row0 + row) * GD_IN, f);
    a0[e] = f2bf(v);
  }
}

// Post-processing of one row's raw outputs into lobes (mu normalised,
// kappa = exp(clamped), softmax weights).
__device__ __forceinline__ void to_lobes(const float *raw, float *L) {
  float mx = -1e30f;
  for (int k = 0; k < GD_LOBES; k++) mx = fmaxf(mx, raw[5 * k + 4]);
  float se = 0.0f;
  for (int k = 0; k < GD_LOBES; k++) se += expf(raw[5 * k + 4] - mx);
  for (int k = 0; k < GD_LOBES; k++) {
    const float *o = raw + 5 * k;
    float nn = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    float ix = nn > 1e-12f ? 1.0f / nn : 0.0f;
    L[5 * k + 0] = nn > 1e-12f ? o[0] * ix : 0.0f;
    L[5 * k + 1] = nn > 1e-12f ? o[1] * ix : 0.0f;
    L[5 * k + 2] = nn > 1e-12f ? o[2] * ix : 1.0f;
    L[5 * k + 3] = expf(fminf(fmaxf(o[3], GD_KAPPA_LO), GD_KAPPA_HI));
    L[5 * k + 4] = expf(o[4] - mx) / se;
  }
}

// Forward pass over rows, transposed form: out^T = W . act^T, with the
// weights as the MFMA A operand (lane l: W[m0 + l%16][k0 + 4(l/16) + i])
// and the activations as the B operand (lane l: act[row l%16][k0 +
// 4(l/16) + i]).  The D fragment of output tile m (lane l: out[row
// l%16][m0 + 4(l/16) + i]) is then exactly the B fragment of the next
// layer's k-tile m, so activations stay in registers between layers.
// The weights (bf16) and biases are staged in LDS once per workgroup,
// which then loops over row groups: every wave reads its weight
// fragments from LDS (rows padded by GD_WPAD so that the 16 rows a
// fragment spans fall in different banks) instead of re-reading 33 KB
// per group from the caches.  COUNT (device) overrides N when given.
#define FT 2              // 16-row tiles per wave and pass
#define GD_FWD_WAVES 8    // waves per forward workgroup
#define GD_WPAD 8         // bf16 padding per LDS weight row
struct FwdLds {
  uint16_t w1[GD_H * (GD_FEAT + GD_WPAD)];
  uint16_t w2[GD_H * (GD_H + GD_WPAD)];
  uint16_t w3[GD_H * (GD_H + GD_WPAD)];
  uint16_t w4[GD_OUTP * (GD_H + GD_WPAD)];
  float b1[GD_H], b2[GD_H], b3[GD_H], b4[GD_OUTP];
  float out[GD_FWD_WAVES][FT * 16 * GD_OUT];  // a pass's raw outputs, row-major
};

// Row-major N x K bf16 matrix from global memory into LDS rows of K +
// GD_WPAD, 4 values (8 bytes) per copy.
template<int N, int K>
__device__ __forceinline__ void stage_weights(const uint16_t *src, uint16_t *dst) {
  for (int e = threadIdx.x; e < N * K / 4; e += blockDim.x) {
    const int row = e / (K / 4), c = e - row * (K / 4);
    *(uint2 *)(dst + row * (K + GD_WPAD) + 4 * c) = *(const uint2 *)(src + row * K + 4 * c);
  }
}

template<int KT, int MT, bool RELU>
__device__ __forceinline__ void layer_t(const uint16_t *W, const float *bias, const s4v (&in)[FT][KT],
                                        f4v (&out)[FT][MT], int lane) {
  const int r = lane & 15, g = lane >> 4;
  constexpr int LDW = KT * 16 + GD_WPAD;
  for (int mt = 0; mt < MT; mt++) {
    f4v acc[FT];
    for (int t = 0; t < FT; t++) acc[t] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    const s4v *wrow = (const s4v *)(W + (mt * 16 + r) * LDW + 4 * g);
#pragma unroll
    for (int kt = 0; kt + 1 < KT; kt += 2) {  // two k-tiles per MFMA (mfma_k)
      const s4v a0 = wrow[kt * 4], a1 = wrow[(kt + 1) * 4];
      for (int t = 0; t < FT; t++)
        acc[t] = mfma_2tiles(a0, a1, in[t][kt], in[t][kt + 1], acc[t]);
    }
    if (KT & 1) {
      const s4v a = wrow[(KT - 1) * 4];
      for (int t = 0; t < FT; t++) acc[t] = mfma_tile(a, in[t][KT - 1], acc[t]);
    }
    for (int t = 0; t < FT; t++)
      for (int i = 0; i < 4; i++) {
        float z = acc[t][i] + bias[mt * 16 + 4 * g + i];
        out[t][mt][i] = RELU ? fmaxf(z, 0.0f) : z;
      }
  }
}

template<int MT>
__device__ __forceinline__ void to_bf16_frag(const f4v (&x)[FT][MT], s4v (&y)[FT][MT]) {
  for (int t = 0; t < FT; t++)
    for (int m = 0; m < MT; m++)
      for (int i = 0; i < 4; i++) y[t][m][i] = (short)f2bf(x[t][m][i]);
}

__global__ void __launch_bounds__(64 * GD_FWD_WAVES)
k_mlp_forward(GuideNet net, const float *in6, uint32_t n0, const uint32_t *count, float *out,
              int lobes) {
  __shared__ FwdLds S;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const uint32_t n = count ? *count : n0;
  const uint32_t ngroups = (n + 16 * FT - 1) / (16 * FT);
  if (blockIdx.x * GD_FWD_WAVES >= ngroups) return;  // the whole block: no group to take
  stage_weights<GD_H, GD_FEAT>(net.w16 + GD_W1, S.w1);
  stage_weights<GD_H, GD_H>(net.w16 + GD_W2, S.w2);
  stage_weights<GD_H, GD_H>(net.w16 + GD_W3, S.w3);
  stage_weights<GD_OUTP, GD_H>(net.w16 + GD_W4, S.w4);
  for (int e = threadIdx.x; e < GD_H; e += blockDim.x) {
    S.b1[e] = net.w32[GD_B1 + e];
    S.b2[e] = net.w32[GD_B2 + e];
    S.b3[e] = net.w32[GD_B3 + e];
  }
  for (int e = threadIdx.x; e < GD_OUTP; e += blockDim.x) S.b4[e] = net.w32[GD_B4 + e];
  __syncthreads();
  float *ob = S.out[wv];
  for (uint32_t grp = blockIdx.x * GD_FWD_WAVES + wv; grp < ngroups;
       grp += gridDim.x * GD_FWD_WAVES) {
    const uint32_t row0 = grp * 16 * FT;
    // encoded features straight into B fragments: this lane's row r of
    // each tile, features kt*16 + 4g + i
    s4v x0[FT][GD_FEAT / 16];
    for (int t = 0; t < FT; t++) {
      const uint32_t row = row0 + t * 16 + r;
      float u[GD_IN];
      for (int k = 0; k < GD_IN; k++) u[k] = row < n ? in6[(size_t)row * GD_IN + k] : 0.0f;
      for (int kt = 0; kt < GD_FEAT / 16; kt++)
        for (int i = 0; i < 4; i++)
          x0[t][kt][i] = (short)f2bf(row < n ? gd_feature_fast(u, kt * 16 + 4 * g + i) : 0.0f);
    }
    f4v h[FT][GD_H / 16];
    s4v hb[FT][GD_H / 16];
    layer_t<GD_FEAT / 16, GD_H / 16, true>(S.w1, S.b1, x0, h, lane);
    to_bf16_frag(h, hb);
    layer_t<GD_H / 16, GD_H / 16, true>(S.w2, S.b2, hb, h, lane);
    to_bf16_frag(h, hb);
    layer_t<GD_H / 16, GD_H / 16, true>(S.w3, S.b3, hb, h, lane);
    to_bf16_frag(h, hb);
    f4v o[FT][GD_OUTP / 16];
    layer_t<GD_H / 16, GD_OUTP / 16, false>(S.w4, S.b4, hb, o, lane);
    // rows gather their 40 outputs through LDS (the padding columns are
    // not kept)
    for (int t = 0; t < FT; t++)
      for (int m = 0; m < GD_OUTP / 16; m++)
        for (int i = 0; i < 4; i++) {
          const int c = m * 16 + 4 * g + i;
          if (c < GD_OUT) ob[(t * 16 + r) * GD_OUT + c] = o[t][m][i];
        }
    WAVE_SYNC();
    // a lane per row turns the row into lobes in place, then the wave
    // writes the pass's rows (contiguous in memory) with coalesced stores
    if (lobes && lane < 16 * FT && row0 + lane < n)
      to_lobes(ob + lane * GD_OUT, ob + lane * GD_OUT);
    WAVE_SYNC();
    const uint32_t rows = n - row0 < 16 * FT ? n - row0 : 16 * FT;
    float *dst = out + (size_t)row0 * GD_OUT;
    for (uint32_t idx = lane; idx < rows * GD_OUT; idx += 64) dst[idx] = ob[idx];
    WAVE_SYNC();
  }
}

// d(-log q)/d(raw outputs) of one row (40 values) for incoming
// direction W; returns log q.
__device__ __forceinline__ float mix_grad(const float *raw, float wx, float wy, float wz, float *g) {
  float L[GD_OUT], v[GD_LOBES];
  to_lobes(raw, L);
  float q = 0.0f;
  for (int k = 0; k < GD_LOBES; k++) {
    const float *o = L + 5 * k;
    v[k] = vmf_pdf(o[3], o[0] * wx + o[1] * wy + o[2] * wz);
    q += o[4] * v[k];
  }
  q = fmaxf(q, 1e-30f);
  const float iq = 1.0f / q;
  for (int k = 0; k < GD_LOBES; k++) {
    const float *o = L + 5 * k;
    const float *ro = raw + 5 * k;
    const float gam = o[4] * v[k] * iq;
    const float kap = o[3];
    const float c = o[0] * wx + o[1] * wy + o[2] * wz;
    // mean direction through the normalisation u -> u / |u|
    const float nn = sqrtf(ro[0] * ro[0] + ro[1] * ro[1] + ro[2] * ro[2]);
    const float s = nn > 1e-12f ? gam * kap / nn : 0.0f;
    g[5 * k + 0] = -s * (wx - c * o[0]);
    g[5 * k + 1] = -s * (wy - c * o[1]);
    g[5 * k + 2] = -s * (wz - c * o[2]);
    // log-concentration (zero where the clamp is active)
    float dk = 0.0f;
    if (ro[3] > GD_KAPPA_LO && ro[3] < GD_KAPPA_HI) {
      const float e2 = expf(-2.0f * kap);
      dk = gam * kap * (1.0f / kap - 2.0f * e2 / (1.0f - e2) + (c - 1.0f));
    }
    g[5 * k + 3] = -dk;
    // mixture logit
    g[5 * k + 4] = -(gam - o[4]);
  }
  return logf(q);
}

// Training rows: forward, output gradient (weighted by the record's
// weight), backward to every layer; activations and gradients stored
// for the weight-gradient pass.
__global__ void __launch_bounds__(64 * GD_WAVES)
k_mlp_rows(GuideNet net, GuideTrain TR, const float *rec, uint32_t n) {
  __shared__ TileLds lds[GD_WAVES];
  __shared__ float rowin[GD_WAVES][16 * GD_IN];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  TileLds &T = lds[wv];
  const uint32_t ntiles = (n + 15) / 16;
  for (uint32_t t = blockIdx.x * GD_WAVES + wv; t < ntiles; t += gridDim.x * GD_WAVES) {
    const uint32_t row0 = t * 16;
    for (int e = lane; e < 16 * GD_IN; e += 64) {
      const int row = e / GD_IN, f = e - row * GD_IN;
      rowin[wv][e] = row0 + row < n ? rec[(size_t)(row0 + row) * GD_REC + f] : 0.0f;
    }
    WAVE_SYNC();
    encode_tile(rowin[wv], 0, 16, T.a0, lane);
    WAVE_SYNC();
    mlp_layer<GD_FEAT, GD_H, true, false>(T.a0, net.w16 + GD_W1, net.w32 + GD_B1, T.a1, nullptr, lane);
    WAVE_SYNC();
    mlp_layer<GD_H, GD_H, true, false>(T.a1, net.w16 + GD_W2, net.w32 + GD_B2, T.a2, nullptr, lane);
    WAVE_SYNC();
    mlp_layer<GD_H, GD_H, true, false>(T.a2, net.w16 + GD_W3, net.w32 + GD_B3, T.a3, nullptr, lane);
    WAVE_SYNC();
    mlp_layer<GD_H, GD_OUTP, false, true>(T.a3, net.w16 + GD_W4, net.w32 + GD_B4, nullptr, T.out,
                                          lane);
    WAVE_SYNC();
    if (lane < 16) {
      float g[GD_OUT];
      float w = 0.0f, lq = 0.0f;
      const uint32_t row = row0 + lane;
      if (row < n) {
        const float *r = rec + (size_t)row * GD_REC;
        w = r[9] > 0.0f ? r[9] : 0.0f;
        lq = mix_grad(T.out + lane * GD_OUTP, r[6], r[7], r[8], g);
      }
      for (int k = 0; k < GD_OUTP; k++)
        T.d[lane * GD_OUTP + k] = (k < GD_OUT && w > 0.0f) ? f2bf(w * g[k]) : (uint16_t)0;
      T.rl[lane] = w > 0.0f ? -w * lq : 0.0f;
      T.rw[lane] = w;
    }
    WAVE_SYNC();
    if (lane == 0) {
      float sl = 0.0f, sw = 0.0f;
      for (int i = 0; i < 16; i++) {
        sl += T.rl[i];
        sw += T.rw[i];
      }
      TR.tile_loss[2 * t] = sl;
      TR.tile_loss[2 * t + 1] = sw;
    }
    // store activations and the output gradient
    for (int e = lane; e < 16 * GD_FEAT; e += 64) TR.a0[(size_t)row0 * GD_FEAT + e] = T.a0[e];
    for (int e = lane; e < 16 * GD_H; e += 64) {
      TR.a1[(size_t)row0 * GD_H + e] = T.a1[e];
      TR.a2[(size_t)row0 * GD_H + e] = T.a2[e];
      TR.a3[(size_t)row0 * GD_H + e] = T.a3[e];
    }
    for (int e = lane; e < 16 * GD_OUTP; e += 64) TR.d4[(size_t)row0 * GD_OUTP + e] = T.d[e];
    WAVE_SYNC();
    mlp_back<GD_OUTP, GD_H>(T.d, net.w16 + GD_W4, T.a3, T.e, lane);
    WAVE_SYNC();
    for (int e = lane; e < 16 * GD_H; e += 64) TR.d3[(size_t)row0 * GD_H + e] = T.e[e];
    mlp_back<GD_H, GD_H>(T.e, net.w16 + GD_W3, T.a2, T.d, lane);
    WAVE_SYNC();
    for (int e = lane; e < 16 * GD_H; e += 64) TR.d2[(size_t)row0 * GD_H + e] = T.d[e];
    mlp_back<GD_H, GD_H>(T.d, net.w16 + GD_W2, T.a1, T.e, lane);
    WAVE_SYNC();
    for (int e = lane; e < 16 * GD_H; e += 64) TR.d1[(size_t)row0 * GD_H + e] = T.e[e];
    WAVE_SYNC();
  }
}

// Weight gradients: dW[n][k] = sum over rows of d[row][n] * a[row][k],
// one 16x16 tile of one layer per wave over one slice of the rows
// (MFMA with the rows as the reduction dimension); bias gradients
// dB[n] = sum over rows of d[row][n].  Partials go to slice-major slots.
struct WTile {
  int layer, mt, kt;
};
__device__ __forceinline__ WTile wtile(int i) {
  // layer 1: 4 x 5 tiles, layers 2, 3: 4 x 4, layer 4: 3 x 4
  if (i < 20) return WTile{1, i / 5, i % 5};
  i -= 20;
  if (i < 16) return WTile{2, i / 4, i % 4};
  i -= 16;
  if (i < 16) return WTile{3, i / 4, i % 4};
  i -= 16;
  return WTile{4, i / 4, i % 4};
}
#define GD_WTILES 64

__global__ void __launch_bounds__(64 * GD_WAVES) k_mlp_wgrad(GuideTrain TR, uint32_t n) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t job = blockIdx.x * GD_WAVES + wv;
  if (job >= GD_WTILES * GD_WG_SLICES) return;
  const int slice = (int)(job / GD_WTILES);
  const WTile T = wtile((int)(job % GD_WTILES));
  const uint16_t *d, *a;
  int nout, kin, woff, boff;
  switch (T.layer) {
    case 1: d = TR.d1; a = TR.a0; nout = GD_H; kin = GD_FEAT; woff = GD_W1; boff = GD_B1; break;
    case 2: d = TR.d2; a = TR.a1; nout = GD_H; kin = GD_H; woff = GD_W2; boff = GD_B2; break;
    case 3: d = TR.d3; a = TR.a2; nout = GD_H; kin = GD_H; woff = GD_W3; boff = GD_B3; break;
    default: d = TR.d4; a = TR.a3; nout = GD_OUTP; kin = GD_H; woff = GD_W4; boff = GD_B4; break;
  }
  const uint32_t rows = (n + 15) / 16 * 16;  // whole tiles (padding rows are zero)
  const uint32_t per = (rows / 16 + GD_WG_SLICES - 1) / GD_WG_SLICES * 16;
  const uint32_t r0 = slice * per, r1 = r0 + per < rows ? r0 + per : rows;
  const int r = lane & 15, g = lane >> 4;
  // k = the rows of the slice, 16 per tile, two tiles per MFMA
  auto frag = [&](const uint16_t *m, int ld, int col, uint32_t rb) {
    s4v v;
    for (int i = 0; i < 4; i++) v[i] = (short)m[(size_t)(rb + 4 * g + i) * ld + col];
    return v;
  };
  const int dc = T.mt * 16 + r, ac = T.kt * 16 + r;
  f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t rb = r0;
  for (; rb + 32 <= r1; rb += 32)
    acc = mfma_2tiles(frag(d, nout, dc, rb), frag(d, nout, dc, rb + 16), frag(a, kin, ac, rb),
                      frag(a, kin, ac, rb + 16), acc);
  if (rb < r1) acc = mfma_tile(frag(d, nout, dc, rb), frag(a, kin, ac, rb), acc);
  float *out = TR.wpart + (size_t)slice * GD_NPARAM;
  for (int i = 0; i < 4; i++) {
    const int m = T.mt * 16 + 4 * g + i, k = T.kt * 16 + r;
    out[woff + m * kin + k] = acc[i];
  }
  if (T.kt == 0 && lane < 16) {
    float s = 0.0f;
    for (uint32_t row = r0; row < r1; row++) s += bf2f(d[(size_t)row * nout + T.mt * 16 + lane]);
    out[boff + T.mt * 16 + lane] = s;
  }
}

// Loss and weight sums over the tiles, in tile order (one wave).
__global__ void __launch_bounds__(64) k_mlp_loss(const float *tile_loss, uint32_t ntiles,
                                                 float *aux) {
  if (threadIdx.x != 0) return;
  float sl = 0.0f, sw = 0.0f;
  for (uint32_t t = 0; t < ntiles; t++) {
    sl += tile_loss[2 * t];
    sw += tile_loss[2 * t + 1];
  }
  aux[0] = sl;
  aux[1] = sw;
}

// grad = (sum of slice partials in slice order) / sum of weights; Adam
// (beta1 0.9, beta2 0.999, eps 1e-8) when lr > 0.
__global__ void __launch_bounds__(256) k_mlp_adam(const float *wpart, const float *aux, float *grad,
                                                  float *w32, uint16_t *w16, float *m, float *v,
                                                  float lr, uint32_t step) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= GD_NPARAM) return;
  float s = 0.0f;
  for (int k = 0; k < GD_WG_SLICES; k++) s += wpart[(size_t)k * GD_NPARAM + p];
  const float sw = aux[1];
  const float gr = sw > 0.0f ? s / sw : 0.0f;
  grad[p] = gr;
  if (lr > 0.0f) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)step));
    const float vh = vv / (1.0f - powf(b2, (float)step));
    const float w = w32[p] - lr * mh / (sqrtf(vh) + 1e-8f);
    w32[p] = w;
    w16[p] = f2bf(w);
  }
}

// Network inputs of the hit rows of a ray queue: hit position
// normalised to the scene box and the outgoing direction -d/|d|.
__global__ void __launch_bounds__(256)
k_guide_inputs(GuideView G, const uint32_t *queue, const uint32_t *count, RayPtrs R,
               const float *ht, const uint32_t *hprim) {
  const uint32_t n = *count;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t p = queue[j];
    float *o = G.in6 + (size_t)j * GD_IN;
    if (hprim[j] == 0xffffffffu) {  // miss (hits are stored by queue position)
      for (int k = 0; k < GD_IN; k++) o[k] = 0.0f;
      continue;
    }
    const float t = ht[j];
    const float dx = R.r[3][p], dy = R.r[4][p], dz = R.r[5][p];
    const float x = R.r[0][p] + dx * t, y = R.r[1][p] + dy * t, z = R.r[2][p] + dz * t;
    const float il = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    o[0] = (x - G.box_lo[0]) * G.box_inv[0];
    o[1] = (y - G.box_lo[1]) * G.box_inv[1];
    o[2] = (z - G.box_lo[2]) * G.box_inv[2];
    o[3] = -dx * il;
    o[4] = -dy * il;
    o[5] = -dz * il;
  }
}

// ---------------------------------------------------------------------------
void launch_guide_inputs(const GuideView &G, const uint32_t *queue, const uint32_t *count,
                         const float *const ray[8], const float *ht, const uint32_t *hprim,
                         unsigned grid, hipStream_t st) {
  RayPtrs R;
  for (int k = 0; k < 8; k++) R.r[k] = ray[k];
  hipLaunchKernelGGL(k_guide_inputs, dim3(grid), dim3(256), 0, st, G, queue, count, R, ht, hprim);
}

void launch_mlp_forward(const GuideNet &net, const float *in6, uint32_t n, const uint32_t *count,
                        float *out, int lobes, unsigned grid, hipStream_t st) {
  // two workgroups fit a CU (their LDS); each loops over row groups
  const unsigned cap = 2u * launch_cus();
  grid = grid < cap ? grid : cap;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64 * GD_FWD_WAVES), 0, st, net, in6, n,
                     count, out, lobes);
}

void launch_mlp_rows(const GuideNet &net, const GuideTrain &T, const float *rec, uint32_t n,
                     hipStream_t st) {
  const uint32_t tiles = (n + 15) / 16;
  unsigned grid = (tiles + GD_WAVES - 1) / GD_WAVES;
  if (grid == 0) grid = 1;
  hipLaunchKernelGGL(k_mlp_rows, dim3(grid), dim3(64 * GD_WAVES), 0, st, net, T, rec, n);
}

void launch_mlp_wgrad(const GuideTrain &T, uint32_t n, hipStream_t st) {
  const unsigned grid = (GD_WTILES * GD_WG_SLICES + GD_WAVES - 1) / GD_WAVES;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(grid), dim3(64 * GD_WAVES), 0, st, T, n);
}

void launch_mlp_adam(const GuideTrain &T, uint32_t n, float *grad, float *aux, float *w32,
                     uint16_t *w16, float *m, float *v, float lr, uint32_t step, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(64), 0, st, T.tile_loss, (n + 15) / 16, aux);
  hipLaunchKernelGGL(k_mlp_adam, dim3((GD_NPARAM + 255) / 256), dim3(256), 0, st, T.wpart, aux,
                     grad, w32, w16, m, v, lr, step);
}

}  // namespace sgr
