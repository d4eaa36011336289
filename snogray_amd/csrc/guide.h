// Learned directional proposal: a small MLP mapping a path vertex
// (position normalised to the scene box, outgoing direction) to a
// mixture of 8 von Mises-Fisher lobes over the incoming direction.
//
// Network: 6 inputs -> encoding (the 6 raw values and sin/cos of
// 2^k * pi * x for k = 0..5: 78 features, padded to 80) -> 3 hidden
// layers of 64 ReLU units -> 40 outputs (per lobe: mean direction
// 3, log concentration 1, mixture logit 1), padded to 48.  Weights are
// kept in fp32 (Adam master copy) and bf16 (MFMA operands); products
// accumulate in fp32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgr {

#define GD_IN      6
#define GD_FEAT    80   // 78 used
#define GD_H       64
#define GD_OUT     40
#define GD_OUTP    48
#define GD_LOBES   8
#define GD_NFREQ   6
#define GD_KAPPA_LO (-3.0f)   // log-concentration clamp
#define GD_KAPPA_HI 6.0f
#define GD_REC     16   // floats per training record

// Parameter offsets (floats): W[out][in] row-major, then the bias.
#define GD_W1 0
#define GD_B1 (GD_W1 + GD_H * GD_FEAT)
#define GD_W2 (GD_B1 + GD_H)
#define GD_B2 (GD_W2 + GD_H * GD_H)
#define GD_W3 (GD_B2 + GD_H)
#define GD_B3 (GD_W3 + GD_H * GD_H)
#define GD_W4 (GD_B3 + GD_H)
#define GD_B4 (GD_W4 + GD_OUTP * GD_H)
#define GD_NPARAM (GD_B4 + GD_OUTP)

// Minibatch workspace of one training step.
#define GD_BATCH     16384
#define GD_WG_SLICES 16       // partial sums of the weight gradients

struct RayPtrs {
  const float *r[8];         // ox oy oz dx dy dz t0 t1
};

struct GuideNet {
  const float *w32;          // fp32 parameters (biases are read here)
  const uint16_t *w16;       // bf16 copy of the parameters
};

// What the render kernels need of the guide.
struct GuideView {
  GuideNet net;
  float *in6;                // per queue row: normalised position, direction
  float *lobes;              // per queue row: 8 x (mu xyz, kappa, weight)
  float *rec;                // training records, GD_REC floats each
  uint32_t rec_stride;       // paths per training record
  uint32_t rec_count;        // records of the current chunk
  uint64_t rec_key;          // shuffles record slots (keyed permutation)
  float box_lo[3], box_inv[3];
  float mix;                 // probability of sampling the BSDF (0.5)
};

// Training workspace (device pointers) of one minibatch.
struct GuideTrain {
  uint16_t *a0, *a1, *a2, *a3;   // activations, bf16
  uint16_t *d4, *d3, *d2, *d1;   // output-side gradients of each layer, bf16
  float *tile_loss;              // per 16-row tile: sum w * -log q, sum w
  float *wpart;                  // GD_WG_SLICES x GD_NPARAM partial gradients
};

void launch_guide_inputs(const GuideView &G, const uint32_t *queue, const uint32_t *count,
                         const float *const ray[8], const float *ht, const uint32_t *hprim,
                         unsigned grid, hipStream_t st);
// Forward pass of N rows of 6 inputs (or of the queue rows when COUNT is
// given) into raw outputs (GD_OUT per row) or lobes (post-processed).
void launch_mlp_forward(const GuideNet &net, const float *in6, uint32_t n, const uint32_t *count,
                        float *out, int lobes, unsigned grid, hipStream_t st);
// One training step's passes over N records (GD_REC floats each, fields
// 0-5 inputs, 6-8 incoming direction, 9 weight).
void launch_mlp_rows(const GuideNet &net, const GuideTrain &T, const float *rec, uint32_t n,
                     hipStream_t st);
void launch_mlp_wgrad(const GuideTrain &T, uint32_t n, hipStream_t st);
// Reduces the partials into GRAD (loss and weight sums in aux[0..1]) and,
// if LR > 0, takes an Adam step on W32 / W16 with moments M, V.
void launch_mlp_adam(const GuideTrain &T, uint32_t n, float *grad, float *aux, float *w32,
                     uint16_t *w16, float *m, float *v, float lr, uint32_t step, hipStream_t st);

}  // namespace sgr
