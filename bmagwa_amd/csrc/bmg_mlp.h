// bmg_mlp.h -- the learned-proposal network (16 -> 256 -> 256 -> 1 MLP,
// bf16 weights and activations, fp32 accumulation on the MFMA units).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_IN 16
#define MLP_HID 256
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_IN * MLP_HID)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_HID)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_HID * MLP_HID)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_HID)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_HID)
#define MLP_NPAR (MLP_OFF_B3 + 1)
#define MLP_ROWS 64          // rows per workgroup

namespace bmg {

// Device buffers of one network.  P/M/V: fp32 master parameters and Adam
// moments; G: fp32 gradient accumulator (cleared by the Adam kernel);
// bf16 copies laid out so that every MFMA operand is a 16-byte load with
// the reduction index contiguous.
#define MLP_GP_WG 64         // workgroups per training launch (partial gradients kept)

struct MlpArgs {
  float* P; float* G; float* M; float* V;
  float* GP;       // [MLP_GP_WG][MLP_NPAR] per-workgroup gradients of one launch
  uint16_t* W1t;   // [256 h][16 k]
  uint16_t* W2t;   // [256 g][256 h]  (forward: n = g, k = h)
  uint16_t* W2b;   // [256 h][256 g]  (backward: n = h, k = g)
  uint16_t* w3b;   // [256]
};

int mlp_alloc(MlpArgs* m);
void mlp_free(MlpArgs* m);
// refresh the bf16 copies from P (after P was written from the host)
int mlp_launch_copies(const MlpArgs& m, hipStream_t s);
// out[r] = network(X[r]) for X: rows x 16 bf16
int mlp_launch_forward(const MlpArgs& m, const uint16_t* X, int64_t rows, float* out, hipStream_t s);
// G += gradient of the mean Huber loss over `rows` (labels: fp32), summed
// in a fixed order (workgroup partials, then an ordered reduction), so the
// result does not depend on scheduling; out (optional) receives the forward
// outputs
int mlp_launch_train(const MlpArgs& m, const uint16_t* X, const float* labels, int64_t rows,
                     float* out, hipStream_t s);
// one Adam step (t counts from 1) from G, which it clears; bf16 copies follow
int mlp_launch_adam(const MlpArgs& m, float lr, int t, hipStream_t s);
// round-to-nearest-even fp32 -> bf16 on the host (the device rule)
uint16_t bf16_bits(float f);

}  // namespace bmg
