// mlp_host.hpp -- host-side hooks of the surrogate network shared between
// the training library part (mlp.hip) and the solver's surrogate stage.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/bats_hip.h"

namespace bats {
namespace mlp {
// Y[rows x 8] = MLP(X[rows x 16]) for device arrays, on `stream`
int forward_device(bats_mlp *m, const float *dX, float *dY, int rows,
                   hipStream_t stream);
// room for nMore samples at the end of the model's device dataset (inputs
// n x 16, targets n x 8, baseline fluxes n x 8, target weights n, FP32);
// commit_data then adds the ones written
int reserve_data(bats_mlp *m, long nMore, float **X, float **T, float **B,
                 float **J);
void commit_data(bats_mlp *m, long nAdded);
// device weights for fused use: fragment-ordered BF16 and FP32 flat
struct Weights {
  const __bf16 *frag;
  const float *w;
};
Weights weights(bats_mlp *m);
}  // namespace mlp
}  // namespace bats
