// Internal interface of the multilayer perceptron (csrc/mlp.hip) for the electron stage.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/pbagpu.h"

namespace pbagpu {
namespace mlp {
// Forward pass of `batch` rows of x into y (FP32, [batch][dims[n_layers]]) on stream s, with
// the device workspace of slot ws_slot; asynchronous.
int forward_into(const pba_mlp &net, int batch, const float *x, float *y, int ws_slot, hipStream_t s);
// PBAGPU_OK if the layer count, sizes and weight pointers are usable for inference.
int check_network(const pba_mlp *net);
}  // namespace mlp
}  // namespace pbagpu
