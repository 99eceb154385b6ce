// Device-side entry of the matrix-core perceptron (felis_mlp.hip) for the
// batched step's surrogate (felis_sur.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "felis.h"

// Y[rows][out] = net(X[rows][in]) for device arrays, enqueued on stream s
// (the network's own stream when s is null)
extern "C" int felis_mlp_forward_dev(felis_mlp *n, int rows, const float *X, float *Y, hipStream_t s);
