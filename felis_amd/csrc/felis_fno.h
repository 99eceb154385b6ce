// Internal entry points of the Fourier neural operator (felis_fno.hip) used
// by the jump path: forward of device maps enqueued after the caller's
// stream, and the network's channel counts.
#pragma once

#include <hip/hip_runtime.h>

#include "felis.h"

// X [B][64*64][cin] on the device; *out: the network's output buffer
// [B][64*64][cout], valid until its next call. The caller's stream st waits
// for the forward.
int felis_fno_forward_dev(felis_fno *n, int B, const float *X, const float **out, hipStream_t st);
int felis_fno_channels(const felis_fno *n, int *cin, int *cout);
