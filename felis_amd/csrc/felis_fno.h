// Internal entry points of the Fourier neural operator (felis_fno.hip) used
// by the jump path: forward of device maps enqueued after the caller's
// stream, and the network's channel counts.
#pragma once

#include <hip/hip_runtime.h>

#include "felis.h"

struct felis_ctx;

// X [B][64*64][cin] on the device; *out: the network's output buffer
// [B][64*64][cout], valid until its next call. The caller's stream st waits
// for the forward.
int felis_fno_forward_dev(felis_fno *n, int B, const float *X, const float **out, hipStream_t st);
int felis_fno_channels(const felis_fno *n, int *cin, int *cout);
// the same forward from the BF16 bits of the affine-applied inputs
// XB [B][64*64][cin] (the fused shape only: width 32, 8 modes, projection
// 64, at most 16 outputs; else FELIS_ERR_ARG)
int felis_fno_forward_bf16_dev(felis_fno *n, int B, const uint16_t *XB, const float **out, hipStream_t st);
// the network's input affine (device, either may be null: identity) and
// FELIS_OK when the BF16-input forward applies (FELIS_ERR_UNSUPPORTED else)
int felis_fno_input_affine_dev(const felis_fno *n, const float **mu, const float **isd);

// whether the context's field can be mapped (batched 3D path, square grid of
// 64 cells or more)
int felis_fno_maps_ok(const felis_ctx *ctx);
// floats of device work memory felis_fno_correct_dev needs
size_t felis_fno_work_floats(const felis_ctx *ctx);
// the learned correction of a jump, enqueued on ctx->stream after its push:
// feat the slices' features before the jump (device), work as above
int felis_fno_correct_dev(felis_ctx *ctx, felis_fno *net, const float *feat, float *work);
