// Internal interface of the convolutional closure network (cnn.hip) for the
// stepper (qgn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "qgn.h"

namespace qcnn {

// Forward pass on device data: x bf16 (n, H, W, cin) channels fastest,
// y fp32 (n, H, W, cout).  Keeps the activations for a following backward.
int forward(qgn_cnn* h, const unsigned short* x, int n, int H, int W, float* y, hipStream_t s);
int channels(const qgn_cnn* h, int* cin, int* cout);
// Allocate the activation buffers for n*H*W pixels ahead of a graph capture.
int reserve(qgn_cnn* h, long npix);
const char* last_error();

}  // namespace qcnn
