// Internal interface of the convolutional closure network (cnn.hip) for the
// stepper (qgn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "qgn.h"

namespace qcnn {

// Forward pass on device data: x bf16 (n, H, W, cin) channels fastest,
// y fp32 (n, H, W, cout).  Keeps the activations for a following backward,
// unless `closure` (a model's RHS ran it), after which backward refuses.
int forward(qgn_cnn* h, const unsigned short* x, int n, int H, int W, float* y, hipStream_t s, bool closure);
int channels(const qgn_cnn* h, int* cin, int* cout);
// Allocate the activation buffers for n*H*W pixels ahead of a graph capture.
int reserve(qgn_cnn* h, long npix);
// A model attaching (+1) or detaching (-1) the network as its closure, or an
// ensemble capturing (+1) or releasing (-1) it: while held, the buffers are
// neither reallocated nor freed (passes on more pixels and
// qgn_cnn_destroy return an error).
void hold(qgn_cnn* h, int delta);
const char* last_error();

}  // namespace qcnn
