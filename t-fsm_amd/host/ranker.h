// Ranker weight files: 1665 little-endian float32 values in the layout of
// tfsm_ranker_weights (w1 [64][24], b1 [64], w2 [64], b2), as written by
// t-fsm_amd/train_ranker.py.
#pragma once

#include "../../include/tfsm_gpu.h"

namespace tfsm {

// Reads the file and installs it on the context; returns a tfsm_status.
int load_ranker_file(tfsm_ctx *ctx, const char *path);

}  // namespace tfsm
