// Tuning files written by t-fsm_amd/autotune.py: a JSON object whose
// "options" member maps search-kernel option names (the fields of
// tfsm_eval_options that shape the kernel) to integers.
#pragma once

#include <string>

#include "../../include/tfsm_gpu.h"

namespace tfsm {

// Reads the "options" object of a tuning file into opt (only the shape
// fields; mode, order, trace and retention are left alone).  Returns false
// and sets err for a missing file, malformed JSON or an unknown option name.
bool load_tuning_file(const char *path, tfsm_eval_options &opt, std::string &err);

}  // namespace tfsm
