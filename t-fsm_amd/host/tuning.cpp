#include "tuning.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tfsm {

namespace {

uint32_t *field(tfsm_eval_options &o, const std::string &name) {
    if (name == "grid_waves_per_cu") return &o.grid_waves_per_cu;
    if (name == "split_steps") return &o.split_steps;
    if (name == "split_fanout") return &o.split_fanout;
    if (name == "long_steps") return &o.long_steps;
    if (name == "long_waves_per_cu") return &o.long_waves_per_cu;
    if (name == "long_donate_every") return &o.long_donate_every;
    if (name == "long_fanout") return &o.long_fanout;
    if (name == "group_lanes") return &o.group_lanes;
    if (name == "match_order") return &o.match_order;
    if (name == "steps_per_round") return &o.steps_per_round;
    if (name == "refill_min") return &o.refill_min;
    if (name == "refill_max") return &o.refill_max;
    if (name == "prune_batch") return &o.prune_batch;
    if (name == "level_mode") return &o.level_mode;
    if (name == "hot_group_lanes") return &o.hot_group_lanes;
    if (name == "ldf_push") return &o.ldf_push;
    return nullptr;
}

}  // namespace

bool load_tuning_file(const char *path, tfsm_eval_options &opt, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open tuning file ") + path;
        return false;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    size_t i = s.find("\"options\"");
    if (i == std::string::npos) {
        err = "tuning file has no \"options\" object";
        return false;
    }
    i = s.find('{', i);
    const size_t end = i == std::string::npos ? i : s.find('}', i);
    if (end == std::string::npos) {
        err = "malformed \"options\" object";
        return false;
    }
    tfsm_eval_options o = opt;
    // "name": integer pairs, comma separated
    size_t p = i + 1;
    while (p < end) {
        const size_t q0 = s.find('"', p);
        if (q0 == std::string::npos || q0 >= end) break;
        const size_t q1 = s.find('"', q0 + 1);
        const size_t colon = s.find(':', q1);
        if (q1 == std::string::npos || colon == std::string::npos || colon >= end) {
            err = "malformed \"options\" object";
            return false;
        }
        const std::string name = s.substr(q0 + 1, q1 - q0 - 1);
        char *stop = nullptr;
        const unsigned long long v = std::strtoull(s.c_str() + colon + 1, &stop, 10);
        uint32_t *f = field(o, name);
        if (!f || stop == s.c_str() + colon + 1 || v > 0xffffffffull) {
            err = "unknown or malformed tuning option \"" + name + "\"";
            return false;
        }
        *f = (uint32_t)v;
        p = (size_t)(stop - s.c_str());
    }
    opt = o;
    return true;
}

}  // namespace tfsm
