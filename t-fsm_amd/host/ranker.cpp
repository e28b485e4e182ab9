#include "ranker.h"

#include <cstdio>

namespace tfsm {

int load_ranker_file(tfsm_ctx *ctx, const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return TFSM_ERR_ARGUMENT;
    tfsm_ranker_weights w;
    const size_t want = sizeof(w) / sizeof(float);
    const size_t got = std::fread(&w, sizeof(float), want, f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != want || extra) return TFSM_ERR_ARGUMENT;
    return tfsm_ranker_set(ctx, &w);
}

}  // namespace tfsm
