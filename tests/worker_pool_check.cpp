// Stress check of the miner's worker pool (t-fsm_amd/host/worker_pool.h):
// parallel_for and Background jobs, interleaved as the miner uses them (one
// background growth per level, waited for after the level's evaluation),
// with every item run exactly once.  Built with a sanitizer by
// tests/test_worker_pool.py; prints "ok" and exits 0 on success.
#include "worker_pool.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

using tfsm::Background;
using tfsm::parallel_for;

static int fail(const char *what, size_t a, size_t b) {
    std::fprintf(stderr, "FAIL %s (%zu, %zu)\n", what, a, b);
    return 1;
}

int main() {
    std::vector<int> hits;
    uint64_t rng = 12345;
    auto next = [&]() {
        rng = rng * 6364136223846793005ull + 1442695040888963407ull;
        return (size_t)(rng >> 33);
    };
    for (int iter = 0; iter < 3000; ++iter) {
        const size_t n = next() % 40;          // includes 0 and 1
        const int threads = 1 + (int)(next() % 8);
        // a background job, with the caller busy meanwhile
        hits.assign(n, 0);
        {
            Background bg(n, threads, [&](size_t i) { hits[i] += 1; });
            volatile size_t spin = 0;
            for (size_t k = 0; k < (next() % 2000); ++k) spin = spin + k;
            if (next() & 1) bg.wait();         // else the destructor waits
        }
        for (size_t i = 0; i < n; ++i)
            if (hits[i] != 1) return fail("background item count", i, (size_t)hits[i]);
        // then a waited-for job on the same pool
        std::vector<size_t> sq(n, 0);
        parallel_for(n, threads, [&](size_t i) { sq[i] = i * i; });
        for (size_t i = 0; i < n; ++i)
            if (sq[i] != i * i) return fail("parallel_for item", i, sq[i]);
    }
    std::printf("ok\n");
    return 0;
}
