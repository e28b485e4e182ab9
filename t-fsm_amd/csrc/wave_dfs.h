// Wave-cooperative existence search: one (u, v) task per wave of 64 lanes.
//
// The same question as the lane-per-task search (is there an injective
// embedding with u -> v? reference gmatch.h:1198-1280) with the work of each
// DFS node spread over the wave: the candidates of depth d are the entries of
// the shortest backward-neighbour row that occur in every other
// backward-neighbour row (generateValidCandidateIndex, gmatch.h:1508-1546);
// the wave tests 64 consecutive entries of that row at once (one lane each),
// keeps the passing ones as a 64-bit mask, and descends into them one at a
// time in ascending order.  The chunk's candidates are staged in LDS; each
// other row is walked by a forward cursor in windows of 64 entries staged in
// LDS, and every lane whose candidate falls inside the window looks it up
// there (merge of two ascending runs, one global load per window instead of
// a binary search in global memory per lane).  Injectivity is checked against
// earlier vertices with the same label.
//
// The search state is uniform across the wave and lives in a WaveStack; every
// lane runs the uniform code with identical values.  Per-lane results travel
// only as ballots (64-bit masks), and staging writes one entry per lane, so
// the code is generic over the wave (W: ballot over the 64 lanes, stage into
// a shared array) and the data accessors (A), and the same source runs on the
// device and in a host check against the CPU oracle.  The includer defines
// TFSM_WAVE_FN (the function qualifiers for its side).
#pragma once

#include <cstdint>

namespace tfsm_wave {

constexpr uint32_t kLanes = 64;

enum WaveResult : uint32_t { kWaveFound = 1, kWaveExhausted = 2, kWaveAborted = 3 };

template <int MV>
struct WaveStack {
    uint32_t idx[MV];                // candidate index per pattern vertex
    uint32_t vid[MV];                // data vertex per depth
    uint32_t base[MV];               // next untested position of depth d's primary row
    uint32_t end[MV];                // end of depth d's primary row
    uint32_t chunk[MV];              // first position of the chunk `mask` refers to
    uint32_t prim[MV];               // which backward-neighbour row is primary
    unsigned long long mask[MV];     // passing candidates of the chunk not yet entered
    uint32_t rlo[MV][MV], rhi[MV][MV];   // backward-neighbour rows of depth d
    uint32_t cur[MV][MV];            // forward cursor into each non-primary row of depth d
    uint32_t cand[kLanes];           // the chunk under test (candidate index per lane)
    uint32_t win[kLanes];            // a window of a non-primary row
};

// Index of the lowest set bit (m != 0).
TFSM_WAVE_FN uint32_t wave_ctz64(unsigned long long m) { return (uint32_t)__builtin_ctzll(m); }

// Is c among the ascending values w[0 .. n-1]?  (LDS on the device)
TFSM_WAVE_FN bool wave_in_window(const uint32_t *w, uint32_t n, uint32_t c) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t x = w[mid];
        if (x == c) return true;
        if (x < c) lo = mid + 1; else hi = mid;
    }
    return false;
}

// Rows of depth d (the candidate indices of its backward neighbours' images)
// and the primary (shortest) one.
template <int MV, class A>
TFSM_WAVE_FN void wave_enter(const A &a, uint32_t u, uint32_t d, WaveStack<MV> &s) {
    const uint32_t nb = a.bn_n(u, d);
    uint32_t best = 0xffffffffu, prim = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        uint32_t lo, hi;
        a.row(u, d, k, s.idx[a.bn_vertex(u, d, k)], lo, hi);
        s.rlo[d][k] = lo;
        s.rhi[d][k] = hi;
        s.cur[d][k] = lo;
        if (hi - lo < best) {
            best = hi - lo;
            prim = k;
        }
    }
    s.prim[d] = prim;
    s.base[d] = s.rlo[d][prim];
    s.end[d] = s.rhi[d][prim];
    s.mask[d] = 0;
}

// A search is either a root task (u -> candidate ci, from depth 1) or a
// piece of one: the candidate indices of the vertices at depths < floor (in
// depth order), a range [lo, hi) of depth floor's primary row still to test
// and a chunk of it already tested (positions cpos + i for the set bits i of
// cmask) still to enter.
template <int MV, class A>
TFSM_WAVE_FN void wave_start_root(const A &a, uint32_t u, uint32_t ci, WaveStack<MV> &s) {
    s.idx[u] = ci;
    s.vid[0] = a.vid(u, ci);
    if (a.nv() > 1) wave_enter<MV>(a, u, 1, s);
}

// prefix: anything with prefix[j] (a pointer, or a reader of a record)
template <int MV, class A, class Prefix>
TFSM_WAVE_FN void wave_start_piece(const A &a, uint32_t u, uint32_t floor, const Prefix &prefix, uint32_t lo,
                                   uint32_t hi, uint32_t cpos, unsigned long long cmask, WaveStack<MV> &s) {
    for (uint32_t j = 0; j < floor; ++j) {
        const uint32_t x = a.order(u, j);
        s.idx[x] = prefix[j];
        s.vid[j] = a.vid(x, prefix[j]);
    }
    wave_enter<MV>(a, u, floor, s);
    s.base[floor] = lo;
    s.end[floor] = hi;
    s.chunk[floor] = cpos;
    s.mask[floor] = cmask;
}

// Runs a started search until it finds an embedding (kWaveFound: s.idx holds
// it), exhausts depth `floor` (kWaveExhausted) or h.abort() says stop
// (kWaveAborted, polled every h.abort_every rounds).  Every h.donate_every
// rounds it offers the pending work of its shallowest depth k that has some
// (untested positions [lo, hi) and tested candidates cpos + bits of cmask) to
// h.donate(k, lo, hi, cpos, cmask, s), keeping work of its own: below k it
// keeps the subtree it is in, at its own depth it keeps the tested candidates
// and gives only the untested range.  When h.donate returns true the offered
// work is gone from this search (the prefix of the piece is s.idx of the
// vertices at depths < k).  Both hooks return the same value on every lane.
// h.entered(d, c) is told of every candidate the search enters (a check hook;
// empty on the device).
template <int MV, class W, class A, class H>
TFSM_WAVE_FN uint32_t wave_run(const W &w, const A &a, uint32_t u, uint32_t floor, WaveStack<MV> &s, H &h) {
    const uint32_t n = a.nv();
    if (n == 1) return kWaveFound;
    uint32_t d = floor;
    for (uint32_t round = 1;; ++round) {
        if (h.abort_every && round % h.abort_every == 0 && h.abort()) return kWaveAborted;
        if (h.donate_every && round % h.donate_every == 0) {
            uint32_t k = floor;
            while (k < d && s.base[k] >= s.end[k] && s.mask[k] == 0) ++k;
            const unsigned long long gm = k < d ? s.mask[k] : 0ull;
            const bool gr = s.base[k] < s.end[k] && (k < d || s.mask[k] != 0);
            if ((gm || gr) && h.donate(k, gr ? s.base[k] : 0u, gr ? s.end[k] : 0u, s.chunk[k], gm, s)) {
                if (gr) s.end[k] = s.base[k];
                s.mask[k] &= ~gm;
            }
        }
        if (s.mask[d] == 0) {
            const uint32_t b0 = s.base[d], e = s.end[d];
            if (b0 >= e) {
                if (d == floor) return kWaveExhausted;
                --d;
                continue;
            }
            // test the next (up to) 64 entries of the primary row, one per lane
            const uint32_t ud = a.order(u, d), nb = a.bn_n(u, d), prim = s.prim[d];
            const uint32_t cn = e - b0 > kLanes ? kLanes : e - b0;
            w.stage(s.cand, cn, [&](uint32_t i) { return a.pool(b0 + i); });
            unsigned long long pass = cn == kLanes ? ~0ull : (1ull << cn) - 1ull;
            for (uint32_t k = 0; k < nb && pass; ++k) {
                if (k == prim) continue;
                // merge with row k: windows of 64 from the cursor; a lane is
                // settled once its candidate is <= the window's last entry
                unsigned long long open = pass, hit = 0;
                uint32_t cur = s.cur[d][k];
                const uint32_t hi = s.rhi[d][k];
                while (open) {
                    const uint32_t wn = hi - cur > kLanes ? kLanes : hi - cur;
                    if (wn == 0) break;                     // row exhausted: the open lanes miss
                    w.stage(s.win, wn, [&](uint32_t i) { return a.pool(cur + i); });
                    const uint32_t wmax = s.win[wn - 1];
                    const unsigned long long settled = open & w.ballot([&](uint32_t lane) -> bool {
                        return s.cand[lane] <= wmax;
                    });
                    hit |= settled & w.ballot([&](uint32_t lane) -> bool {
                        return ((settled >> lane) & 1ull) && wave_in_window(s.win, wn, s.cand[lane]);
                    });
                    open &= ~settled;
                    if (open) cur += wn;                    // later candidates lie beyond this window
                }
                s.cur[d][k] = cur;                          // later chunks start no earlier
                pass &= hit;
            }
            const uint32_t before = a.same_before(u, d);
            if (before && pass) {
                pass &= w.ballot([&](uint32_t lane) -> bool {
                    if (!((pass >> lane) & 1ull)) return false;
                    const uint32_t v = a.vid(ud, s.cand[lane]);
                    for (uint32_t m = before; m; m &= m - 1)
                        if (s.vid[wave_ctz64(m)] == v) return false;
                    return true;
                });
            }
            s.mask[d] = pass;
            s.chunk[d] = b0;
            s.base[d] = e - b0 > kLanes ? b0 + kLanes : e;
            continue;
        }
        // enter the lowest passing candidate of the chunk
        const unsigned long long m = s.mask[d];
        const uint32_t i = wave_ctz64(m);
        s.mask[d] = m & (m - 1);
        const uint32_t ud = a.order(u, d);
        const uint32_t c = a.pool(s.chunk[d] + i);   // s.cand may hold a deeper depth's chunk by now
        h.entered(d, c);
        s.idx[ud] = c;
        s.vid[d] = a.vid(ud, c);
        if (d + 1 == n) return kWaveFound;
        ++d;
        wave_enter<MV>(a, u, d, s);
    }
}

}  // namespace tfsm_wave
