// Lane-per-task existence search: is there an injective embedding of the
// pattern with u -> v?  (reference LFTJ, gmatch.h:1198-1280, with the
// candidate sets of generateValidCandidateIndex, gmatch.h:1508-1546.)
//
// One lane runs one (u, v) task, or a piece of one, as an iterative DFS.  At
// each depth the candidates of the depth's pattern vertex w are the entries
// of the shortest of its backward rows (the table rows A[x -> w][idx(x)] of
// its already matched neighbours x) that occur in every other backward row.
// The primary row is walked entry by entry; every other row keeps a
// forward-only cursor (candidates arrive in ascending order), so membership
// is a galloping merge.  Injectivity is checked against the matched vertices
// that share w's label (vertices with different labels never coincide).
//
// Two matching orders:
//   reference: the GQL order fixed per start vertex (generateGQLQueryPlan /
//              generateBN, gmatch.h:692-727, 803-828), taken from the plan;
//   adaptive:  at each descent the next vertex is the unmatched neighbour of
//              the matched set whose shortest backward row under the current
//              partial embedding is shortest (a row of length 0 backtracks
//              at once).  The row bounds of every cut edge are loaded
//              together, so choosing costs one round of loads, like entering
//              a fixed depth.  Whether an embedding exists does not depend on
//              the order, so the verdicts are those of the reference order.
//
// The search state is per lane: columns (C) indexed by pattern vertex (idx,
// and vid in the adaptive order) or by depth (pos/end of the primary row,
// vid in the reference order, meta), and cursor slots (cur/cend) handed out
// to the non-primary rows depth after depth.  The code is generic over the
// columns (LDS on the device) and the data accessors (A), so the same source
// runs in the search kernel and in a host check against the CPU oracle.  The
// includer defines TFSM_LANE_FN (the function qualifiers for its side).
#pragma once

#include <cstdint>

namespace tfsm_lane {

// Pattern edges the adaptive order handles (a pattern with more, or with more
// non-tree edges than cursor slots, is searched in the reference order).
constexpr uint32_t kCutMax = 12;
constexpr uint32_t kNone = 0xffu;

enum LaneStep : uint32_t { kLaneRunning = 0, kLaneFound = 1, kLaneExhausted = 2 };

// Depth metadata, 16 bits: the primary row among the depth's backward rows
// (bits 0-3), the depth's pattern vertex (4-7; adaptive order), the first
// cursor slot of its other rows (8-11) and their number (12-15).
TFSM_LANE_FN uint32_t meta_pack(uint32_t prim, uint32_t w, uint32_t secb, uint32_t secn) {
    return (prim & 15u) | ((w & 15u) << 4) | ((secb & 15u) << 8) | ((secn & 15u) << 12);
}
TFSM_LANE_FN uint32_t meta_prim(uint32_t m) { return m & 15u; }
TFSM_LANE_FN uint32_t meta_vertex(uint32_t m) { return (m >> 4) & 15u; }
TFSM_LANE_FN uint32_t meta_secb(uint32_t m) { return (m >> 8) & 15u; }
TFSM_LANE_FN uint32_t meta_secn(uint32_t m) { return (m >> 12) & 15u; }

TFSM_LANE_FN uint32_t lane_ctz(uint32_t m) { return (uint32_t)__builtin_ctz(m); }

// First position in [lo, hi) whose entry is >= c (galloping from lo).
template <class A>
TFSM_LANE_FN uint32_t lane_gallop(const A &a, uint32_t lo, uint32_t hi, uint32_t c) {
    if (lo >= hi || a.pool(lo) >= c) return lo;
    uint32_t step = 1, prev = lo;
    while (lo + step < hi && a.pool(lo + step) < c) {
        prev = lo + step;
        step <<= 1;
    }
    uint32_t x = prev + 1, y = lo + step < hi ? lo + step : hi;
    while (x < y) {
        const uint32_t mid = (x + y) >> 1;
        if (a.pool(mid) < c) x = mid + 1; else y = mid;
    }
    return x;
}

// A: nv(), adaptive(), pool(pos), vid(x, c), row(slot, r, lo, hi) (absolute
// bounds of row r of a slot's table), reference plan order(u, d), bn_n(u, d),
// bn_slot(u, d, k), slot_src(s), sec_base(u, d), same_before(u, d),
// same_after(u, d); adaptive order: ne(), edge_x(e), edge_y(e), slot_xy(e)
// (rows over C(x)), slot_yx(e), same_lab(x) (other vertices with x's label).
// C: idx(x), vid(j), pos(d), end(d), cur(g), cend(g), meta(d), all lvalues.
template <int MV>
struct Lane {
    uint32_t u = 0, n = 0, d = 0, floor = 1;
    uint32_t vis = 0;              // adaptive order: pattern vertices of depths 0 .. d
    bool adaptive = false;

    // Reference order: rows of depth dd, read once with their loads in
    // flight together; primary = `prim`, or the shortest when prim is kNone.
    template <class C, class A>
    TFSM_LANE_FN void enter_reference(C &c, const A &a, uint32_t dd, uint32_t prim) {
        const uint32_t nb = a.bn_n(u, dd);
        uint32_t rl[MV - 1], rh[MV - 1];
#pragma unroll
        for (uint32_t k = 0; k < MV - 1; ++k) {
            rl[k] = rh[k] = 0;
            if (k < nb) {
                const uint32_t s = a.bn_slot(u, dd, k);
                a.row(s, c.idx(a.slot_src(s)), rl[k], rh[k]);
            }
        }
        if (prim >= nb) {
            uint32_t best = 0xffffffffu;
#pragma unroll
            for (uint32_t k = 0; k < MV - 1; ++k)
                if (k < nb && rh[k] - rl[k] < best) {
                    best = rh[k] - rl[k];
                    prim = k;
                }
        }
        uint32_t g = a.sec_base(u, dd);
#pragma unroll
        for (uint32_t k = 0; k < MV - 1; ++k) {
            if (k >= nb) continue;
            if (k == prim) {
                c.pos(dd) = rl[k];
                c.end(dd) = rh[k];
                continue;
            }
            if (g < MV) {
                c.cur(g) = rl[k];
                c.cend(g) = rh[k];
            }
            ++g;
        }
        c.meta(dd) = meta_pack(prim, 0, 0, 0);
    }

    // Adaptive order: depth dd gets vertex w (chosen here when kNone: the
    // one with the shortest cut row) with primary row `prim` (the shortest
    // of w's rows when kNone).  w's backward rows are numbered in pattern
    // edge order, which a piece's receiver repeats.
    template <class C, class A>
    TFSM_LANE_FN void enter_adaptive(C &c, const A &a, uint32_t dd, uint32_t w, uint32_t prim) {
        const uint32_t ne = a.ne();
        uint32_t lo[kCutMax], hi[kCutMax], to[kCutMax];
#pragma unroll
        for (uint32_t e = 0; e < kCutMax; ++e) {
            lo[e] = hi[e] = 0;
            to[e] = kNone;
            if (e < ne) {
                const uint32_t x = a.edge_x(e), y = a.edge_y(e);
                const bool mx = (vis >> x) & 1u, my = (vis >> y) & 1u;
                const uint32_t dst = mx ? y : x;
                if (mx != my && (w == kNone || dst == w)) {
                    a.row(mx ? a.slot_xy(e) : a.slot_yx(e), c.idx(mx ? x : y), lo[e], hi[e]);
                    to[e] = dst;
                }
            }
        }
        if (w == kNone) {
            uint32_t best = 0xffffffffu;
#pragma unroll
            for (uint32_t e = 0; e < kCutMax; ++e)
                if (to[e] != kNone && hi[e] - lo[e] < best) {
                    best = hi[e] - lo[e];
                    w = to[e];
                }
        }
        uint32_t nb = 0, pick = 0, plen = 0xffffffffu;
#pragma unroll
        for (uint32_t e = 0; e < kCutMax; ++e)
            if (to[e] == w && w != kNone) {
                if (hi[e] - lo[e] < plen) {
                    plen = hi[e] - lo[e];
                    pick = nb;
                }
                ++nb;
            }
        if (prim == kNone) prim = pick;
        const uint32_t pm = dd ? c.meta(dd - 1) : 0u;
        const uint32_t secb = dd ? meta_secb(pm) + meta_secn(pm) : 0u;
        uint32_t k = 0, g = secb;
        c.pos(dd) = c.end(dd) = 0;           // no rows: the depth is exhausted at once
#pragma unroll
        for (uint32_t e = 0; e < kCutMax; ++e)
            if (to[e] == w && w != kNone) {
                if (k == prim) {
                    c.pos(dd) = lo[e];
                    c.end(dd) = hi[e];
                } else {
                    if (g < MV) {
                        c.cur(g) = lo[e];
                        c.cend(g) = hi[e];
                    }
                    ++g;
                }
                ++k;
            }
        c.meta(dd) = meta_pack(prim, w, secb, nb ? nb - 1 : 0);
        if (w != kNone) vis |= 1u << w;
    }

    template <class C, class A>
    TFSM_LANE_FN void start_root(C &c, const A &a, uint32_t u_, uint32_t ci) {
        u = u_;
        n = a.nv();
        adaptive = a.adaptive();
        floor = 1;
        d = 1;
        c.idx(u) = ci;
        if (adaptive) {
            vis = 1u << u;
            c.meta(0) = meta_pack(0, u, 0, 0);
            if (a.same_lab(u)) c.vid(u) = a.vid(u, ci);
            enter_adaptive(c, a, 1, kNone, kNone);
        } else {
            c.vid(0) = a.vid(u, ci);
            enter_reference(c, a, 1, kNone);
        }
    }

    // A piece: the candidates [lo, hi) of row `prim` at depth floor_, below
    // the candidate indices prefix[j] of depths j < floor_ (pattern vertices
    // pord[j], adaptive order, j <= floor_).
    template <class C, class A>
    TFSM_LANE_FN void start_piece(C &c, const A &a, uint32_t u_, uint32_t floor_, const uint32_t *prefix,
                                  const uint8_t *pord, uint32_t prim, uint32_t lo, uint32_t hi) {
        u = u_;
        n = a.nv();
        adaptive = a.adaptive();
        floor = floor_;
        d = floor_;
        if (adaptive) {
            vis = 0;
            for (uint32_t j = 0; j < floor_; ++j) {
                const uint32_t w = pord[j] & 15u;
                c.idx(w) = prefix[j];
                c.meta(j) = meta_pack(0, w, 0, 0);
                if (a.same_lab(w)) c.vid(w) = a.vid(w, prefix[j]);
                vis |= 1u << w;
            }
            enter_adaptive(c, a, floor_, pord[floor_] & 15u, prim);
        } else {
            for (uint32_t j = 0; j < floor_; ++j) {
                const uint32_t oj = a.order(u, j);
                c.idx(oj) = prefix[j];
                c.vid(j) = a.vid(oj, prefix[j]);
            }
            enter_reference(c, a, floor_, prim);
        }
        c.pos(floor_) = lo;
        c.end(floor_) = hi;
    }

    // Pattern vertex of depth j (j <= d).
    template <class C, class A>
    TFSM_LANE_FN uint32_t vertex_at(C &c, const A &a, uint32_t j) const {
        return adaptive ? meta_vertex(c.meta(j)) : a.order(u, j);
    }

    // Shallowest depth >= floor with untried candidates (d + 1 if none).
    template <class C>
    TFSM_LANE_FN uint32_t open_depth(C &c) const {
        uint32_t k = floor;
        while (k <= d && c.pos(k) >= c.end(k)) ++k;
        return k;
    }

    // One DFS step: test one candidate (and descend when it passes) or
    // backtrack one depth.
    template <class C, class A>
    TFSM_LANE_FN uint32_t step(C &c, const A &a) {
        const uint32_t pos = c.pos(d);
        if (pos >= c.end(d)) {
            if (adaptive) vis &= ~(1u << meta_vertex(c.meta(d)));
            --d;
            return d < floor ? kLaneExhausted : kLaneRunning;
        }
        c.pos(d) = pos + 1;
        const uint32_t cand = a.pool(pos);
        const uint32_t m = c.meta(d);
        const uint32_t prim = meta_prim(m);
        uint32_t w, nb, g0;
        if (adaptive) {
            w = meta_vertex(m);
            nb = meta_secn(m) + 1;
            g0 = meta_secb(m);
        } else {
            w = a.order(u, d);
            nb = a.bn_n(u, d);
            g0 = a.sec_base(u, d);
        }
        for (uint32_t k = 0, j = 0; k < nb; ++k) {
            if (k == prim) continue;
            const uint32_t g = g0 + j++;
            uint32_t at, hi;
            if (g < MV) {
                hi = c.cend(g);
                at = lane_gallop(a, c.cur(g), hi, cand);
                c.cur(g) = at;
            } else {
                // reference order only: more non-primary rows than cursor slots
                uint32_t lo;
                const uint32_t s = a.bn_slot(u, d, k);
                a.row(s, c.idx(a.slot_src(s)), lo, hi);
                at = lane_gallop(a, lo, hi, cand);
            }
            if (!(at < hi && a.pool(at) == cand)) return kLaneRunning;
        }
        if (adaptive) {
            const uint32_t sl = a.same_lab(w);
            if (sl) {
                const uint32_t v = a.vid(w, cand);
                for (uint32_t q = sl & vis & ~(1u << w); q; q &= q - 1)
                    if (c.vid(lane_ctz(q)) == v) return kLaneRunning;
                c.vid(w) = v;
            }
        } else {
            const uint32_t before = a.same_before(u, d);
            if (before | a.same_after(u, d)) {
                const uint32_t v = a.vid(w, cand);
                for (uint32_t q = before; q; q &= q - 1)
                    if (c.vid(lane_ctz(q)) == v) return kLaneRunning;
                c.vid(d) = v;
            }
        }
        c.idx(w) = cand;
        if (d + 1 == n) return kLaneFound;
        ++d;
        if (adaptive) enter_adaptive(c, a, d, kNone, kNone);
        else enter_reference(c, a, d, kNone);
        return kLaneRunning;
    }
};

}  // namespace tfsm_lane
