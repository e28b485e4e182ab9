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
// vid in the reference order, meta), and cursor slots (cur/cend/curv) handed
// out to the non-primary rows depth after depth.  Each step is one chain of
// dependent loads, kept short: a cursor slot caches the entry under its
// cursor (curv), so a candidate below it is rejected without a load; the
// next primary entry is prefetched during the current step; the vertex id
// for the injectivity test is loaded beside the membership test; entering a
// depth loads the row offsets of every cut edge in one round.  The code is
// generic over the columns (LDS on the device) and the data accessors (A),
// so the same source runs in the search kernel and in a host check against
// the CPU oracle.  The includer defines TFSM_LANE_FN (the function
// qualifiers for its side).
#pragma once

#include <cstdint>

namespace tfsm_lane {

// Pattern edges the adaptive order handles (a pattern with more, or with more
// non-tree edges than cursor slots, is searched in the reference order).
constexpr uint32_t kCutMax = 12;
constexpr uint32_t kNone = 0xffu;
constexpr uint32_t kUnknown = 0xffffffffu;   // curv: entry under the cursor not loaded yet

enum LaneStep : uint32_t { kLaneRunning = 0, kLaneFound = 1, kLaneExhausted = 2 };

// The matching orders a search instance handles: either, per pattern at run
// time, or only one (the other's code is then not compiled in).
enum LaneOrder : int { kOrderMixed = 0, kOrderAdaptive = 1, kOrderReference = 2 };

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

// First position in [lo, hi) whose entry is >= c (galloping from lo); `val`
// gets that entry (kUnknown at hi), known from the probes without a reload.
template <class A>
TFSM_LANE_FN uint32_t lane_gallop(const A &a, uint32_t lo, uint32_t hi, uint32_t c, uint32_t &val) {
    val = kUnknown;
    if (lo >= hi) return lo;
    uint32_t v = a.pool(lo);
    if (v >= c) {
        val = v;
        return lo;
    }
    uint32_t step = 1, prev = lo, yv = kUnknown;
    while (lo + step < hi) {
        v = a.pool(lo + step);
        if (v >= c) {
            yv = v;
            break;
        }
        prev = lo + step;
        step <<= 1;
    }
    uint32_t x = prev + 1, y = lo + step < hi ? lo + step : hi;
    while (x < y) {
        const uint32_t mid = (x + y) >> 1;
        const uint32_t m = a.pool(mid);
        if (m < c) {
            x = mid + 1;
        } else {
            y = mid;
            yv = m;
        }
    }
    if (x < hi) val = yv;
    return x;
}

// A: nv(), adaptive(), pool(pos), vid(x, c), row(slot, r, lo, hi) (absolute
// bounds of row r of a slot's table), reference plan order(u, d), bn_n(u, d),
// bn_slot(u, d, k), slot_src(s), sec_base(u, d), same_before(u, d),
// same_after(u, d); adaptive order: edge_pack(i) (pattern edges, one byte
// each, x | y << 4, four to a word, 0xff past the last), cut_ao(s) (where
// the row offsets of slot s start; slot 2e is edge e's table from x, over
// C(x), and 2e + 1 the one from y) and row_at(ao, r, lo, hi) (row r of the
// table whose offsets start at ao), same_lab(x) (other vertices with x's
// label), edge_max() (at least the edge count of every pattern the
// caller serves; uniform on the device).
// C: idx(x), vid(j), pos(d), end(d), cur(g), cend(g), curv(g), meta(d), all
// lvalues.
// ME: the most pattern edges this instance's adaptive order handles (a
// multiple of 4 up to kCutMax, the edges packed four to a word); a launch
// whose patterns have at most 8 edges runs the ME = 8 instance, whose
// per-edge arrays take half the registers.
template <int MV, int ORD = kOrderMixed, int ME = (int)kCutMax>
struct Lane {
    static_assert(ME % 4 == 0 && ME > 0 && ME <= (int)kCutMax, "edges packed four to a word");
    uint32_t u = 0, n = 0, d = 0, floor = 1;
    // interleaved root search: of depth `floor`'s candidates this lane tests
    // every stride-th one (the other lanes of its group take the others)
    uint32_t stride = 1;
    uint32_t vis = 0;              // adaptive order: pattern vertices of depths 0 .. d
    bool adaptive_rt = false;      // the pattern's order (kOrderMixed)
    TFSM_LANE_FN bool adaptive() const { return ORD == kOrderAdaptive || (ORD == kOrderMixed && adaptive_rt); }
    uint32_t pf_pos = kUnknown, pf_val = 0;   // prefetched pool entry (the pool is read-only)
    // the depth and candidate the last step entered (kNone: it entered none);
    // read by the host check only (dead on the device)
    uint32_t last_depth = kNone, last_cand = 0;
    // Adaptive order: the pattern's constants, loaded once per task (the
    // search kernel's occupancy is set by its LDS, so registers are spare):
    // the packed edges, where each slot's row pairs start, and for patterns
    // of up to 8 vertices the same-label masks (8 bits per vertex).
    uint32_t ep[ME / 4] = {};
    uint32_t aor[2 * ME] = {};
    unsigned long long slab = 0;     // a value, not an array: a run-time index would put it in scratch

    template <class A>
    TFSM_LANE_FN void load_edges(const A &a) {
#pragma unroll
        for (uint32_t i = 0; i < (uint32_t)ME / 4; ++i) ep[i] = a.edge_pack(i);
        const uint32_t me = a.edge_max();
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            // slots of absent edges are never selected; slot 0 always exists
            const bool edge = ((ep[e >> 2] >> ((e & 3u) * 8u)) & 0xffu) != 0xffu;
            aor[2 * e] = a.cut_ao(edge ? 2 * e : 0u);
            aor[2 * e + 1] = a.cut_ao(edge ? 2 * e + 1 : 0u);
        }
        if (MV <= 8) {
            slab = 0;
            for (uint32_t x = 0; x < n && x < 8; ++x) slab |= (unsigned long long)(a.same_lab(x) & 0xffu) << (8u * x);
        }
    }
    template <class A>
    TFSM_LANE_FN uint32_t same_lab_of(const A &a, uint32_t x) const {
        if (MV <= 8) return (uint32_t)(slab >> (8u * x)) & 0xffu;
        return a.same_lab(x);
    }

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
                c.curv(g) = kUnknown;
            }
            ++g;
        }
        c.meta(dd) = meta_pack(prim, 0, 0, 0);
    }

    // Adaptive order: depth dd gets vertex w (chosen here when kNone: the
    // one with the shortest cut row) with primary row `prim` (the shortest
    // of w's rows when kNone).  w's backward rows are numbered in pattern
    // edge order, which a piece's receiver repeats.  The edges come from
    // registers.  Every edge issues its loads whether it is cut or not (a
    // non-cut edge reads row 0 of slot 0, which always exists) and nothing
    // loaded is used before all are issued, so the offsets of all cut rows
    // arrive in one round trip instead of one per edge.
    template <class C, class A>
    TFSM_LANE_FN void enter_adaptive(C &c, const A &a, uint32_t dd, uint32_t w, uint32_t prim) {
        // me: the most edges of any pattern of the launch (uniform), so the
        // wave skips the slots no pattern has
        const uint32_t me = a.edge_max();
        uint32_t sel[ME], rr[ME];
        uint32_t cut = 0, dlo = 0, dhi = 0;     // cut edges; their unmatched ends, 4 bits each
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            const uint32_t xy = (ep[e >> 2] >> ((e & 3u) * 8u)) & 0xffu;
            const bool edge = xy != 0xffu;
            const uint32_t x = edge ? xy & 15u : 0u, y = edge ? xy >> 4 : 0u;
            const uint32_t mx = (vis >> x) & 1u, my = (vis >> y) & 1u;
            const uint32_t dst = mx ? y : x;
            const bool is_cut = edge && mx != my && (w == kNone || dst == w);
            sel[e] = is_cut ? 2u * e + (mx ? 0u : 1u) : 0u;
            rr[e] = c.idx(is_cut ? (mx ? x : y) : 0u);
            if (!is_cut) rr[e] = 0;
            cut |= (is_cut ? 1u : 0u) << e;
            if (e < 8) dlo |= (is_cut ? dst : 0u) << (4u * e);
            else dhi |= (is_cut ? dst : 0u) << (4u * (e - 8));
        }
        uint32_t ao[ME];
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            ao[e] = (sel[e] & 1u) ? aor[2 * e + 1] : (sel[e] ? aor[2 * e] : aor[0]);
        }
        uint32_t lo[ME], hi[ME];
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            a.row_at(ao[e], rr[e], lo[e], hi[e]);
        }
        auto dst_of = [&](uint32_t e) { return e < 8 ? (dlo >> (4u * e)) & 15u : (dhi >> (4u * (e - 8))) & 15u; };
        auto len_of = [&](uint32_t e) { return ((cut >> e) & 1u) ? hi[e] - lo[e] : 0xffffffffu; };
        if (w == kNone) {
            uint32_t best = 0xffffffffu;
#pragma unroll
            for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
                if (e >= me) continue;
                const uint32_t l = len_of(e);
                if (((cut >> e) & 1u) && l < best) {
                    best = l;
                    w = dst_of(e);
                }
            }
        }
        uint32_t nb = 0, pick = 0, plen = 0xffffffffu;
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            if (((cut >> e) & 1u) && dst_of(e) == w) {
                if (len_of(e) < plen) {
                    plen = len_of(e);
                    pick = nb;
                }
                ++nb;
            }
        }
        if (prim == kNone) prim = pick;
        const uint32_t pm = dd ? c.meta(dd - 1) : 0u;
        const uint32_t secb = dd ? meta_secb(pm) + meta_secn(pm) : 0u;
        uint32_t k = 0, g = secb, plo = 0, phi = 0;
#pragma unroll
        for (uint32_t e = 0; e < (uint32_t)ME; ++e) {
            if (e >= me) continue;
            if (((cut >> e) & 1u) && dst_of(e) == w) {
                if (k == prim) {
                    plo = lo[e];
                    phi = hi[e];
                } else {
                    if (g < MV) {
                        c.cur(g) = lo[e];
                        c.cend(g) = hi[e];
                        c.curv(g) = kUnknown;
                    }
                    ++g;
                }
                ++k;
            }
        }
        c.pos(dd) = plo;                     // no rows: the depth is exhausted at once
        c.end(dd) = phi;
        c.meta(dd) = meta_pack(prim, w, secb, nb ? nb - 1 : 0);
        if (w != kNone) vis |= 1u << w;
    }

    // A root task (u, ci).  With stride g > 1 and phase p < g, this lane
    // tests only depth 1's candidates p, p + g, p + 2g, ... (the order of the
    // depth-1 row and every deeper depth are as without the stride), so g
    // lanes with phases 0 .. g-1 together search exactly the task's tree.
    template <class C, class A>
    TFSM_LANE_FN void start_root(C &c, const A &a, uint32_t u_, uint32_t ci, uint32_t stride_ = 1,
                                 uint32_t phase = 0) {
        u = u_;
        n = a.nv();
        adaptive_rt = a.adaptive();
        floor = 1;
        d = 1;
        stride = stride_;
        c.idx(u) = ci;
        if (adaptive()) {
            load_edges(a);
            vis = 1u << u;
            c.meta(0) = meta_pack(0, u, 0, 0);
            if (same_lab_of(a, u)) c.vid(u) = a.vid(u, ci);
            enter_adaptive(c, a, 1, kNone, kNone);
        } else {
            c.vid(0) = a.vid(u, ci);
            enter_reference(c, a, 1, kNone);
        }
        // positions are below 2^32 - 64, so the shifted start cannot wrap;
        // a start at or past the row's end is an exhausted share
        c.pos(1) = c.pos(1) + phase;
    }

    // A piece: the candidates [lo, hi) of row `prim` at depth floor_, below
    // the candidate indices prefix[j] of depths j < floor_ (pattern vertices
    // pord[j], adaptive order, j <= floor_).
    template <class C, class A>
    TFSM_LANE_FN void start_piece(C &c, const A &a, uint32_t u_, uint32_t floor_, const uint32_t *prefix,
                                  const uint8_t *pord, uint32_t prim, uint32_t lo, uint32_t hi) {
        u = u_;
        n = a.nv();
        adaptive_rt = a.adaptive();
        floor = floor_;
        d = floor_;
        stride = 1;
        if (adaptive()) {
            load_edges(a);
            vis = 0;
            for (uint32_t j = 0; j < floor_; ++j) {
                const uint32_t w = pord[j] & 15u;
                c.idx(w) = prefix[j];
                c.meta(j) = meta_pack(0, w, 0, 0);
                if (same_lab_of(a, w)) c.vid(w) = a.vid(w, prefix[j]);
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
        return adaptive() ? meta_vertex(c.meta(j)) : a.order(u, j);
    }

    // Shallowest depth >= floor with untried candidates (d + 1 if none),
    // whose untried candidates form the contiguous range [pos, end): an
    // interleaved floor (stride > 1) is skipped.
    template <class C>
    TFSM_LANE_FN uint32_t open_depth(C &c) const {
        uint32_t k = stride > 1 ? floor + 1 : floor;
        while (k <= d && c.pos(k) >= c.end(k)) ++k;
        return k;
    }

    // One DFS step: test one candidate (and descend when it passes) or
    // backtrack one depth.
    template <class C, class A>
    TFSM_LANE_FN uint32_t step(C &c, const A &a) {
        last_depth = kNone;
        uint32_t pos = c.pos(d), end = c.end(d);
        // backtrack out of every exhausted depth, then test the next
        // candidate in the same step
        while (pos >= end) {
            if (adaptive()) vis &= ~(1u << meta_vertex(c.meta(d)));
            --d;
            if (d < floor) return kLaneExhausted;
            pos = c.pos(d);
            end = c.end(d);
        }
        const uint32_t adv = d == floor ? stride : 1u;
        c.pos(d) = pos + adv;
        const uint32_t m = c.meta(d);
        const uint32_t prim = meta_prim(m);
        uint32_t w, nb, g0, sl;
        if (adaptive()) {
            w = meta_vertex(m);
            nb = meta_secn(m) + 1;
            g0 = meta_secb(m);
            sl = same_lab_of(a, w);
        } else {
            w = a.order(u, d);
            nb = a.bn_n(u, d);
            g0 = a.sec_base(u, d);
            sl = a.same_before(u, d) | a.same_after(u, d);
        }
        const uint32_t cand = pos == pf_pos ? pf_val : a.pool(pos);
        // in flight beside this step's loads: the next primary entry and the
        // candidate's vertex id
        if (pos + adv < end) {
            pf_pos = pos + adv;
            pf_val = a.pool(pos + adv);
        }
        const uint32_t v = sl ? a.vid(w, cand) : 0u;
        for (uint32_t k = 0, j = 0; k < nb; ++k) {
            if (k == prim) continue;
            const uint32_t g = g0 + j++;
            bool member;
            // the adaptive order's rows always have cursor slots
            if (ORD == kOrderAdaptive || g < MV) {
                const uint32_t cur = c.cur(g), hi = c.cend(g);
                // a spent row: no later candidate of this depth is in it
                if (cur >= hi) {
                    c.pos(d) = end;
                    return kLaneRunning;
                }
                uint32_t cv = c.curv(g), at = cur;
                if (cv == kUnknown) at = lane_gallop(a, cur, hi, cand, cv);
                else if (cv < cand) at = lane_gallop(a, cur + 1, hi, cand, cv);
                c.cur(g) = at;
                c.curv(g) = cv;
                member = at < hi && cv == cand;
                if (!member) {
                    if (at >= hi) {
                        c.pos(d) = end;
                        return kLaneRunning;
                    }
                    // leapfrog: no primary entry below cv is in this row
                    if (adv == 1) {
                        const uint32_t nx = pos + 1;
                        if (nx < end && !(pf_pos == nx && pf_val >= cv)) {
                            uint32_t v2;
                            const uint32_t np = lane_gallop(a, nx, end, cv, v2);
                            c.pos(d) = np;
                            if (np < end) {
                                pf_pos = np;
                                pf_val = v2;
                            }
                        }
                    }
                    return kLaneRunning;
                }
            } else {
                // reference order only: more non-primary rows than cursor slots
                uint32_t lo, hi, cv;
                const uint32_t s = a.bn_slot(u, d, k);
                a.row(s, c.idx(a.slot_src(s)), lo, hi);
                const uint32_t at = lane_gallop(a, lo, hi, cand, cv);
                member = at < hi && cv == cand;
            }
            if (!member) return kLaneRunning;
        }
        if (sl) {
            if (adaptive()) {
                for (uint32_t q = sl & vis & ~(1u << w); q; q &= q - 1)
                    if (c.vid(lane_ctz(q)) == v) return kLaneRunning;
                c.vid(w) = v;
            } else {
                for (uint32_t q = a.same_before(u, d); q; q &= q - 1)
                    if (c.vid(lane_ctz(q)) == v) return kLaneRunning;
                c.vid(d) = v;
            }
        }
        c.idx(w) = cand;
        last_depth = d;
        last_cand = cand;
        if (d + 1 == n) return kLaneFound;
        ++d;
        if (adaptive()) enter_adaptive(c, a, d, kNone, kNone);
        else enter_reference(c, a, d, kNone);
        return kLaneRunning;
    }
};

}  // namespace tfsm_lane
