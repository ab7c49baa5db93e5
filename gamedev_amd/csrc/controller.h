// Host bath controller and power-of-two scheduler of the collision operator.
//
// The controller formulas (change_limit, choose_bath, the Bernstein envelopes of finish_bath) keep the
// floating-point expression order of the reference's _col_bound.cuh:22-216, because bath durations and
// audit decisions must coincide with it. The integer parts (merged size bins, the safety-factor streaks,
// the scheduler of _col_sched.cuh:17-183) are written from the rules the reference documents
// (doc/guide_swarm.md 8.7.1, 8.7.3, 8.7.5); their results are fixed by those rules, so they coincide too
// (the host tests compare them with the oracle's restatement on hand-made and random inputs).
// Failures raise status_error with the library status of the reference's std::runtime_error.
#ifndef GDX_CONTROLLER_H
#define GDX_CONTROLLER_H

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

#include "gdx_internal.h"

namespace gdx {

struct rate_bin { double mass = 0, weighted_rate = 0, weighted_change = 0, weighted_second = 0; int owner_count = 0, invalid_count = 0; };
struct audit_bin
{
    double mass = 0, predicted_f = 0, predicted_e = 0, predicted_var_f = 0, predicted_var_e = 0, predicted_g = 0,
        predicted_var_g = 0, maximum_weight = 0, maximum_g_jump = 0, touched = 0, event_weight = 0, growth = 0,
        end_mass = 0;
    int owner_count = 0, invalid_count = 0;
};
struct bath_result
{
    double max_f = 0, max_e = 0, max_touched = 0, max_events = 0, max_g = 0, max_g_upper = 0, d_bath = 0;
    bool activity_overshoot = false, distribution_overshoot = false, persistent_overshoot = false;
};
using bath_state = context::bath_state;

// A: mass-weighted mean |d ln s| rate, B: second-moment rate; bound h A and sqrt(h B) by epsilon
inline std::pair<double, int> change_limit(double A, double B, double epsilon, double horizon)
{
    if (!std::isfinite(A) || A < 0 || !std::isfinite(B) || B < 0 || !std::isfinite(epsilon) || !(epsilon > 0)
        || !std::isfinite(horizon) || !(horizon > 0))
        throw status_error(8, "invalid change-based refresh inputs");
    std::pair<double, int> r{horizon, 0};
    if (A > 0 && epsilon / A < r.first) r = {epsilon / A, 1};
    if (B > 0 && epsilon*epsilon / B < r.first) r = {epsilon*epsilon / B, 2};
    if (!(r.first > 0)) throw status_error(8, "change-based timestep underflow");
    return r;
}

// Merged size bins (the rule of the reference's controller, doc/guide_swarm.md 8.7.1): inside each
// spatial group, adjacent raw size bins are merged from small to large sizes, a merged bin being closed
// as soon as it holds COL_BIN_MIN owners; a tail that stays below COL_BIN_MIN joins the merged bin before
// it (if there is one); the bins above the largest occupied one form one further, empty merged bin, and
// a group with no owners is one empty merged bin. Merged bins are numbered consecutively over groups.
// Returns the number of merged bins; map[r] is the merged bin of raw bin r.
inline int build_binmap(const std::vector<int> &raw_count, std::vector<int> &map)
{
    const int groups = int(raw_count.size()) / COL_BIN_S;
    map.assign(raw_count.size(), -1);
    int total = 0;
    int local[COL_BIN_S];
    for (int gidx = 0; gidx < groups; ++gidx)
    {
        const int *cnt = raw_count.data() + size_t(gidx)*COL_BIN_S;
        int top = COL_BIN_S - 1;
        while (top >= 0 && cnt[top] <= 0) --top;
        // group-local merged indices of the occupied range [0, top]
        int closed = 0, fill = 0;
        for (int b = 0; b <= top; ++b)
        {
            local[b] = closed;
            fill += cnt[b];
            if (fill >= COL_BIN_MIN && b < top) { ++closed; fill = 0; }
        }
        int used = top >= 0 ? closed + 1 : 0;
        if (top >= 0 && fill < COL_BIN_MIN && closed > 0)
        {
            for (int b = 0; b <= top; ++b) if (local[b] == closed) local[b] = closed - 1;
            used = closed;
        }
        for (int b = top + 1; b < COL_BIN_S; ++b) local[b] = used;   // the empty bins above
        if (top + 1 < COL_BIN_S) ++used;
        int *out = map.data() + size_t(gidx)*COL_BIN_S;
        for (int b = 0; b < COL_BIN_S; ++b) out[b] = total + local[b];
        total += used;
    }
    return total;
}

// binding: 0 horizon or COL_BATH_MAX, 1 mean-change bound, 2 variance bound (the last bin that shortened it)
inline double choose_bath(const rate_bin *bins, int count, double remaining, double limit_scale,
                          int *binding = nullptr)
{
    double duration = std::min(remaining, COL_BATH_MAX);
    const double tolerance = COL_BATH_EPS*limit_scale;
    int reason = 0;
    for (int b = 0; b < count; ++b)
    {
        const rate_bin &v = bins[b];
        if (v.invalid_count) throw status_error(8, "invalid change-based rate moments");
        if (!(v.mass > 0)) continue;
        auto bound = change_limit(v.weighted_change / v.mass, v.weighted_second / v.mass, tolerance, duration);
        if (bound.first < duration)
        {
            duration = bound.first;
            reason = bound.second;
        }
    }
    if (binding) *binding = reason;
    return duration;
}

// The group's safety factor after an audited bath (doc/guide_swarm.md 8.7.5): halved (floor 0.25) after a
// distribution overshoot or the second overshooting bath in a row of either kind; raised by 1.25 (to at
// most 1) after the third quiet bath in a row; each change restarts the streak that caused it. Returns
// whether this bath is the second overshooting bath in a row at the floor (a persistent overshoot).
inline bool update_safety(bath_state &state, bool distribution, bool activity)
{
    const bool overshoot = distribution || activity;
    const bool at_floor = state.limit_scale <= 0.25*(1.0 + 8.0*std::numeric_limits<double>::epsilon());
    state.floor_streak = (overshoot && at_floor) ? state.floor_streak + 1 : 0;
    state.activity_streak = overshoot ? state.activity_streak + 1 : 0;
    state.quiet_streak = overshoot ? 0 : state.quiet_streak + 1;
    if (distribution || state.activity_streak == 2)
    {
        state.limit_scale = std::max(0.25, 0.5*state.limit_scale);
        state.activity_streak = 0;
    }
    else if (state.quiet_streak == 3)
    {
        state.limit_scale = std::min(1.0, 1.25*state.limit_scale);
        state.quiet_streak = 0;
    }
    return state.floor_streak >= 2;
}

// realized activity and size change against Bernstein envelopes; adapt the group's safety factor
inline bath_result finish_bath(const audit_bin *bins, int count, bath_state &state)
{
    bath_result r;
    int active = 0;
    double total = 0.0;
    for (int b = 0; b < count; ++b)
    {
        if (bins[b].invalid_count != 0) throw status_error(9, "collision bath controller returned an invalid audit state");
        if (bins[b].mass > 0.0) active++;
        total += bins[b].mass;
    }
    if (!(total > 0.0)) throw status_error(9, "collision bath controller found no active represented mass");
    const double L = std::log(2.0*double(active) / COL_BATH_ALPHA);
    double mass_difference = 0.0;
    for (int b = 0; b < count; ++b)
    {
        const audit_bin &v = bins[b];
        mass_difference += std::abs(v.end_mass - v.mass);
        if (!(v.mass > 0.0)) continue;
        const double msq = v.mass*v.mass;
        const double pf = v.predicted_f / v.mass, pe = v.predicted_e / v.mass, pg = v.predicted_g / v.mass;
        const double touched = v.touched / v.mass, events = v.event_weight / v.mass, growth = v.growth / v.mass;
        const double wmax = v.maximum_weight / v.mass, jmax = v.maximum_g_jump / v.mass;
        const double upper_f = std::min(1.0, pf + std::sqrt(2.0*v.predicted_var_f / msq*L) + wmax*L / 3.0);
        const double upper_e = pe + std::sqrt(2.0*v.predicted_var_e / msq*L) + wmax*L / 3.0;
        const double upper_g = pg + std::sqrt(2.0*v.predicted_var_g / msq*L) + jmax*L / 3.0;
        r.max_f = std::max(r.max_f, pf);
        r.max_e = std::max(r.max_e, pe);
        r.max_touched = std::max(r.max_touched, touched);
        r.max_events = std::max(r.max_events, events);
        r.max_g = std::max(r.max_g, growth);
        r.max_g_upper = std::max(r.max_g_upper, upper_g);
        r.activity_overshoot = r.activity_overshoot || std::min(touched, 1.0) > upper_f || events > upper_e;
        r.distribution_overshoot = r.distribution_overshoot || growth > std::max(COL_BATH_EPS, upper_g);
    }
    r.d_bath = mass_difference / (2.0*total);
    r.distribution_overshoot = r.distribution_overshoot || r.d_bath > COL_BATH_EPS;
    r.persistent_overshoot = update_safety(state, r.distribution_overshoot, r.activity_overshoot);
    return r;
}

// Power-of-two scheduler (the rules of doc/guide_swarm.md 8.7.3). Group c advances by step h / 2^level[c]
// on a lattice of 2^52 ticks per horizon h; its level is the coarsest whose step does not exceed the
// requested duration. Linked groups keep levels within 4 of each other (16:1), the coarser one refined.
// At a refresh a due group may coarsen only if its audit passed, is refined until its step aligns with
// the tick, is capped so that no pending endpoint of a group that is not due must move, and due groups
// are then refined again until every link holds. Endpoints: next[c], advanced by step[c].
struct schedule
{
    static constexpr int FINEST = 52;
    static constexpr int LINK_LEVELS = 4;
    static constexpr int FAR = 1 << 20;       // hop count of unreachable pairs
    static constexpr int MAX_GROUPS = 64;     // one bit per group in the link masks
    double horizon;
    int n;
    std::vector<uint64_t> adj;                // linked groups (an edge either way), one bit per group
    std::vector<int> level;
    std::vector<uint64_t> step, next;
    uint64_t end;

    // the coarsest level whose step does not exceed `requested`
    int level_for(double requested, const char *invalid, const char *too_fine) const
    {
        if (!(requested > 0) || !std::isfinite(requested)) throw status_error(10, invalid);
        for (int lv = 0; lv <= FINEST; ++lv)
            if (std::ldexp(horizon, -lv) <= requested) return lv;
        throw status_error(10, too_fine);
    }

    // hop counts d[0..n) from `src` along links whose interior groups are in the mask `relay` (FAR if
    // unreachable): a breadth-first search one hop level at a time over the link masks (the host's
    // share of every refresh wave, so it allocates nothing)
    void hops_from(int src, uint64_t relay, int *d) const
    {
        for (int i = 0; i < n; ++i) d[i] = FAR;
        d[src] = 0;
        uint64_t seen = uint64_t(1) << src, frontier = seen;
        const uint64_t pass = relay | seen;
        for (int depth = 1; frontier; ++depth)
        {
            uint64_t reach = 0;
            for (uint64_t f = frontier & pass; f; f &= f - 1) reach |= adj[size_t(__builtin_ctzll(f))];
            reach &= ~seen;
            seen |= reach;
            for (uint64_t m = reach; m; m &= m - 1) d[__builtin_ctzll(m)] = depth;
            frontier = reach;
        }
    }

    // the finest level any group forces on c through the 16:1 links: max over e of lv[e] - 4 hops(c, e),
    // relayed only through groups in the mask `mutable_set` (the others keep lv and pass nothing on)
    std::vector<int> raise_for_links(const std::vector<int> &lv, uint64_t mutable_set) const
    {
        std::vector<int> out = lv;
        int d[MAX_GROUPS];
        for (int c = 0; c < n; ++c)
        {
            if (!((mutable_set >> c) & 1u)) continue;
            hops_from(c, mutable_set, d);
            for (int e = 0; e < n; ++e)
                if (d[e] != FAR) out[c] = std::max(out[c], lv[e] - LINK_LEVELS*d[e]);
        }
        return out;
    }

    schedule(double h, const std::vector<double> &requested, const std::vector<uint32_t> &edges)
        : horizon(h), n(int(requested.size())), adj(requested.size(), 0), level(requested.size(), 0),
          step(requested.size()), next(requested.size(), 0)
    {
        if (!(h > 0) || !std::isfinite(h)) throw status_error(10, "invalid local collision horizon");
        if (n > MAX_GROUPS) throw status_error(GDX_COL_ERR_ARG, "the scheduler takes at most 64 groups");
        const int words = (n + 31) / 32;
        for (int c = 0; c < n; ++c)
            for (int d = 0; d < n; ++d)
                if (c != d && (((edges[size_t(c)*words + d / 32] >> (d % 32)) & 1u)
                               || ((edges[size_t(d)*words + c / 32] >> (c % 32)) & 1u)))
                    adj[size_t(c)] |= uint64_t(1) << d;
        for (int c = 0; c < n; ++c)
            level[c] = level_for(requested[c], "invalid local collision timestep",
                                 "local collision timestep exceeds time resolution");
        level = raise_for_links(level, n == 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1);
        end = uint64_t(1) << FINEST;
        for (int c = 0; c < n; ++c) step[c] = uint64_t(1) << (FINEST - level[c]);
    }
    uint64_t time() const { return *std::min_element(next.begin(), next.end()); }
    double seconds(uint64_t tick) const { return horizon*(double(tick) / double(end)); }
    std::vector<int> due(uint64_t tick) const
    {
        std::vector<int> out;
        for (int c = 0; c < n; ++c) if (next[c] == tick) out.push_back(c);
        return out;
    }
    void adapt(uint64_t tick, const std::vector<int> &groups, const std::vector<double> &requested,
               const std::vector<bool> &passed)
    {
        uint64_t is_due = 0;
        for (int c : groups) is_due |= uint64_t(1) << c;
        // finest level that still leaves every pending endpoint of a group that is not due in place:
        // min over such groups e of level[e] + 4 hops(c, e), through due groups only
        std::vector<int> cap(n, FINEST);
        int d[MAX_GROUPS];
        for (int c : groups)
        {
            hops_from(c, is_due, d);
            for (int e = 0; e < n; ++e)
                if (!((is_due >> e) & 1u) && d[e] != FAR) cap[c] = std::min(cap[c], level[e] + LINK_LEVELS*d[e]);
        }
        std::vector<int> cand = level;
        for (int c : groups)
        {
            const int want = level_for(requested[c], "invalid adaptive collision timestep",
                                       "adaptive collision time resolution exceeded");
            int lv = passed[c] ? want : std::max(want, level[c]);
            // a step aligned with the tick: 2^(52 - lv) divides it
            if (tick != 0) lv = std::max(lv, FINEST - __builtin_ctzll(tick));
            cand[c] = std::min(lv, cap[c]);
        }
        cand = raise_for_links(cand, is_due);
        for (int c : groups)
        {
            level[c] = cand[c];
            step[c] = uint64_t(1) << (FINEST - level[c]);
        }
    }
    void advance(const std::vector<int> &groups) { for (int c : groups) next[c] += step[c]; }
};

}  // namespace gdx

#endif  // GDX_CONTROLLER_H
