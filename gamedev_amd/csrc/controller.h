// Host bath controller and power-of-two scheduler of the collision operator.
//
// Follows the reference's host logic: change_limit, _build_col_binmap, _choose_col_bath and
// _finish_col_bath (inc/swarm/_col_bound.cuh:22-216) and local_schedule (inc/swarm/_col_sched.cuh:17-183).
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

// merge adjacent sparse size bins inside every spatial group until each holds COL_BIN_MIN owners
inline int build_binmap(const std::vector<int> &raw_count, std::vector<int> &map)
{
    const int groups = int(raw_count.size()) / COL_BIN_S;
    map.assign(raw_count.size(), -1);
    int merged = 0;
    for (int gidx = 0; gidx < groups; ++gidx)
    {
        const int *cnt = raw_count.data() + gidx*COL_BIN_S;
        int *out = map.data() + gidx*COL_BIN_S;
        int last = -1;
        for (int b = 0; b < COL_BIN_S; ++b) if (cnt[b] > 0) last = b;
        if (last < 0)
        {
            int id = merged++;
            for (int b = 0; b < COL_BIN_S; ++b) out[b] = id;
            continue;
        }
        int b0 = 0, prev = -1;
        while (b0 <= last)
        {
            int b1 = b0, sum = 0;
            while (b1 <= last && sum < COL_BIN_MIN) sum += cnt[b1++];
            bool tail = b1 > last && sum < COL_BIN_MIN && prev >= 0;
            int id = tail ? prev : merged++;
            for (int b = b0; b < b1; ++b) out[b] = id;
            prev = id;
            b0 = b1;
        }
        if (last + 1 < COL_BIN_S)
        {
            int id = merged++;
            for (int b = last + 1; b < COL_BIN_S; ++b) out[b] = id;
        }
    }
    return merged;
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
    const bool issue = r.activity_overshoot || r.distribution_overshoot;
    const bool at_floor = state.limit_scale <= 0.25*(1.0 + 8.0*std::numeric_limits<double>::epsilon());
    if (issue)
    {
        state.activity_streak++;
        state.quiet_streak = 0;
        state.floor_streak = at_floor ? state.floor_streak + 1 : 0;
    }
    else
    {
        state.activity_streak = 0;
        state.quiet_streak++;
        state.floor_streak = 0;
    }
    if (r.distribution_overshoot || state.activity_streak >= 2)
    {
        state.limit_scale = std::max(0.25, 0.5*state.limit_scale);
        state.activity_streak = 0;
    }
    else if (state.quiet_streak >= 3)
    {
        state.limit_scale = std::min(1.0, 1.25*state.limit_scale);
        state.quiet_streak = 0;
    }
    r.persistent_overshoot = state.floor_streak >= 2;
    return r;
}

// groups advance on power-of-two fractions of the horizon, on a lattice of 2^52 ticks; linked
// groups keep step ratios of at most 16
struct schedule
{
    static constexpr int FINEST = 52;
    double horizon;
    std::vector<std::pair<int, int>> links;
    std::vector<int> level;
    std::vector<uint64_t> step, next;
    uint64_t end;

    static int level_for(double h, double requested, const char *what)
    {
        if (!(requested > 0) || !std::isfinite(requested)) throw status_error(10, what);
        int lv = 0;
        double dt = h;
        while (dt > requested)
        {
            if (++lv > FINEST) throw status_error(10, "local collision timestep exceeds time resolution");
            dt *= 0.5;
        }
        return lv;
    }

    schedule(double h, const std::vector<double> &requested, const std::vector<uint32_t> &edges)
        : horizon(h), level(requested.size(), 0), step(requested.size()), next(requested.size(), 0)
    {
        if (!(h > 0) || !std::isfinite(h)) throw status_error(10, "invalid local collision horizon");
        const int n = int(level.size());
        const int words = (n + 31) / 32;
        auto edge = [&](int c, int d) { return (edges[size_t(c)*words + d / 32] >> (d % 32)) & 1u; };
        for (int c = 0; c < n; ++c) level[c] = level_for(h, requested[c], "invalid local collision timestep");
        for (int c = 0; c < n; ++c)
            for (int d = c + 1; d < n; ++d)
                if (edge(c, d) || edge(d, c)) links.emplace_back(c, d);
        bool changed;
        do
        {
            changed = false;
            for (int c = 0; c < n; ++c)
                for (int d = 0; d < n; ++d)
                {
                    if (!edge(c, d)) continue;
                    if (level[c] < level[d] - 4) { level[c] = level[d] - 4; changed = true; }
                    if (level[d] < level[c] - 4) { level[d] = level[c] - 4; changed = true; }
                }
        } while (changed);
        end = uint64_t(1) << FINEST;
        for (int c = 0; c < n; ++c) step[c] = uint64_t(1) << (FINEST - level[c]);
    }
    uint64_t time() const { return *std::min_element(next.begin(), next.end()); }
    double seconds(uint64_t tick) const { return horizon*(double(tick) / double(end)); }
    std::vector<int> due(uint64_t tick) const
    {
        std::vector<int> out;
        for (int c = 0; c < int(next.size()); ++c) if (next[c] == tick) out.push_back(c);
        return out;
    }
    void adapt(uint64_t tick, const std::vector<int> &groups, const std::vector<double> &requested,
               const std::vector<bool> &passed)
    {
        const int n = int(level.size());
        std::vector<bool> active(n, false);
        for (int c : groups) active[c] = true;
        std::vector<int> cap(n, FINEST), cand = level;
        for (int c = 0; c < n; ++c) if (!active[c]) cap[c] = level[c];
        bool changed;
        do
        {
            changed = false;
            for (const auto &l : links)
            {
                int c = l.first, d = l.second;
                if (active[c] && cap[c] > cap[d] + 4) { cap[c] = cap[d] + 4; changed = true; }
                if (active[d] && cap[d] > cap[c] + 4) { cap[d] = cap[c] + 4; changed = true; }
            }
        } while (changed);
        for (int c : groups)
        {
            if (!(requested[c] > 0) || !std::isfinite(requested[c])) throw status_error(10, "invalid adaptive collision timestep");
            int want = 0;
            double h = horizon;
            while (h > requested[c])
            {
                if (++want > FINEST) throw status_error(10, "adaptive collision time resolution exceeded");
                h *= 0.5;
            }
            cand[c] = passed[c] ? want : std::max(want, level[c]);
            while (tick % (uint64_t(1) << (FINEST - cand[c]))) ++cand[c];
            cand[c] = std::min(cand[c], cap[c]);
        }
        do
        {
            changed = false;
            for (const auto &l : links)
            {
                int c = l.first, d = l.second;
                if (active[c] && cand[c] < cand[d] - 4) { cand[c] = cand[d] - 4; changed = true; }
                if (active[d] && cand[d] < cand[c] - 4) { cand[d] = cand[c] - 4; changed = true; }
            }
        } while (changed);
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
