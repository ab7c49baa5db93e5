// Collision diagnostics (the reference's COL_DIAGNOSTICS build): the controller archive of every
// audited bath and the per-operator JSON-lines record with scheduler statistics per group and the
// event totals per outcome category.
//
// Formats follow save_col_controller (inc/swarm/_col_bound.cuh:242-299, schema 2) and the local
// collision log of evolve_local_collisions (inc/swarm/_col_chain.cuh:488-538, schema 1): the same
// keys in the same order, numbers written with 17 significant digits, so both builds' files can be
// compared with one reader.
#ifndef GDX_DIAGNOSTICS_H
#define GDX_DIAGNOSTICS_H

#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <limits>
#include <string>
#include <vector>

#include "controller.h"

namespace gdx {

constexpr int EVENT_CATEGORIES = GDX_EVENT_CATEGORIES;

struct bath_record
{
    int group_index = -1, operator_index = 0, bath_index = 0, merged_bins = 0, continuation_launches = 0;
    double duration = 0.0, limit_before = 1.0, limit_after = 1.0;
    bath_result result;
};

struct controller_summary
{
    int operator_count = 0, bath_count = 0, wave_count = 0, continuation_launches = 0;
    int activity_overshoots = 0, distribution_overshoots = 0, persistent_overshoots = 0;
    double minimum_duration = std::numeric_limits<double>::infinity(), maximum_duration = 0.0;
    double minimum_limit_scale = 1.0;
    double maximum_f = 0.0, maximum_e = 0.0, maximum_touched = 0.0, maximum_events = 0.0, maximum_g = 0.0;
    double maximum_g_upper = 0.0, maximum_d_bath = 0.0;
    std::vector<bath_record> baths;
};

// _record_col_bath (_col_bound.cuh:219-239)
inline void record_bath(controller_summary &s, const bath_record &rec)
{
    const bath_result &r = rec.result;
    s.bath_count++;
    s.continuation_launches += rec.continuation_launches;
    s.activity_overshoots += r.activity_overshoot ? 1 : 0;
    s.distribution_overshoots += r.distribution_overshoot ? 1 : 0;
    s.persistent_overshoots += r.persistent_overshoot ? 1 : 0;
    s.minimum_duration = std::min(s.minimum_duration, rec.duration);
    s.maximum_duration = std::max(s.maximum_duration, rec.duration);
    s.minimum_limit_scale = std::min(s.minimum_limit_scale, rec.limit_after);
    s.maximum_f = std::max(s.maximum_f, r.max_f);
    s.maximum_e = std::max(s.maximum_e, r.max_e);
    s.maximum_touched = std::max(s.maximum_touched, r.max_touched);
    s.maximum_events = std::max(s.maximum_events, r.max_events);
    s.maximum_g = std::max(s.maximum_g, r.max_g);
    s.maximum_g_upper = std::max(s.maximum_g_upper, r.max_g_upper);
    s.maximum_d_bath = std::max(s.maximum_d_bath, r.d_bath);
    s.baths.push_back(rec);
}

// collision_chain_<frame>.json, schema 2
inline bool save_controller(const std::string &path, const controller_summary &s)
{
    std::ofstream f(path);
    if (!f) return false;
    double minimum_duration = std::isfinite(s.minimum_duration) ? s.minimum_duration : 0.0;
    auto flag = [](bool b) { return b ? "true" : "false"; };
    f << std::setprecision(17);
    f << "{\n"
      << "  \"schema\": 2,\n"
      << "  \"operator_count\": " << s.operator_count << ",\n"
      << "  \"bath_count\": " << s.bath_count << ",\n"
      << "  \"wave_count\": " << s.wave_count << ",\n"
      << "  \"continuation_launches\": " << s.continuation_launches << ",\n"
      << "  \"activity_overshoots\": " << s.activity_overshoots << ",\n"
      << "  \"distribution_overshoots\": " << s.distribution_overshoots << ",\n"
      << "  \"persistent_overshoots\": " << s.persistent_overshoots << ",\n"
      << "  \"minimum_duration\": " << minimum_duration << ",\n"
      << "  \"maximum_duration\": " << s.maximum_duration << ",\n"
      << "  \"minimum_limit_scale\": " << s.minimum_limit_scale << ",\n"
      << "  \"maximum_f\": " << s.maximum_f << ",\n"
      << "  \"maximum_e\": " << s.maximum_e << ",\n"
      << "  \"maximum_touched\": " << s.maximum_touched << ",\n"
      << "  \"maximum_events\": " << s.maximum_events << ",\n"
      << "  \"maximum_g\": " << s.maximum_g << ",\n"
      << "  \"maximum_g_upper\": " << s.maximum_g_upper << ",\n"
      << "  \"maximum_d_bath\": " << s.maximum_d_bath << ",\n"
      << "  \"baths\": [\n";
    for (size_t i = 0; i < s.baths.size(); ++i)
    {
        const bath_record &b = s.baths[i];
        const bath_result &r = b.result;
        f << "    {\"operator\": " << b.operator_index << ", \"group\": " << b.group_index
          << ", \"bath\": " << b.bath_index << ", \"merged_bins\": " << b.merged_bins
          << ", \"continuation_launches\": " << b.continuation_launches << ", \"duration\": " << b.duration
          << ", \"limit_before\": " << b.limit_before << ", \"limit_after\": " << b.limit_after
          << ", \"max_f\": " << r.max_f << ", \"max_e\": " << r.max_e << ", \"max_touched\": " << r.max_touched
          << ", \"max_events\": " << r.max_events << ", \"max_g\": " << r.max_g
          << ", \"max_g_upper\": " << r.max_g_upper << ", \"d_bath\": " << r.d_bath
          << ", \"activity_overshoot\": " << flag(r.activity_overshoot)
          << ", \"distribution_overshoot\": " << flag(r.distribution_overshoot)
          << ", \"persistent_overshoot\": " << flag(r.persistent_overshoot) << "}"
          << (i + 1 < s.baths.size() ? "," : "") << "\n";
    }
    f << "  ]\n}\n";
    return static_cast<bool>(f);
}

// scheduler statistics of one group over one operator call (_col_chain.cuh:26-31, 262-449)
struct group_stats
{
    uint64_t updates = 0, coarsened = 0, refined = 0, constrained = 0;
    int overshoots = 0, persistent = 0, initial_level = 0, binding = 0;
    double max_age = 0, max_growth = 0, max_activity = 0, max_requested_ratio = 0;
    double min_request = 0, max_request = 0, initial_request = 0;
};

struct operator_record
{
    double clock_sim = 0, duration = 0;
    uint64_t waves = 0, owner_updates = 0, chain_blocks = 0, chain_launches = 0, finest_ticks = 0;
    double chain_seconds = 0, audit_seconds = 0, scheduler_wall_seconds = 0;
    std::vector<int> owners, level;
    std::vector<double> dt, final_request;
    std::vector<group_stats> groups;
    unsigned long long event_count[EVENT_CATEGORIES] = {};
    double event_log_mass[EVENT_CATEGORIES] = {};
};

// one line of collision_local_<stamp>.jsonl, schema 1
inline void write_operator_line(std::ostream &o, uint64_t operator_id, const operator_record &r)
{
    o << "{\"schema\":1,\"method\":\"local_cached_collision\",\"operator\":" << operator_id
      << ",\"clock_sim\":" << r.clock_sim << ",\"duration\":" << r.duration << ",\"waves\":" << r.waves
      << ",\"owner_updates\":" << r.owner_updates << ",\"chain_blocks\":" << r.chain_blocks
      << ",\"chain_launches\":" << r.chain_launches << ",\"chain_seconds\":" << r.chain_seconds
      << ",\"audit_seconds\":" << r.audit_seconds << ",\"scheduler_wall_seconds\":" << r.scheduler_wall_seconds
      << ",\"finest_ticks\":" << r.finest_ticks << ",\"groups\":[";
    for (size_t c = 0; c < r.groups.size(); ++c)
    {
        const group_stats &s = r.groups[c];
        if (c) o << ',';
        o << "{\"id\":" << c << ",\"owners\":" << r.owners[c] << ",\"level\":" << r.level[c]
          << ",\"dt\":" << r.dt[c] << ",\"initial_level\":" << s.initial_level
          << ",\"coarsened_updates\":" << s.coarsened << ",\"refined_updates\":" << s.refined
          << ",\"neighbor_constrained_updates\":" << s.constrained
          << ",\"minimum_requested_dt\":" << s.min_request << ",\"maximum_requested_dt\":" << s.max_request
          << ",\"final_requested_dt\":" << r.final_request[c] << ",\"initial_requested_dt\":" << s.initial_request
          << ",\"initial_binding_constraint\":" << s.binding << ",\"updates\":" << s.updates
          << ",\"overshoots\":" << s.overshoots << ",\"persistent_overshoots\":" << s.persistent
          << ",\"max_snapshot_age_at_start\":" << s.max_age << ",\"max_requested_dt_ratio\":" << s.max_requested_ratio
          << ",\"max_growth\":" << s.max_growth << ",\"max_predicted_activity\":" << s.max_activity << '}';
    }
    o << "],\"event_counts\":[";
    for (int k = 0; k < EVENT_CATEGORIES; ++k) o << (k ? "," : "") << r.event_count[k];
    o << "],\"event_log_mass_sums\":[";
    for (int k = 0; k < EVENT_CATEGORIES; ++k) o << (k ? "," : "") << r.event_log_mass[k];
    o << "]}\n";
}

// diagnostics state of one handle: the controller archive since the last reset and the log
struct diag_state
{
    controller_summary summary;
    std::ofstream log;
    uint64_t operator_id = 0;
    double clock_sim = 0.0;
    gdx_event_work totals{};                // event totals of the last operator call
};

}  // namespace gdx

#endif  // GDX_DIAGNOSTICS_H
