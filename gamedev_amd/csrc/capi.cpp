// C ABI of the collision library and the host driver of one operator call.
//
// gdx_col_operator follows evolve_local_collisions (inc/swarm/_col_chain.cuh:108-539): publish and
// rate every group at tick zero, request per-group bath durations, schedule them on the power-of-two
// lattice, then advance the earliest due groups wave by wave with a rating pass (tick > 0), the fused
// screen + chain + audit kernel, and the host audit. Two host synchronisations per wave.
#include "../../include/gdx_col.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "controller.h"
#include "diagnostics.h"
#include "gdx_internal.h"

struct gdx_col_ctx
{
    gdx::context c;
    gdx::diag_state diag;
    gdx::run_buffers rb;
};

struct gdx_morton_ctx
{
    gdx::morton_index ix;
    gdx::sort_scratch scratch;
    gdx::dbuf<unsigned int> overflow;
};

namespace gdx {

int state_flag(context &c);

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(context *c, F &&f)
{
    try
    {
        f();
        return GDX_COL_OK;
    }
    catch (const status_error &e)
    {
        (c ? c->last_error : g_last_error) = e.what();
        return e.code;
    }
    catch (const std::bad_alloc &)
    {
        (c ? c->last_error : g_last_error) = "host allocation failure";
        return GDX_COL_ERR_ALLOC;
    }
    catch (const std::exception &e)
    {
        (c ? c->last_error : g_last_error) = e.what();
        return GDX_COL_ERR_HIP;
    }
}

void require(bool ok, int code, const char *what)
{
    if (!ok) throw status_error(code, what);
}

// the group-table entry of group g of swarm `run` of a run whose table starts that swarm at tbase
gdx_gref make_gref(context &c, int g, run_buffers &rb, int run, int tbase)
{
    gdx_gref r{};
    r.st = c.st.p;
    r.res = c.res.p;
    r.nbr = c.nbr.p;
    r.measure = c.measure.p;
    r.env = c.env.p;
    r.hgs = c.hgs.p;
    r.cached = c.cached.p;
    r.rng = c.rng.p;
    r.slot_pidx = c.slot_pidx.p;
    r.queue = c.queue.p;
    r.qsum = c.qsum.p;
    r.qevents = c.qevents.p;
    r.evw = c.diag ? c.evw.p : nullptr;
    r.raw_rate = rb.raw_rate.p + size_t(tbase + g)*COL_BIN_S;
    r.raw_audit = rb.raw_audit.p + size_t(tbase + g)*COL_BIN_S;
    r.qcount = rb.qcount.p + tbase + g;
    r.hcount = rb.hcount.p + tbase + g;
    r.ccount = rb.ccount.p + tbase + g;
    r.cqueue = c.cqueue.p;
    r.cstart = c.cstart.p;
    r.scr_min = c.scr_min.p + g;
    r.scr_max = c.scr_max.p + g;
    r.lower = c.lower.p + g;
    r.upper = c.upper.p + g;
    r.counters = rb.counters.p + size_t(run)*GDX_CTRS;
    r.wave_ctr = rb.wave_ctr.p + size_t(run)*GDX_WAVE_CTRS;
    r.err = rb.err.p + run;
    r.lambda_0 = c.n / static_cast<double>(N_K) / c.total_mass;
    r.seed = c.seed;
    r.mlp = c.mlp;
    r.safety = c.safety*c.safety_scale;
    r.thin_stride = c.thin_stride;
    r.lam = c.lam.p;
    r.lam_add = c.lam_add.p;
    r.dirty = c.dirty.p;
    r.exact_walk = c.exact_walk;
    const bool track = c.model && c.rev_valid;
    r.stale = track ? c.stale.p : nullptr;
    r.rev_beg = track ? c.rev_beg.p : nullptr;
    r.rev_end = track ? c.rev_end.p : nullptr;
    r.rev = track ? c.rev.p : nullptr;
    r.bound_reuse = c.bound_reuse;
    r.res32 = c.res32.p;
    r.qstart = c.qstart.p;
    return r;
}

struct phase_entry
{
    int ref, beg, cnt;
    double dt;
    int sub;              // learned swarm: exact-subsample phase of this bath; -1 plain
};

// due entries of one phase in device memory; launch shapes: 1024 and 256 owners per block, the groups
// rated one lane per owner (256 per block), the learned groups' owners (MAJ_SLOTS per block) and the
// groups rated by lane groups (RATE_SUB_OWNERS per block): learned groups' exact subsamples, and plain
// groups when the phase is narrow (at most narrow_owners due owners)
struct phase_wave
{
    gdx_wave w1024, w256, wsub, wmaj, wlsub;
    int blocks1024 = 0, blocks256 = 0, blocks_sub = 0, blocks_maj = 0, blocks_lsub = 0;
    bool narrow = false;
};

// geometry epochs timed for each candidate value of one top-K launch parameter (`field`, the tile or the
// block shape), the fastest kept: the epoch's results do not depend on either, only its time does
void autotune_epoch(context &c, int &field, const int *candidates, int count, int repeats, int *best, double *ms)
{
    require(c.user.bound, GDX_COL_ERR_STATE, "no particle state is bound");
    if (c.state_loaded) store_state(c);
    hip_check(hipStreamSynchronize(c.stream), "autotune");
    int pick = field;
    double fastest = 0.0;
    for (int i = 0; i < count; ++i)
    {
        field = candidates[i];
        geometry_epoch(c);
        double t = 0.0;
        for (int r = 0; r < repeats; ++r)
        {
            hip_check(hipStreamSynchronize(c.stream), "autotune");
            auto t0 = std::chrono::steady_clock::now();
            geometry_epoch(c);
            hip_check(hipStreamSynchronize(c.stream), "autotune");
            double v = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            t = r ? std::min(t, v) : v;
        }
        if (ms) ms[i] = t;
        if (i == 0 || t < fastest) { fastest = t; pick = candidates[i]; }
    }
    field = pick;
    geometry_epoch(c);
    hip_check(hipStreamSynchronize(c.stream), "autotune");
    if (best) *best = pick;
}

// the host's wait for the end of a phase: polling the stream rather than sleeping in the runtime's
// wait, since the next phase's launches depend on what the host reads back (every microsecond of
// wake-up is a microsecond of idle GPU, several times per operator call)
void wait_stream(hipStream_t st, const char *what)
{
    hipError_t e;
    while ((e = hipStreamQuery(st)) == hipErrorNotReady) {}
    hip_check(e, what);
}

// fills the phase's due list and stages it on the device together with clearing the ranges `clear`
phase_wave upload_phase(run_buffers &rb, const std::vector<phase_entry> &e, hipStream_t st, const zero_ranges &clear,
                        int narrow_owners)
{
    const int n = int(e.size());
    rb.reserve(n);
    int *ref = rb.h_int, *beg = ref + n, *cnt = beg + n, *sub = cnt + n, *b1024 = sub + n, *b256 = b1024 + n + 1;
    int *bsub = b256 + n + 1, *bmaj = bsub + n + 1, *blsub = bmaj + n + 1;
    int total = 0, x = 0, y = 0, z = 0, u = 0, v = 0, learned = 0;
    long long due = 0;
    for (const phase_entry &p : e) due += p.cnt;
    const bool narrow = due <= narrow_owners;
    for (int i = 0; i < n; ++i)
    {
        ref[i] = e[i].ref;
        beg[i] = e[i].beg;
        cnt[i] = e[i].cnt;
        sub[i] = e[i].sub;
        rb.h_dt[i] = e[i].dt;
        b1024[i] = x;
        b256[i] = y;
        bsub[i] = z;
        bmaj[i] = u;
        blsub[i] = v;
        x += (e[i].cnt + 1023) / 1024;
        y += (e[i].cnt + 255) / 256;
        // plain groups: every owner, 256 per block (RATE_SUB_OWNERS per block in a narrow phase); learned
        // groups: the exact subsample (runs sub, sub + stride, ... of THIN_RUN slots, as exact_owner on the
        // device), RATE_SUB_OWNERS per block
        if (e[i].sub < 0 && !narrow) z += (e[i].cnt + 255) / 256;
        else if (e[i].sub < 0) v += (e[i].cnt + RATE_SUB_OWNERS - 1) / RATE_SUB_OWNERS;
        else v += (thin_rated(e[i].cnt, e[i].sub, rb.h_table[size_t(e[i].ref)].thin_stride) + RATE_SUB_OWNERS - 1)
                  / RATE_SUB_OWNERS;
        u += e[i].sub < 0 ? 0 : (e[i].cnt + MAJ_SLOTS - 1) / MAJ_SLOTS;
        learned |= e[i].sub >= 0;
        total += e[i].cnt;
    }
    b1024[n] = x;
    b256[n] = y;
    bsub[n] = z;
    bmaj[n] = u;
    blsub[n] = v;
    // ints and bath ends in one launch
    launch_stage(st, clear, rb.due_d.p, rb.due_h.p, rb.due_ints_bytes + sizeof(double)*std::max(1, n));
    phase_wave pw;
    const int *d = rb.d_int.p;
    auto wave = [&](const int *bstart) {
        return gdx_wave{n, total, rb.table.p, d, d + n, d + 2*n, bstart, rb.d_dt.p, rb.qprefix.p, d + 3*n, learned};
    };
    pw.w1024 = wave(d + (b1024 - ref));
    pw.w256 = wave(d + (b256 - ref));
    pw.wsub = wave(d + (bsub - ref));
    pw.wmaj = wave(d + (bmaj - ref));
    pw.wlsub = wave(d + (blsub - ref));
    pw.blocks1024 = x;
    pw.blocks256 = y;
    pw.blocks_sub = z;
    pw.blocks_maj = u;
    pw.blocks_lsub = v;
    pw.narrow = narrow;
    return pw;
}

void copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "device to host copy");
}

// HIP-event timing of one launch on the library stream; resolved at the wave's next synchronisation
struct kernel_timer
{
    context &c;
    struct pending { hipEvent_t a, b; int kind; long long owners; };
    std::vector<pending> open;
    explicit kernel_timer(context &ctx) : c(ctx) {}
    // events of a call that ended early go back to the pool untimed
    ~kernel_timer()
    {
        for (const pending &p : open)
        {
            c.events.push_back(p.a);
            c.events.push_back(p.b);
        }
    }
    bool on() const { return (c.flags & GDX_COL_PROFILE) != 0; }
    hipEvent_t take()
    {
        if (c.events.empty())
        {
            hipEvent_t e;
            hip_check(hipEventCreate(&e), "create event");
            return e;
        }
        hipEvent_t e = c.events.back();
        c.events.pop_back();
        return e;
    }
    template <typename F>
    void time(int kind, long long owners, F &&launch)
    {
        if (!on() || owners == 0) { launch(); return; }
        pending p{take(), take(), kind, owners};
        hip_check(hipEventRecord(p.a, c.stream), "record event");
        launch();
        hip_check(hipEventRecord(p.b, c.stream), "record event");
        open.push_back(p);
    }
    // call after the stream has been synchronised
    void resolve()
    {
        for (const pending &p : open)
        {
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, p.a, p.b), "event time");
            c.timing.ms[p.kind] += ms;
            c.timing.launches[p.kind] += 1;
            c.timing.owners[p.kind] += p.owners;
            c.events.push_back(p.a);
            c.events.push_back(p.b);
        }
        open.clear();
    }
};

}  // namespace

// One operator call of one swarm (inc/swarm/_col_chain.cuh:108-539) as a sequence of phases: a rate
// phase (publish, bath-start rates, binning; host: bath durations and scheduler) and a chain phase
// (screen, chains, audit; host: bath audit and scheduler advance) per refresh wave. With
// diagnostics it also keeps the reference's controller archive and per-operator record
// (_col_chain.cuh:148-538). run_swarms drives several of these through shared launches.
struct op_run
{
    using wall = std::chrono::steady_clock;
    enum phase_t { RATE, CHAIN, DONE };

    context &c;
    diag_state *dg;
    double duration;
    bool diag;
    int tbase = 0, index = 0;                 // first table entry and swarm index within the run
    phase_t phase = RATE;
    bool initial = true;
    gdx_col_stats stats{};
    const gdx_raw_rate *raw_rate = nullptr;   // this swarm's raw bins in the run's host copies
    const gdx_raw_audit *raw_audit = nullptr;
    std::vector<int> counts, binmap, groups;
    std::vector<rate_bin> ratebin;
    std::vector<audit_bin> audit;
    int merged = 0;
    std::unique_ptr<schedule> sched;
    std::vector<double> requested, steps, current;
    std::vector<bool> passed;
    double dt_col = 0.0;
    uint64_t tick = 0;
    const int *err = nullptr;
    const unsigned long long *wave_ctr = nullptr;
    std::vector<group_stats> gs;
    operator_record rec;
    std::vector<double> published;
    int operator_index = 0, bath_index = 0;
    wall::time_point begin, chain_begin;
    // learned majorant: in use for this call (switched off for the rest of the call when the exact
    // checks of a wave find the bound too often below the exact rate); subsample phase per group
    bool learned = false, rate_learned = false, chain_learned = false, table_dirty = false;
    std::vector<int> sub_phase;

    op_run(context &ctx, diag_state *d, double h)
        : c(ctx), dg(d), duration(h), diag(ctx.diag && d),
          counts(RAW_BINS), binmap(RAW_BINS), ratebin(RAW_BINS), audit(RAW_BINS), steps(LOCAL_GROUPS),
          passed(LOCAL_GROUPS, true), published(LOCAL_GROUPS, 0.0)
    {
        require(c.state_loaded, GDX_COL_ERR_STATE, "collision state has not been loaded");
        if (!(duration > 0) || !std::isfinite(duration))
            throw status_error(GDX_COL_ERR_SCHEDULE, "invalid local collision horizon");
        begin = wall::now();
        learned = c.mlp != nullptr && c.n_active >= c.majorant_min_owners;
        c.thin_stride = c.n_active >= THIN_LARGE_OWNERS ? THIN_STRIDE_LARGE : THIN_STRIDE;
        sub_phase.assign(LOCAL_GROUPS, 0);
        stats.active_owners = c.n_active;
        dt_col = duration;
        groups.resize(LOCAL_GROUPS);
        for (int g = 0; g < LOCAL_GROUPS; ++g) groups[g] = g;
        if (diag) gs.assign(LOCAL_GROUPS, group_stats{});
    }

    void start(hipStream_t st, run_buffers &rb)
    {
        raw_rate = rb.h_rate.data() + size_t(tbase)*COL_BIN_S;
        raw_audit = rb.h_audit.data() + size_t(tbase)*COL_BIN_S;
        err = rb.h_err.data() + index;
        wave_ctr = rb.h_wave.data() + size_t(index)*GDX_WAVE_CTRS;
        if (diag)
            hip_check(hipMemsetAsync(c.evw.p, 0, sizeof(gdx_event_work)*std::max(1, c.n_active), st), "zero event work");
    }

    void entries(std::vector<phase_entry> &e, bool chain)
    {
        (chain ? chain_learned : rate_learned) = learned;
        for (int g : groups)
            if (c.group_cnt[g] > 0)
                e.push_back(phase_entry{tbase + g, c.group_beg[g], c.group_cnt[g], chain ? steps[g] : 0.0,
                                        learned ? sub_phase[g] % c.thin_stride : -1});
    }
    long long due_owners() const
    {
        long long n = 0;
        for (int g : groups) n += std::max(0, c.group_cnt[g]);
        return n;
    }

    int bin_end(int g) const { return g + 1 < LOCAL_GROUPS ? binmap[(g + 1)*COL_BIN_S] : merged; }
    double requested_step(int g, int *binding = nullptr) const
    {
        if (c.fixed_bath > 0.0)
        {
            if (binding) *binding = 0;
            return std::min(duration, c.fixed_bath);
        }
        int first = binmap[g*COL_BIN_S];
        return choose_bath(ratebin.data() + first, bin_end(g) - first, duration, c.state[g].limit_scale, binding);
    }

    // the next due groups with listed owners (_col_chain.cuh:291-313), or DONE at the horizon
    void next_groups()
    {
        while (sched->time() < sched->end)
        {
            tick = sched->time();
            groups = sched->due(tick);
            const double time = sched->seconds(tick);
            long long listed = 0;
            for (int g : groups)
            {
                listed += c.group_all[g];
                published[g] = time;
            }
            if (listed == 0)
            {
                sched->advance(groups);
                continue;
            }
            phase = tick != 0 ? RATE : CHAIN;
            if (phase == CHAIN) before_chain();
            return;
        }
        phase = DONE;
    }

    // merged rate bins of the rated groups (_col_chain.cuh:205-243)
    void after_rate()
    {
        for (int r = 0; r < RAW_BINS; ++r)
        {
            stats.pair_evals += (long long)raw_rate[r].count*N_K;
            stats.rated_owners += raw_rate[r].count;
        }
        // the exact subsample tests the bound: more than 5% above it and the swarm reverts to the plain
        // method for the rest of the call (a full rate pass of the same groups follows); otherwise the
        // bound's scale follows the violation rate towards 0.2%
        if (rate_learned)
        {
            const double checks = double(wave_ctr[GDX_WAVE_RATE_CHECKS]);
            const double viol = double(wave_ctr[GDX_WAVE_RATE_VIOLATIONS]);
            if (checks >= 64 && viol*20.0 > checks)
            {
                learned = false;
                adapt_bound(2.0);
                return;
            }
            if (viol > 0.002*checks + 2.0*std::sqrt(0.002*checks) + 1.0) adapt_bound(1.25);
            else if (viol*2000.0 < checks) adapt_bound(0.97);
        }
        // a learned rate pass rated one owner in thin_stride: bins merge on the population estimate
        for (int r = 0; r < RAW_BINS; ++r) counts[r] = raw_rate[r].count*(rate_learned ? c.thin_stride : 1);
        merged = build_binmap(counts, binmap);
        std::fill(ratebin.begin(), ratebin.begin() + merged, rate_bin{});
        for (int r = 0; r < RAW_BINS; ++r)
        {
            rate_bin &b = ratebin[binmap[r]];
            b.mass += raw_rate[r].mass;
            b.weighted_rate += raw_rate[r].weighted_rate;
            b.weighted_change += raw_rate[r].weighted_change;
            b.weighted_second += raw_rate[r].weighted_second;
            b.owner_count += raw_rate[r].owner_count;
            b.invalid_count += raw_rate[r].invalid_count;
        }
        if (initial)
        {
            initial = false;
            requested.assign(LOCAL_GROUPS, 0.0);
            for (int g = 0; g < LOCAL_GROUPS; ++g) requested[g] = requested_step(g, diag ? &gs[g].binding : nullptr);
            sched = std::make_unique<schedule>(duration, requested, c.edges);
            for (int g = 0; g < LOCAL_GROUPS; ++g) steps[g] = sched->seconds(sched->step[g]);
            current = requested;
            if (diag)
            {
                for (int g = 0; g < LOCAL_GROUPS; ++g)
                {
                    gs[g].initial_level = sched->level[g];
                    gs[g].min_request = gs[g].max_request = gs[g].initial_request = requested[g];
                }
                operator_index = ++dg->summary.operator_count;
            }
            next_groups();
            return;
        }
        for (int g : groups)
        {
            current[g] = requested_step(g);
            if (diag)
            {
                gs[g].min_request = std::min(gs[g].min_request, current[g]);
                gs[g].max_request = std::max(gs[g].max_request, current[g]);
            }
        }
        std::vector<int> previous = sched->level;
        sched->adapt(tick, groups, current, passed);
        for (int g : groups)
        {
            steps[g] = sched->seconds(sched->step[g]);
            if (diag)
            {
                gs[g].coarsened += sched->level[g] < previous[g];
                gs[g].refined += sched->level[g] > previous[g];
                gs[g].constrained += steps[g] > current[g];
            }
        }
        phase = CHAIN;
        before_chain();
    }

    // scale the swarm's bound (within [1, 16]); the group table is uploaded again before the next rate pass
    void adapt_bound(double f)
    {
        double v = std::min(16.0, std::max(1.0, c.safety_scale*f));
        if (v != c.safety_scale)
        {
            c.safety_scale = v;
            table_dirty = true;
        }
    }

    void before_chain()
    {
        if (diag)
        {
            const double time = sched->seconds(tick);
            for (int g : groups)
            {
                if (c.group_all[g] == 0) continue;
                for (int d = 0; d < LOCAL_GROUPS; ++d)
                    if (c.edges[size_t(g)*LOCAL_WORDS + d / 32] & (1u << (d % 32)))
                        gs[g].max_age = std::max(gs[g].max_age, time - published[d]);
                gs[g].max_requested_ratio = std::max(gs[g].max_requested_ratio, steps[g] / requested_step(g));
            }
        }
        chain_begin = wall::now();
    }

    // bath audits (_col_chain.cuh:386-470), then the scheduler advance
    void after_chain()
    {
        if (*err) throw status_error(*err, "local collision chain error " + std::to_string(*err));
        const auto audit_begin = wall::now();
        if (diag) rec.chain_seconds += std::chrono::duration<double>(audit_begin - chain_begin).count();
        std::fill(audit.begin(), audit.begin() + merged, audit_bin{});
        for (int r = 0; r < RAW_BINS; ++r)
        {
            const gdx_raw_audit &v = raw_audit[r];
            audit_bin &b = audit[binmap[r]];
            b.mass += v.mass;
            b.predicted_f += v.predicted_f;
            b.predicted_e += v.predicted_e;
            b.predicted_var_f += v.predicted_var_f;
            b.predicted_var_e += v.predicted_var_e;
            b.predicted_g += v.predicted_g;
            b.predicted_var_g += v.predicted_var_g;
            double mw, mg;
            std::memcpy(&mw, &v.maximum_weight_bits, sizeof(double));
            std::memcpy(&mg, &v.maximum_g_jump_bits, sizeof(double));
            b.maximum_weight = std::max(b.maximum_weight, mw);
            b.maximum_g_jump = std::max(b.maximum_g_jump, mg);
            b.touched += v.touched;
            b.event_weight += v.event_weight;
            b.growth += v.growth;
            b.end_mass += v.end_mass;
            b.owner_count += v.owner_count;
            b.invalid_count += v.invalid_count;
        }
        for (int g : groups)
        {
            if (c.group_all[g] == 0) continue;
            int first = binmap[g*COL_BIN_S], last = bin_end(g);
            double mass = 0.0;
            for (int b = first; b < last; ++b)
            {
                if (audit[b].invalid_count) throw status_error(GDX_COL_ERR_AUDIT, "invalid local audit state");
                mass += audit[b].mass;
            }
            if (!(mass > 0)) continue;
            const double before = c.state[g].limit_scale;
            bath_result r = finish_bath(audit.data() + first, last - first, c.state[g]);
            passed[g] = !(r.activity_overshoot || r.distribution_overshoot || r.persistent_overshoot);
            stats.baths++;
            if (diag)
            {
                group_stats &s = gs[g];
                ++s.updates;
                s.overshoots += r.activity_overshoot || r.distribution_overshoot;
                s.persistent += r.persistent_overshoot;
                s.max_growth = std::max(s.max_growth, r.max_g);
                s.max_activity = std::max(s.max_activity, r.max_f);
                bath_record b;
                b.group_index = g;
                b.operator_index = operator_index;
                b.bath_index = ++bath_index;
                b.merged_bins = last - first;
                b.duration = steps[g];
                b.limit_before = before;
                b.limit_after = c.state[g].limit_scale;
                b.result = r;
                record_bath(dg->summary, b);
            }
        }
        if (diag)
        {
            // launches the reference needs: one per COL_EVENT_CAP accepted events of the longest chain
            const int continuations = wave_ctr[2] ? int(wave_ctr[0] / COL_EVENT_CAP + 1) : 0;
            dg->summary.continuation_launches += continuations;
            dg->summary.wave_count++;
            rec.chain_launches += continuations;
            rec.chain_blocks += wave_ctr[1];
            rec.audit_seconds += std::chrono::duration<double>(wall::now() - audit_begin).count();
        }
        stats.waves++;
        stats.owner_updates += due_owners();
        dt_col = duration;
        for (int g : groups) dt_col = std::min(dt_col, steps[g]);
        if (chain_learned)
        {
            for (int g : groups)
            {
                const int cnt = std::max(0, c.group_cnt[g]), sub = sub_phase[g] % c.thin_stride;
                stats.thinned += cnt - thin_rated(cnt, sub, c.thin_stride);
                sub_phase[g]++;
            }
            // candidates above their bound (each taken exactly by superposition) also raise the scale
            const double checks = double(wave_ctr[GDX_WAVE_CHECKS]), viol = double(wave_ctr[GDX_WAVE_VIOLATIONS]);
            if (viol > 0.01*checks + 3.0) adapt_bound(1.25);
        }
        sched->advance(groups);
        next_groups();
    }

    // counters: this swarm's four run counters, read back after the run
    void finish(hipStream_t st, const unsigned long long *counters, gdx_col_stats *out)
    {
        if (diag)
        {
            launch_event_sum(c, st);
            copy_d2h(&dg->totals, c.ev_sum.p, sizeof(gdx_event_work), st);
        }
        hip_check(hipStreamSynchronize(st), "operator");
        if (!(c.flags & GDX_COL_DEFER_WRITEBACK))
        {
            store_state(c);
            hip_check(hipStreamSynchronize(c.stream), "write-back");
        }
        stats.events = (long long)counters[GDX_CTR_EVENTS];
        stats.chain_owners = (long long)counters[GDX_CTR_CHAINED];
        stats.pair_evals += (long long)counters[GDX_CTR_PAIRS];
        stats.max_chain_events = (long long)counters[GDX_CTR_LONGEST];

        stats.candidates = (long long)counters[GDX_CTR_CANDIDATES];
        stats.accepted = (long long)counters[GDX_CTR_ACCEPTED];
        stats.exact_checks = (long long)counters[GDX_CTR_CHECKS];
        stats.violations = (long long)counters[GDX_CTR_VIOLATIONS];
        stats.candidate_pairs = (long long)counters[GDX_CTR_CAND_PAIRS];
        stats.bound_evals = 0;
        for (int k = 0; k < GDX_CTR_SPREAD; ++k) stats.bound_evals += (long long)counters[GDX_CTR_BOUND_EVALS + k];
        stats.dt_col = dt_col;
        stats.clock_dyn = duration;
        c.last_waves = stats.waves;
        if (diag && dg->log.is_open())
        {
            rec.clock_sim = dg->clock_sim;
            rec.duration = duration;
            rec.waves = uint64_t(stats.waves);
            rec.owner_updates = uint64_t(stats.owner_updates);
            rec.finest_ticks = sched ? sched->end : 0;
            rec.owners.assign(c.group_all.begin(), c.group_all.end());
            rec.level = sched ? sched->level : std::vector<int>(LOCAL_GROUPS, 0);
            rec.dt = steps;
            rec.final_request = current.empty() ? std::vector<double>(LOCAL_GROUPS, 0.0) : current;
            rec.groups = gs;
            for (int k = 0; k < GDX_EVENT_CATEGORIES; ++k)
            {
                rec.event_count[k] = dg->totals.count[k];
                rec.event_log_mass[k] = dg->totals.log_mass[k];
            }
            rec.scheduler_wall_seconds = std::chrono::duration<double>(wall::now() - begin).count();
            write_operator_line(dg->log, ++dg->operator_id, rec);
            dg->log.flush();
            if (!dg->log) throw status_error(GDX_COL_ERR_ARG, "cannot write local collision diagnostics");
        }
        if (out) *out = stats;
    }
};

// Advance several swarms through one operator horizon each: every loop runs one rate phase over the
// due groups of all swarms that need rates and one chain phase over those whose baths are set, with
// one launch per kernel and one host synchronisation per phase for all of them.
void run_swarms(std::vector<op_run *> &runs, run_buffers &rb, hipStream_t st, kernel_timer &timer)
{
    const int nr = int(runs.size());
    const int ng = nr*LOCAL_GROUPS;
    const int narrow_owners = runs[0]->c.narrow_owners;
    rb.size_run(ng, nr);
    rb.h_table.clear();
    for (op_run *r : runs)
    {
        // swarms with a majorant track which bounds their state changes touch: the reverse neighbour
        // lists of the epoch once, and a full evaluation after the model's parameters change
        context &c = r->c;
        if (c.model && !c.rev_valid) build_reverse(c);
        else if (c.model && majorant_version(c.model) != c.model_version) invalidate_bounds(c);
    }
    bool track = false;
    for (op_run *r : runs) track |= r->c.model && r->c.rev_valid;
    for (int i = 0; i < nr; ++i)
    {
        op_run *r = runs[i];
        r->index = i;
        r->tbase = i*LOCAL_GROUPS;
        for (int g = 0; g < LOCAL_GROUPS; ++g) rb.h_table.push_back(make_gref(r->c, g, rb, i, r->tbase));
    }
    rb.table.alloc(ng);
    rb.reserve(ng);
    hip_check(hipMemcpyAsync(rb.table.p, rb.h_table.data(), sizeof(gdx_gref)*ng, hipMemcpyHostToDevice, st),
              "group table");
    hip_check(hipMemsetAsync(rb.counters.p, 0, sizeof(unsigned long long)*GDX_CTRS*nr, st), "zero counters");
    if (!rb.resident) rb.resident = chain_resident_blocks();
    for (op_run *r : runs) r->start(st, rb);
    launch_reset_scratch(st, rb.table.p, ng);
    // swarms whose bound scale changed: their table entries again, before the next rate pass
    auto upload_changed = [&](const std::vector<op_run *> &rs) {
        for (op_run *r : rs)
        {
            if (!r->table_dirty) continue;
            r->table_dirty = false;
            for (int g = 0; g < LOCAL_GROUPS; ++g)
                rb.h_table[size_t(r->tbase + g)] = make_gref(r->c, g, rb, r->index, r->tbase);
            hip_check(hipMemcpyAsync(rb.table.p + r->tbase, rb.h_table.data() + r->tbase,
                                     sizeof(gdx_gref)*LOCAL_GROUPS, hipMemcpyHostToDevice, st), "group table");
        }
    };
    std::vector<phase_entry> e;
    std::vector<op_run *> now;
    while (true)
    {
        bool work = false;
        // ---- rate phase ----
        e.clear();
        now.clear();
        for (op_run *r : runs)
            if (r->phase == op_run::RATE) { now.push_back(r); r->entries(e, false); }
        if (!now.empty())
        {
            work = true;
            zero_ranges z;
            z.add(rb.raw_rate.p, sizeof(gdx_raw_rate)*COL_BIN_S*ng);
            z.add(rb.wave_ctr.p, sizeof(unsigned long long)*GDX_WAVE_CTRS*nr);
            z.add(rb.nchanged.p, sizeof(int)*CHANGED_LISTS);
            phase_wave pw = upload_phase(rb, e, st, z, narrow_owners);
            timer.time(GDX_COL_KERNEL_PUBLISH, pw.w256.total, [&] { launch_publish(st, pw.w1024, pw.blocks1024, rb); });
            if (track && pw.blocks1024)
                timer.time(GDX_COL_KERNEL_STALE, pw.w256.total, [&] { launch_mark_stale(st, pw.w1024, rb); });
            if (pw.blocks_maj > 0)
                timer.time(GDX_COL_KERNEL_MAJORANT, pw.w256.total, [&] { launch_majorant(st, pw.wmaj, pw.blocks_maj); });
            timer.time(GDX_COL_KERNEL_RATE, pw.w256.total, [&] {
                launch_bath_rate(st, pw.wsub, pw.blocks_sub);
                launch_bath_rate_grouped(st, pw.wlsub, pw.blocks_lsub);
            });
            rb.read_back(rb.off_rate, rb.off_err, st);             // rate bins and wave counters
            wait_stream(st, "rating pass");
            for (op_run *r : now) r->after_rate();
            upload_changed(now);
        }
        // ---- chain phase ----
        e.clear();
        now.clear();
        for (op_run *r : runs)
            if (r->phase == op_run::CHAIN) { now.push_back(r); r->entries(e, true); }
        if (!now.empty())
        {
            work = true;
            zero_ranges z;
            z.add(rb.raw_audit.p, sizeof(gdx_raw_audit)*COL_BIN_S*ng);
            z.add(rb.err.p, sizeof(int)*nr);
            z.add(rb.qcount.p, sizeof(int)*ng);
            z.add(rb.hcount.p, sizeof(int)*ng);
            z.add(rb.ccount.p, sizeof(int)*ng);
            z.add(rb.wave_ctr.p, sizeof(unsigned long long)*GDX_WAVE_CTRS*nr);
            phase_wave pw = upload_phase(rb, e, st, z, narrow_owners);
            timer.time(GDX_COL_KERNEL_SCREEN, pw.w256.total, [&] {
                if (pw.narrow) launch_screen(st, pw.w256, pw.blocks256, true);
                else launch_screen(st, pw.w1024, pw.blocks1024, false);
            });
            if (pw.w256.learned && pw.blocks256)
                timer.time(GDX_COL_KERNEL_CANDIDATES, pw.w256.total, [&] { launch_candidates(st, pw.w256, pw.blocks256, rb.cprefix.p); });
            if (pw.blocks256)
                timer.time(GDX_COL_KERNEL_PREFIX, pw.w256.total, [&] { launch_queue_prefix(st, pw.w256, rb.qprefix.p); });
            timer.time(GDX_COL_KERNEL_CHAIN, pw.w256.total, [&] { launch_chain(st, pw.w256, rb.resident); });
            timer.time(GDX_COL_KERNEL_AUDIT, pw.w256.total, [&] { launch_chain_audit(st, pw.w1024, pw.blocks1024); });
            rb.read_back(rb.off_wave, rb.back_bytes, st);          // wave counters, errors, audit bins
            wait_stream(st, "chain pass");
            for (op_run *r : now) r->after_chain();
            upload_changed(now);
        }
        if (!work) break;
    }
    copy_d2h(rb.h_counters.data(), rb.counters.p, sizeof(unsigned long long)*GDX_CTRS*nr, st);
}

void run_operator(context &c, diag_state *dg, run_buffers &rb, double duration, gdx_col_stats *out)
{
    op_run r(c, dg, duration);
    std::vector<op_run *> runs{&r};
    kernel_timer timer(c);
    run_swarms(runs, rb, c.stream, timer);
    r.finish(c.stream, rb.h_counters.data(), out);
    // the call's kernel times, once the stream has drained (not between phases, where the host's
    // event queries would hold up the next launch)
    timer.resolve();
}

}  // namespace gdx

using namespace gdx;

namespace {

// one batched call: every swarm's device work on the first handle's stream and run buffers
void run_batch(std::vector<gdx_col_ctx *> &hs, double duration, const std::vector<gdx_col_stats *> &out)
{
    gdx_col_ctx *h0 = hs[0];
    std::vector<std::unique_ptr<op_run>> owned;
    std::vector<op_run *> runs;
    for (gdx_col_ctx *h : hs)
    {
        owned.push_back(std::make_unique<op_run>(h->c, &h->diag, duration));
        runs.push_back(owned.back().get());
    }
    for (size_t i = 1; i < hs.size(); ++i) hip_check(hipStreamSynchronize(hs[i]->c.stream), "batch stream");
    kernel_timer timer(h0->c);
    run_swarms(runs, h0->rb, h0->c.stream, timer);
    for (size_t i = 0; i < hs.size(); ++i)
        runs[i]->finish(h0->c.stream, h0->rb.h_counters.data() + size_t(GDX_CTRS)*i, out[i]);
    timer.resolve();
}

// streams of concurrent ensemble groups, per device: created on first use at the default priority and
// kept until the process ends
constexpr int ENSEMBLE_STREAMS = 4, ENSEMBLE_MAX_STREAMS = 8;
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams[64];

const std::vector<hipStream_t> &ensemble_streams(int count)
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "current device");
    require(dev >= 0 && dev < 64, GDX_COL_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_stream_mu);
    std::vector<hipStream_t> &v = g_streams[dev];
    while (int(v.size()) < count)
    {
        hipStream_t st = nullptr;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "ensemble stream");
        v.push_back(st);
    }
    return v;
}

}  // namespace

extern "C" {

int gdx_col_model_info(gdx_col_model *o)
{
    if (!o) return GDX_COL_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    std::strncpy(o->model, GDX_MODEL_NAME, sizeof(o->model) - 1);
    o->n_k = N_K; o->dim = DIM; o->coag_kernel = COAG_KERNEL; o->n_x = N_X; o->n_y = N_Y; o->n_z = N_Z;
    o->local_groups = LOCAL_GROUPS; o->col_bin_s = COL_BIN_S;
    o->morton_leaf_target = MORTON_LEAF_TARGET; o->morton_max_level = MORTON_MAX_LEVEL;
    o->x_min = X_MIN; o->x_max = X_MAX; o->y_min = Y_MIN; o->y_max = Y_MAX; o->z_min = Z_MIN; o->z_max = Z_MAX;
    o->h_search = H_SEARCH; o->v_frag = V_FRAG; o->col_bath_max = COL_BATH_MAX; o->col_bath_eps = COL_BATH_EPS;
    o->col_bath_alpha = COL_BATH_ALPHA; o->init_smin = INIT_SMIN; o->rho_0 = RHO_0;
    return GDX_COL_OK;
}

const char *gdx_col_last_error(gdx_col_handle h)
{
    return h ? h->c.last_error.c_str() : g_last_error.c_str();
}

int gdx_col_create(const gdx_col_config *cfg, gdx_col_handle *out)
{
    if (!cfg || !out) return GDX_COL_ERR_ARG;
    *out = nullptr;
    if (cfg->n <= 0 || cfg->n > INT_MAX / 3)
    {
        g_last_error = "representative count must be in [1, INT_MAX/3]";
        return GDX_COL_ERR_ARG;
    }
    if (!(cfg->total_dust_mass > 0) || !std::isfinite(cfg->total_dust_mass))
    {
        g_last_error = "total dust mass must be positive and finite";
        return GDX_COL_ERR_ARG;
    }
    if (cfg->search != GDX_COL_SEARCH_MORTON && cfg->search != GDX_COL_SEARCH_KDTREE)
    {
        g_last_error = "search must be GDX_COL_SEARCH_MORTON or GDX_COL_SEARCH_KDTREE";
        return GDX_COL_ERR_ARG;
    }
    gdx_col_ctx *h = nullptr;
    int rc = guard(nullptr, [&] {
        h = new gdx_col_ctx();
        context &c = h->c;
        c.n = cfg->n;
        c.total_mass = cfg->total_dust_mass;
        c.seed = cfg->seed;
        c.flags = cfg->flags;
        c.search = cfg->search;
        c.diag = (cfg->flags & GDX_COL_DIAGNOSTICS) != 0;
        if (c.diag) c.ev_sum.alloc(1);
        c.state.assign(LOCAL_GROUPS, context::bath_state{});
        c.rng_p.alloc(size_t(c.n));
        hip_check(hipMemset(c.rng_p.p, 0, sizeof(uint64_t)*c.n), "zero counters");
        c.scr_min.alloc(LOCAL_GROUPS);
        c.scr_max.alloc(LOCAL_GROUPS);
        c.lower.alloc(LOCAL_GROUPS);
        c.upper.alloc(LOCAL_GROUPS);
        hip_check(hipMemset(c.lower.p, 0, sizeof(double)*LOCAL_GROUPS), "zero bounds");
        hip_check(hipMemset(c.upper.p, 0, sizeof(double)*LOCAL_GROUPS), "zero bounds");
    });
    if (rc != GDX_COL_OK)
    {
        delete h;
        return rc;
    }
    *out = h;
    return GDX_COL_OK;
}

int gdx_col_destroy(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    (void)hipStreamSynchronize(h->c.stream);
    for (hipEvent_t e : h->c.events) (void)hipEventDestroy(e);
    delete h;
    return GDX_COL_OK;
}

int gdx_col_bind_soa(gdx_col_handle h, const double *x, const double *y, const double *z, double *size, double *numr)
{
    if (!h || !x || !y || !z || !size || !numr) return GDX_COL_ERR_ARG;
    bound_state &u = h->c.user;
    u = bound_state{};
    u.x = {reinterpret_cast<const char *>(x), 8};
    u.y = {reinterpret_cast<const char *>(y), 8};
    u.z = {reinterpret_cast<const char *>(z), 8};
    u.size = {reinterpret_cast<const char *>(size), 8};
    u.numr = {reinterpret_cast<const char *>(numr), 8};
    u.bound = true;
    h->c.geometry_valid = h->c.state_loaded = false;
    return GDX_COL_OK;
}

int gdx_col_bind_swarm(gdx_col_handle h, void *records)
{
    if (!h || !records) return GDX_COL_ERR_ARG;
    const char *b = static_cast<const char *>(records);
    bound_state &u = h->c.user;
    u = bound_state{};
    // swarm layout (inc/swarm/const_defs.cuh:194-203): position xyz, velocity xyz, par_size, par_numr
    u.x = {b + 0, 64}; u.y = {b + 8, 64}; u.z = {b + 16, 64};
    u.vx = {b + 24, 64}; u.vy = {b + 32, 64}; u.vz = {b + 40, 64};
    u.size = {b + 48, 64}; u.numr = {b + 56, 64};
    u.bound = true;
    h->c.geometry_valid = h->c.state_loaded = false;
    return GDX_COL_OK;
}

int gdx_col_set_stream(gdx_col_handle h, void *stream)
{
    if (!h) return GDX_COL_ERR_ARG;
    h->c.stream = static_cast<hipStream_t>(stream);
    return GDX_COL_OK;
}

int gdx_col_geometry(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(h->c.user.bound, GDX_COL_ERR_STATE, "no particle state is bound");
        geometry_epoch(h->c);
    });
}

int gdx_col_check_state(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(h->c.user.bound, GDX_COL_ERR_STATE, "no particle state is bound");
        int bad = state_flag(h->c);
        if (bad) throw status_error(GDX_COL_ERR_NONFINITE,
            "non-finite particle state before collision search at particle " + std::to_string(bad - 1));
    });
}

int gdx_col_load_state(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] { load_state(h->c); });
}

int gdx_col_store_state(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        store_state(h->c);
        hip_check(hipStreamSynchronize(h->c.stream), "store state");
    });
}

int gdx_col_operator(gdx_col_handle h, double duration, gdx_col_stats *out)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] { run_operator(h->c, &h->diag, h->rb, duration, out); });
}

int gdx_col_operator_batch(const gdx_col_handle *handles, int count, double duration, gdx_col_stats *out)
{
    if (!handles || count <= 0 || !handles[0]) return GDX_COL_ERR_ARG;
    for (int i = 0; i < count; ++i)
    {
        if (!handles[i]) return GDX_COL_ERR_ARG;
        for (int j = 0; j < i; ++j)
            if (handles[j] == handles[i]) return GDX_COL_ERR_ARG;
    }
    std::vector<gdx_col_ctx *> hs(handles, handles + count);
    std::vector<gdx_col_stats *> outs(static_cast<size_t>(count), nullptr);
    if (out)
        for (int i = 0; i < count; ++i) outs[size_t(i)] = out + i;
    return guard(&handles[0]->c, [&] { run_batch(hs, duration, outs); });
}

int gdx_col_operator_ensemble(const gdx_col_handle *handles, int count, double duration, int steps, int streams,
                              gdx_col_stats *out)
{
    if (!handles || count <= 0 || !handles[0] || steps <= 0) return GDX_COL_ERR_ARG;
    for (int i = 0; i < count; ++i)
    {
        if (!handles[i]) return GDX_COL_ERR_ARG;
        for (int j = 0; j < i; ++j)
            if (handles[j] == handles[i]) return GDX_COL_ERR_ARG;
    }
    gdx_col_ctx *h0 = handles[0];
    return guard(&h0->c, [&] {
        const int lanes = std::max(1, std::min({streams > 0 ? streams : ENSEMBLE_STREAMS, count, ENSEMBLE_MAX_STREAMS}));
        // a realization's critical path is its sequence of refresh waves, each a few dependent launches and
        // host synchronisations: realizations ordered by the waves of their last call (before any call, by
        // active-owner count) and dealt in turn, so the heaviest ones land on different streams (measured on
        // 64-realization ensembles: better than greedy balancing of the wave sums, which stacks the many
        // light realizations onto the streams of the heavy ones)
        std::vector<int> order(static_cast<size_t>(count));
        for (int i = 0; i < count; ++i) order[size_t(i)] = i;
        bool known = true;
        for (int i = 0; i < count; ++i) known &= handles[i]->c.last_waves > 0;
        auto cost = [&](int i) -> long long {
            return known ? (long long)handles[i]->c.last_waves : (long long)handles[i]->c.n_active;
        };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
        std::vector<std::vector<int>> part(static_cast<size_t>(lanes));
        for (int p = 0; p < count; ++p) part[size_t(p % lanes)].push_back(order[size_t(p)]);
        std::vector<hipStream_t> saved(static_cast<size_t>(count));
        for (int i = 0; i < count; ++i)
        {
            saved[size_t(i)] = handles[i]->c.stream;
            hip_check(hipStreamSynchronize(handles[i]->c.stream), "ensemble stream");
        }
        const std::vector<hipStream_t> &pool = ensemble_streams(lanes);
        // one host thread per stream; each group advances `steps` batched calls on its own stream, independent
        // of the other groups (the realizations are independent)
        std::vector<int> rc(static_cast<size_t>(lanes), GDX_COL_OK);
        auto work = [&](int g) {
            std::vector<gdx_col_ctx *> hs;
            for (int i : part[size_t(g)])
            {
                handles[i]->c.stream = pool[size_t(g)];
                hs.push_back(handles[i]);
            }
            if (hs.empty()) return;
            rc[size_t(g)] = guard(&hs[0]->c, [&] {
                std::vector<gdx_col_stats *> outs(hs.size(), nullptr);
                for (int s = 0; s < steps; ++s)
                {
                    if (out)
                        for (size_t k = 0; k < hs.size(); ++k)
                            outs[k] = out + size_t(s)*size_t(count) + size_t(part[size_t(g)][k]);
                    run_batch(hs, duration, outs);
                }
            });
        };
        std::vector<std::thread> th;
        for (int g = 1; g < lanes; ++g) th.emplace_back(work, g);
        work(0);
        for (std::thread &t : th) t.join();
        for (int i = 0; i < count; ++i) handles[i]->c.stream = saved[size_t(i)];
        for (int g = 0; g < lanes; ++g)
            if (rc[size_t(g)] != GDX_COL_OK)
                throw status_error(rc[size_t(g)], "ensemble stream " + std::to_string(g) + ": "
                                   + handles[part[size_t(g)][0]]->c.last_error);
    });
}

int gdx_col_reset_controller(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    h->c.state.assign(LOCAL_GROUPS, context::bath_state{});
    // the learned bound's adaptive scale is controller memory too: a restart from a checkpoint written
    // at this boundary continues with the same scale
    h->c.safety_scale = 1.0;
    h->diag.summary = controller_summary{};
    return GDX_COL_OK;
}

int gdx_col_open_log(gdx_col_handle h, const char *dir)
{
    if (!h || !dir) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(h->c.diag, GDX_COL_ERR_STATE, "the handle was created without GDX_COL_DIAGNOSTICS");
        auto stamp = std::chrono::duration_cast<std::chrono::microseconds>(
            std::chrono::system_clock::now().time_since_epoch()).count();
        std::string path = std::string(dir) + "collision_local_" + std::to_string(stamp) + ".jsonl";
        if (h->diag.log.is_open()) h->diag.log.close();
        h->diag.log.clear();
        h->diag.log.open(path);
        if (!h->diag.log) throw status_error(GDX_COL_ERR_ARG, "cannot open local collision diagnostics " + path);
        h->diag.log << std::setprecision(17);
    });
}

int gdx_col_set_clock(gdx_col_handle h, double clock_sim)
{
    if (!h) return GDX_COL_ERR_ARG;
    h->diag.clock_sim = clock_sim;
    return GDX_COL_OK;
}

int gdx_col_save_controller(gdx_col_handle h, const char *file_name)
{
    if (!h || !file_name) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(h->c.diag, GDX_COL_ERR_STATE, "the handle was created without GDX_COL_DIAGNOSTICS");
        if (!save_controller(file_name, h->diag.summary))
            throw status_error(GDX_COL_ERR_ARG, std::string("cannot save controller archive ") + file_name);
    });
}

int gdx_col_event_totals(gdx_col_handle h, unsigned long long *counts, double *log_mass)
{
    if (!h || !counts || !log_mass) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(h->c.diag, GDX_COL_ERR_STATE, "the handle was created without GDX_COL_DIAGNOSTICS");
        for (int k = 0; k < GDX_EVENT_CATEGORIES; ++k)
        {
            counts[k] = h->diag.totals.count[k];
            log_mass[k] = h->diag.totals.log_mass[k];
        }
    });
}

int gdx_col_synchronize(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] { hip_check(hipStreamSynchronize(h->c.stream), "synchronize"); });
}

int gdx_col_kernel_times(gdx_col_handle h, double *ms, long long *launches, long long *owners)
{
    if (!h) return GDX_COL_ERR_ARG;
    for (int k = 0; k < GDX_COL_KERNELS; ++k)
    {
        if (ms) ms[k] = h->c.timing.ms[k];
        if (launches) launches[k] = h->c.timing.launches[k];
        if (owners) owners[k] = h->c.timing.owners[k];
    }
    return GDX_COL_OK;
}

int gdx_col_reset_kernel_times(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    h->c.timing = context::timing_t{};
    return GDX_COL_OK;
}

int gdx_col_active_count(gdx_col_handle h, int *n_active)
{
    if (!h || !n_active) return GDX_COL_ERR_ARG;
    if (!h->c.geometry_valid) { h->c.last_error = "collision geometry has not been built"; return GDX_COL_ERR_STATE; }
    *n_active = h->c.n_active;
    return GDX_COL_OK;
}

int gdx_col_record_info(gdx_col_handle h, int *records, int *unique_ids, float *q_max)
{
    if (!h) return GDX_COL_ERR_ARG;
    if (!h->c.geometry_valid) { h->c.last_error = "collision geometry has not been built"; return GDX_COL_ERR_STATE; }
    if (records) *records = h->c.search == GDX_COL_SEARCH_KDTREE ? h->c.kd.n : h->c.index.n;
    if (unique_ids) *unique_ids = h->c.unique_ids ? 1 : 0;
    if (q_max) *q_max = h->c.q_max;
    return GDX_COL_OK;
}

int gdx_col_get_neighbors(gdx_col_handle h, int *codes, double *measure)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.geometry_valid, GDX_COL_ERR_STATE, "collision geometry has not been built");
        const size_t na = size_t(c.n_active);
        std::vector<int> nbr(na*N_K), slot_pidx(na), pidx_slot(c.n);
        std::vector<unsigned char> image(X_WEDGE ? na*N_K : 0);
        std::vector<double> meas(na);
        hipStream_t st = c.stream;
        if (na)
        {
            copy_d2h(nbr.data(), c.nbr.p, sizeof(int)*na*N_K, st);
            if (X_WEDGE) copy_d2h(image.data(), c.nbr_image.p, na*N_K, st);
            copy_d2h(meas.data(), c.measure.p, sizeof(double)*na, st);
            copy_d2h(slot_pidx.data(), c.slot_pidx.p, sizeof(int)*na, st);
        }
        copy_d2h(pidx_slot.data(), c.pidx_slot.p, sizeof(int)*c.n, st);
        hip_check(hipStreamSynchronize(st), "read neighbours");
        for (int p = 0; p < c.n; ++p)
        {
            int s = pidx_slot[p];
            for (int k = 0; k < N_K; ++k)
            {
                int v = -1;
                if (s >= 0)
                {
                    const size_t e = size_t(s)*N_K + bitrev_slot(k);
                    int j = nbr[e];
                    v = j < 0 ? -1 : 3*slot_pidx[j] + (X_WEDGE ? int(image[e]) : 0);
                }
                if (codes) codes[size_t(p)*N_K + k] = v;
            }
            if (measure) measure[p] = s >= 0 ? meas[s] : 0.0;
        }
    });
}

int gdx_col_get_sites(gdx_col_handle h, float *pts4, unsigned char *active, unsigned char *group)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.geometry_valid, GDX_COL_ERR_STATE, "collision geometry has not been built");
        if (pts4) copy_d2h(pts4, c.pts.p, sizeof(float4)*c.n, c.stream);
        if (active) copy_d2h(active, c.active.p, c.n, c.stream);
        if (group) copy_d2h(group, c.group.p, c.n, c.stream);
        hip_check(hipStreamSynchronize(c.stream), "read sites");
    });
}

int gdx_col_get_morton(gdx_col_handle h, unsigned long long *keys, int *idx)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.geometry_valid, GDX_COL_ERR_STATE, "collision geometry has not been built");
        if (keys) copy_d2h(keys, c.index.keys.p, sizeof(uint64_t)*c.index.n, c.stream);
        if (idx) copy_d2h(idx, c.index.idx.p, sizeof(int)*c.index.n, c.stream);
        hip_check(hipStreamSynchronize(c.stream), "read morton");
    });
}

int gdx_col_get_kdtree(gdx_col_handle h, int *node_idx, int *node_dim, float *world6)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.geometry_valid && c.search == GDX_COL_SEARCH_KDTREE, GDX_COL_ERR_STATE,
                "no k-d tree: geometry not built with GDX_COL_SEARCH_KDTREE");
        const int n = c.kd.n;
        if (node_idx) copy_d2h(node_idx, c.kd.idx.p, sizeof(int)*n, c.stream);
        if (world6) copy_d2h(world6, c.kd.world.p, sizeof(float)*6, c.stream);
        if (node_dim)
        {
            std::vector<float4> nodes(n);
            copy_d2h(nodes.data(), c.kd.node.p, sizeof(float4)*n, c.stream);
            hip_check(hipStreamSynchronize(c.stream), "read k-d nodes");
            for (int i = 0; i < n; ++i)
            {
                int d;
                std::memcpy(&d, &nodes[i].w, sizeof d);
                node_dim[i] = d;
            }
        }
        hip_check(hipStreamSynchronize(c.stream), "read k-d tree");
    });
}

int gdx_col_get_cached(gdx_col_handle h, double *cached4)
{
    if (!h || !cached4) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.state_loaded, GDX_COL_ERR_STATE, "collision state has not been loaded");
        dbuf<double> tmp;
        tmp.alloc(size_t(4)*c.n);
        hip_check(hipMemsetAsync(tmp.p, 0, sizeof(double)*4*c.n, c.stream), "zero cached");
        launch_gather_cached(c, tmp.p);
        copy_d2h(cached4, tmp.p, sizeof(double)*4*c.n, c.stream);
        hip_check(hipStreamSynchronize(c.stream), "read cached");
    });
}

int gdx_col_get_rng(gdx_col_handle h, unsigned long long *counters)
{
    if (!h || !counters) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        copy_d2h(counters, c.rng_p.p, sizeof(uint64_t)*c.n, c.stream);
        if (c.state_loaded && c.n_active)
        {
            std::vector<uint64_t> slot(c.n_active);
            std::vector<int> sp(c.n_active);
            copy_d2h(slot.data(), c.rng.p, sizeof(uint64_t)*c.n_active, c.stream);
            copy_d2h(sp.data(), c.slot_pidx.p, sizeof(int)*c.n_active, c.stream);
            hip_check(hipStreamSynchronize(c.stream), "read rng");
            for (int s = 0; s < c.n_active; ++s) counters[sp[s]] = slot[s];
        }
        hip_check(hipStreamSynchronize(c.stream), "read rng");
    });
}

int gdx_col_set_rng(gdx_col_handle h, const unsigned long long *counters)
{
    if (!h || !counters) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        hip_check(hipMemcpyAsync(c.rng_p.p, counters, sizeof(uint64_t)*c.n, hipMemcpyHostToDevice, c.stream), "write rng");
        if (c.state_loaded && c.n_active)
        {
            std::vector<int> sp(c.n_active);
            copy_d2h(sp.data(), c.slot_pidx.p, sizeof(int)*c.n_active, c.stream);
            hip_check(hipStreamSynchronize(c.stream), "read slots");
            std::vector<uint64_t> slot(c.n_active);
            for (int s = 0; s < c.n_active; ++s) slot[s] = counters[sp[s]];
            hip_check(hipMemcpyAsync(c.rng.p, slot.data(), sizeof(uint64_t)*c.n_active, hipMemcpyHostToDevice, c.stream), "write rng");
        }
        hip_check(hipStreamSynchronize(c.stream), "write rng");
    });
}

int gdx_col_get_graph(gdx_col_handle h, unsigned long long *rows)
{
    if (!h || !rows) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.geometry_valid, GDX_COL_ERR_STATE, "collision geometry has not been built");
        copy_d2h(rows, c.graph.p, sizeof(unsigned long long)*LOCAL_GROUPS, c.stream);
        hip_check(hipStreamSynchronize(c.stream), "read graph");
    });
}

int gdx_col_rate_pass(gdx_col_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(c.state_loaded, GDX_COL_ERR_STATE, "collision state has not been loaded");
        run_buffers &rb = h->rb;
        rb.size_run(LOCAL_GROUPS, 1);
        rb.reserve(LOCAL_GROUPS);
        rb.h_table.clear();
        for (int g = 0; g < LOCAL_GROUPS; ++g) rb.h_table.push_back(make_gref(c, g, rb, 0, 0));
        rb.table.alloc(LOCAL_GROUPS);
        hip_check(hipMemcpyAsync(rb.table.p, rb.h_table.data(), sizeof(gdx_gref)*LOCAL_GROUPS, hipMemcpyHostToDevice,
                                 c.stream), "group table");
        std::vector<phase_entry> e;
        for (int g = 0; g < LOCAL_GROUPS; ++g)
            if (c.group_cnt[g] > 0) e.push_back(phase_entry{g, c.group_beg[g], c.group_cnt[g], 0.0, -1});
        launch_reset_scratch(c.stream, rb.table.p, LOCAL_GROUPS);
        zero_ranges z;
        z.add(rb.raw_rate.p, sizeof(gdx_raw_rate)*RAW_BINS);
        z.add(rb.nchanged.p, sizeof(int)*CHANGED_LISTS);
        phase_wave pw = upload_phase(rb, e, c.stream, z, c.narrow_owners);
        launch_publish(c.stream, pw.w1024, pw.blocks1024, rb);
        if (c.model && c.rev_valid && pw.blocks1024) launch_mark_stale(c.stream, pw.w1024, rb);
        launch_bath_rate(c.stream, pw.wsub, pw.blocks_sub);
        launch_bath_rate_grouped(c.stream, pw.wlsub, pw.blocks_lsub);
        hip_check(hipStreamSynchronize(c.stream), "rating pass");
    });
}

int gdx_col_pair_rates(gdx_col_handle h, double *rates)
{
    if (!h || !rates) return GDX_COL_ERR_ARG;
    int rc = gdx_col_rate_pass(h);
    if (rc != GDX_COL_OK) return rc;
    return guard(&h->c, [&] {
        context &c = h->c;
        dbuf<double> tmp;
        tmp.alloc(size_t(c.n)*N_K);
        hip_check(hipMemsetAsync(tmp.p, 0, sizeof(double)*size_t(c.n)*N_K, c.stream), "zero pair rates");
        launch_pair_rates(c, make_gref(c, 0, h->rb, 0, 0), tmp.p);
        copy_d2h(rates, tmp.p, sizeof(double)*size_t(c.n)*N_K, c.stream);
        hip_check(hipStreamSynchronize(c.stream), "read pair rates");
    });
}

int gdx_col_majorant_sample(gdx_col_handle h, float *features, double *scale, double *rate)
{
    if (!h) return GDX_COL_ERR_ARG;
    int rc = gdx_col_rate_pass(h);
    if (rc != GDX_COL_OK) return rc;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(features && scale && rate, GDX_COL_ERR_ARG, "null sample output");
        hip_check(hipMemsetAsync(features, 0, sizeof(float)*GDX_COL_MAJORANT_FEATURES*size_t(c.n), c.stream), "samples");
        hip_check(hipMemsetAsync(scale, 0, sizeof(double)*size_t(c.n), c.stream), "samples");
        hip_check(hipMemsetAsync(rate, 0, sizeof(double)*size_t(c.n), c.stream), "samples");
        launch_sample(c, make_gref(c, 0, h->rb, 0, 0), features, scale, rate);
        hip_check(hipStreamSynchronize(c.stream), "majorant samples");
    });
}

int gdx_col_set_leaf_target(gdx_col_handle h, int leaf_target)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(leaf_target >= 1 && leaf_target <= 4096, GDX_COL_ERR_MORTON_ARGS, "leaf target must be in [1, 4096]");
        h->c.leaf_target = leaf_target;
        h->c.geometry_valid = h->c.state_loaded = false;
    });
}

int gdx_col_autotune(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        context &c = h->c;
        require(candidates && count > 0 && repeats > 0, GDX_COL_ERR_ARG, "invalid autotuning arguments");
        require(c.user.bound, GDX_COL_ERR_STATE, "no particle state is bound");
        for (int i = 0; i < count; ++i)
            require(candidates[i] >= 1 && candidates[i] <= 4096, GDX_COL_ERR_MORTON_ARGS, "leaf target must be in [1, 4096]");
        // the epoch reloads the bound state: a deferred state goes back to the bound arrays first
        if (c.state_loaded) store_state(c);
        hip_check(hipStreamSynchronize(c.stream), "autotune");
        int pick = c.leaf_target;
        double fastest = 0.0;
        for (int i = 0; i < count; ++i)
        {
            c.leaf_target = candidates[i];
            geometry_epoch(c);                         // first epoch sizes the buffers
            double t = 0.0;
            for (int r = 0; r < repeats; ++r)
            {
                hip_check(hipStreamSynchronize(c.stream), "autotune");
                auto t0 = std::chrono::steady_clock::now();
                geometry_epoch(c);
                hip_check(hipStreamSynchronize(c.stream), "autotune");
                double v = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                t = r ? std::min(t, v) : v;
            }
            if (ms) ms[i] = t;
            if (i == 0 || t < fastest) { fastest = t; pick = candidates[i]; }
        }
        c.leaf_target = pick;
        geometry_epoch(c);
        hip_check(hipStreamSynchronize(c.stream), "autotune");
        if (best) *best = pick;
    });
}

int gdx_col_set_topk_tile(gdx_col_handle h, int tile)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(tile == 0 || tile == 8 || tile == 16 || tile == 32, GDX_COL_ERR_ARG, "top-K tile must be 0, 8, 16 or 32");
        h->c.topk_tile = tile;
        h->c.geometry_valid = h->c.state_loaded = false;
    });
}

int gdx_col_autotune_tile(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(candidates && count > 0 && repeats > 0, GDX_COL_ERR_ARG, "invalid autotuning arguments");
        for (int i = 0; i < count; ++i)
            require(candidates[i] == 0 || candidates[i] == 8 || candidates[i] == 16 || candidates[i] == 32,
                    GDX_COL_ERR_ARG, "top-K tile must be 0, 8, 16 or 32");
        autotune_epoch(h->c, h->c.topk_tile, candidates, count, repeats, best, ms);
    });
}

int gdx_col_set_topk_block(gdx_col_handle h, int waves)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(waves == 1 || waves == 2 || waves == 4, GDX_COL_ERR_ARG, "top-K block must be 1, 2 or 4 wavefronts");
        h->c.topk_waves = waves;
        h->c.geometry_valid = h->c.state_loaded = false;
    });
}

int gdx_col_autotune_block(gdx_col_handle h, const int *candidates, int count, int repeats, int *best, double *ms)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(candidates && count > 0 && repeats > 0, GDX_COL_ERR_ARG, "invalid autotuning arguments");
        for (int i = 0; i < count; ++i)
            require(candidates[i] == 1 || candidates[i] == 2 || candidates[i] == 4, GDX_COL_ERR_ARG,
                    "top-K block must be 1, 2 or 4 wavefronts");
        autotune_epoch(h->c, h->c.topk_waves, candidates, count, repeats, best, ms);
    });
}

int gdx_col_set_profile(gdx_col_handle h, int on)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        if (on) h->c.flags |= GDX_COL_PROFILE;
        else h->c.flags &= ~GDX_COL_PROFILE;
    });
}

int gdx_col_set_bound_reuse(gdx_col_handle h, int on)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        h->c.bound_reuse = on ? 1 : 0;
        invalidate_bounds(h->c);
    });
}

int gdx_col_set_fixed_bath(gdx_col_handle h, double duration)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(std::isfinite(duration) && duration >= 0.0, GDX_COL_ERR_ARG, "fixed bath duration must be >= 0");
        h->c.fixed_bath = duration;
    });
}

int gdx_col_set_exact_walk(gdx_col_handle h, int on)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] { h->c.exact_walk = on ? 1 : 0; });
}

int gdx_col_set_narrow_waves(gdx_col_handle h, int owners)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(owners >= 0, GDX_COL_ERR_ARG, "narrow-wave owner limit must be >= 0");
        h->c.narrow_owners = owners;
    });
}

int gdx_col_use_majorant(gdx_col_handle h, gdx_col_majorant m, double safety, int min_owners)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(&h->c, [&] {
        require(!m || (std::isfinite(safety) && safety >= 1.0), GDX_COL_ERR_ARG, "majorant safety factor must be >= 1");
        require(!m || (long long)h->c.n <= MAJORANT_MAX_OWNERS, GDX_COL_ERR_ARG,
                "a swarm with a majorant holds at most 2^32 / N_K representatives");
        h->c.mlp = majorant_pack(m);
        h->c.model = m;
        invalidate_bounds(h->c);
        h->c.safety = m ? safety : 1.0;
        h->c.majorant_min_owners = min_owners > 0 ? min_owners : GDX_COL_MAJORANT_MIN_OWNERS;
    });
}

int gdx_col_majorant_max_owners(long long *out)
{
    if (!out) return GDX_COL_ERR_ARG;
    *out = MAJORANT_MAX_OWNERS;
    return GDX_COL_OK;
}

// ---- host controller and scheduler ----
int gdx_col_host_binmap(const int *raw_count, int raw_bins, int *raw_to_merged, int *merged)
{
    if (!raw_count || !raw_to_merged || !merged || raw_bins <= 0 || raw_bins % COL_BIN_S) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        std::vector<int> rc(raw_count, raw_count + raw_bins), map;
        *merged = build_binmap(rc, map);
        std::copy(map.begin(), map.end(), raw_to_merged);
    });
}

int gdx_col_host_choose(const double *bins6, int n, double remaining, double limit_scale, double *duration)
{
    if (!bins6 || !duration || n < 0) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        std::vector<rate_bin> b(n);
        for (int i = 0; i < n; ++i)
        {
            const double *v = bins6 + 6*i;
            b[i] = rate_bin{v[0], v[1], v[2], v[3], int(v[4]), int(v[5])};
        }
        *duration = choose_bath(b.data(), n, remaining, limit_scale);
    });
}

int gdx_col_host_finish(const double *audit15, int n, double *state4, double *result10)
{
    if (!audit15 || !state4 || !result10 || n < 0) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        std::vector<audit_bin> b(n);
        for (int i = 0; i < n; ++i)
        {
            const double *v = audit15 + 15*i;
            audit_bin &o = b[i];
            o.mass = v[0]; o.predicted_f = v[1]; o.predicted_e = v[2]; o.predicted_var_f = v[3];
            o.predicted_var_e = v[4]; o.predicted_g = v[5]; o.predicted_var_g = v[6]; o.maximum_weight = v[7];
            o.maximum_g_jump = v[8]; o.touched = v[9]; o.event_weight = v[10]; o.growth = v[11];
            o.end_mass = v[12]; o.owner_count = int(v[13]); o.invalid_count = int(v[14]);
        }
        bath_state st{state4[0], int(state4[1]), int(state4[2]), int(state4[3])};
        bath_result r = finish_bath(b.data(), n, st);
        state4[0] = st.limit_scale; state4[1] = st.activity_streak; state4[2] = st.quiet_streak;
        state4[3] = st.floor_streak;
        const double out[10] = {r.max_f, r.max_e, r.max_touched, r.max_events, r.max_g, r.max_g_upper, r.d_bath,
                                double(r.activity_overshoot), double(r.distribution_overshoot),
                                double(r.persistent_overshoot)};
        std::copy(out, out + 10, result10);
    });
}

int gdx_col_host_schedule(int n, double horizon, const double *requested, const unsigned int *edges, int max_waves,
                          unsigned long long *ticks, int *due_counts, int *levels, int *waves)
{
    if (n <= 0 || !requested || !edges || !ticks || !due_counts || !levels || !waves) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        std::vector<double> req(requested, requested + n);
        std::vector<uint32_t> e(edges, edges + size_t(n)*((n + 31) / 32));
        schedule s(horizon, req, e);
        for (int c = 0; c < n; ++c) levels[c] = s.level[c];
        std::vector<bool> passed(n, true);
        int w = 0;
        while (s.time() < s.end)
        {
            if (w >= max_waves) throw status_error(GDX_COL_ERR_ARG, "schedule exceeds max_waves");
            uint64_t tick = s.time();
            std::vector<int> groups = s.due(tick);
            if (tick != 0) s.adapt(tick, groups, req, passed);
            ticks[w] = tick;
            due_counts[w] = int(groups.size());
            ++w;
            s.advance(groups);
        }
        *waves = w;
    });
}

int gdx_col_host_schedule_adaptive(int n, double horizon, const double *requested, const unsigned int *edges,
                                   int max_waves, const double *wave_requested, const unsigned char *wave_passed,
                                   unsigned long long *ticks, int *due_counts, int *wave_levels, int *waves)
{
    if (n <= 0 || !requested || !edges || !wave_requested || !wave_passed || !ticks || !due_counts || !wave_levels
        || !waves)
        return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        std::vector<double> req(requested, requested + n);
        std::vector<uint32_t> e(edges, edges + size_t(n)*((n + 31) / 32));
        schedule s(horizon, req, e);
        std::vector<bool> passed(n, true);
        int w = 0;
        while (s.time() < s.end)
        {
            if (w >= max_waves) throw status_error(GDX_COL_ERR_ARG, "schedule exceeds max_waves");
            uint64_t tick = s.time();
            std::vector<int> groups = s.due(tick);
            if (tick != 0)
            {
                for (int c = 0; c < n; ++c)
                {
                    req[c] = wave_requested[size_t(w)*n + c];
                    passed[c] = wave_passed[size_t(w)*n + c] != 0;
                }
                s.adapt(tick, groups, req, passed);
            }
            ticks[w] = tick;
            due_counts[w] = int(groups.size());
            for (int c = 0; c < n; ++c) wave_levels[size_t(w)*n + c] = s.level[c];
            ++w;
            s.advance(groups);
        }
        *waves = w;
    });
}

// ---- stand-alone Morton index ----
int gdx_morton_build(const void *points4, const unsigned char *active, int n, float ox, float oy, float oz,
                     float width, int dim, int leaf_target, int max_level, gdx_morton_handle *out)
{
    if (!out || !points4) return GDX_COL_ERR_ARG;
    *out = nullptr;
    gdx_morton_ctx *m = nullptr;
    int rc = guard(nullptr, [&] {
        m = new gdx_morton_ctx();
        morton_build(m->ix, static_cast<const float4 *>(points4), active, n, ox, oy, oz, width, dim, leaf_target,
                     max_level, m->scratch, nullptr);
        m->overflow.alloc(1);
    });
    if (rc != GDX_COL_OK) { delete m; return rc; }
    *out = m;
    return GDX_COL_OK;
}

int gdx_morton_destroy(gdx_morton_handle h)
{
    if (!h) return GDX_COL_ERR_ARG;
    delete h;
    return GDX_COL_OK;
}

int gdx_morton_node_count(gdx_morton_handle h, int *nodes)
{
    if (!h || !nodes) return GDX_COL_ERR_ARG;
    *nodes = h->ix.node_count;
    return GDX_COL_OK;
}

int gdx_morton_get(gdx_morton_handle h, unsigned long long *keys, int *idx, float *lower3, float *width, int *begin,
                   int *count, int *child8, int *child_count)
{
    if (!h) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        const morton_index &ix = h->ix;
        if (keys) copy_d2h(keys, ix.keys.p, sizeof(uint64_t)*ix.n, nullptr);
        if (idx) copy_d2h(idx, ix.idx.p, sizeof(int)*ix.n, nullptr);
        std::vector<gdx_node> nodes(ix.node_count);
        copy_d2h(nodes.data(), ix.nodes.p, sizeof(gdx_node)*ix.node_count, nullptr);
        hip_check(hipStreamSynchronize(nullptr), "read index");
        for (int i = 0; i < ix.node_count; ++i)
        {
            const gdx_node &nd = nodes[i];
            if (lower3) { lower3[3*i] = nd.lx; lower3[3*i + 1] = nd.ly; lower3[3*i + 2] = nd.lz; }
            if (width) width[i] = nd.width;
            if (begin) begin[i] = nd.begin;
            if (count) count[i] = nd.count;
            if (child_count) child_count[i] = nd.child_count & 15;
            if (child8)
            {
                // children were allocated in code order; bits 8..15 hold their codes
                int mask = (nd.child_count >> 8) & 0xff, ch = 0;
                for (int k = 0; k < 8; ++k) child8[8*i + k] = (mask >> k & 1) ? nd.first_child + ch++ : -1;
            }
        }
    });
}

int gdx_morton_knn(gdx_morton_handle h, const void *queries4, int nq, int k, int *codes, float *d2)
{
    if (!h || !queries4 || !codes || !d2) return GDX_COL_ERR_ARG;
    return guard(nullptr, [&] {
        hip_check(hipMemset(h->overflow.p, 0, sizeof(unsigned int)), "zero overflow");
        knn_query(h->ix, static_cast<const float4 *>(queries4), nq, k, codes, d2, h->overflow.p, nullptr);
        unsigned int ovf = 0;
        copy_d2h(&ovf, h->overflow.p, sizeof(ovf), nullptr);
        hip_check(hipStreamSynchronize(nullptr), "knn");
        if (ovf) throw status_error(GDX_COL_ERR_MORTON_OVERFLOW, "Morton traversal stack overflow");
    });
}

}  // extern "C"
