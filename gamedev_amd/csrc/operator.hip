// Frozen-bath collision operator kernels on gfx950.
//
// One refresh wave of evolve_local_collisions (inc/swarm/_col_chain.cuh:291-485) runs three kernels
// over the due groups' contiguous owner slots:
//   k_publish    col_bath_init + col_comp_zero + col_size_scan (src/swarm/col_bath_init.cu,
//                col_comp_zero.cu, col_size_scan.cu): freeze the reservoir, record size extrema
//   k_bath_rate  col_size_bnds + col_bath_rate + col_count_bin + col_rate_bins (col_size_bnds.cu,
//                col_bath_rate.cu, col_count_bin.cu, col_rate_bins.cu): bath-start rate moments and
//                per-raw-bin controller sums
//   k_chain      col_skip_scan + col_chain_run + col_audit_bin (col_skip_scan.cu, col_chain_run.cu,
//                col_audit_bin.cu): screen, run every event chain to the bath end, audit sums
// Merged-bin sums are sums of raw-bin sums, so the host merges after the binmap is known.
// A wavefront segment of N_K lanes holds one owner, lane k evaluating neighbour slot k; the
// xor-butterfly sums equal the reference's generic upper-half fold (col_bath_rate.cu:105-118).
#include "gdx_internal.h"
#include "gdx_device.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace gdx {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / WAVE;
constexpr int PUBLISH_OWNERS = 256;
constexpr int RATE_OWNERS = 64;
constexpr int CHAIN_OWNERS = 256;
// events one owner may take in one bath: the reference allows COL_EVENT_CAP events per launch and
// 1e6 continuation launches (inc/swarm/_col_chain.cuh:381-384); exceeding this bound is status 7
constexpr int EVENT_LIMIT = 1 << 22;

__device__ __forceinline__ bool due_block(const gdx_due &due, int per_block, int &gi, int &s0, int &s1)
{
    int b = blockIdx.x;
    gi = 0;
    while (gi < due.n && b >= due.bstart[gi + 1]) ++gi;
    if (gi >= due.n) return false;
    int local = b - due.bstart[gi];
    s0 = due.beg[gi] + local*per_block;
    s1 = min(due.beg[gi] + due.cnt[gi], s0 + per_block);
    (void)per_block;
    return s0 < s1;
}

__global__ void k_reset_scratch(unsigned long long *scr_min, unsigned long long *scr_max)
{
    int g = threadIdx.x;
    if (g < LOCAL_GROUPS)
    {
        scr_min[g] = dbits(__builtin_inf());
        scr_max[g] = 0ull;
    }
}

__global__ __launch_bounds__(BLOCK) void k_publish(gdx_due due, const double2 *st, double2 *res,
                                                  unsigned long long *scr_min, unsigned long long *scr_max)
{
    int gi, s0, s1;
    if (!due_block(due, PUBLISH_OWNERS, gi, s0, s1)) return;
    const int g = due.gid[gi];
    int s = s0 + threadIdx.x;
    unsigned long long lo = ~0ull, hi = 0ull;
    if (s < s1)
    {
        double2 v = st[s];
        res[s] = v;
        // positive doubles order like their bit patterns (col_size_scan.cu:20-24)
        if (v.x > 0 && isfinite(v.x)) { lo = dbits(v.x); hi = lo; }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
    {
        lo = min(lo, __shfl_xor(lo, m, WAVE));
        hi = max(hi, __shfl_xor(hi, m, WAVE));
    }
    if (threadIdx.x % WAVE == 0 && hi != 0ull)
    {
        atomicMin(scr_min + g, lo);
        atomicMax(scr_max + g, hi);
    }
}

// one slot's pair term at owner size size_i (col_bath_rate.cu:30-58, col_chain_run.cu:136-165)
struct pair_term
{
    double rate, j1, j2, jmax, vrel, size_j;
};

__device__ __forceinline__ pair_term eval_pair(int j, const gdx_env &e, double hg_i, const double *hgs,
                                               const double2 *res, double size_i, double m, double lambda_0)
{
    pair_term t{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (j < 0) return t;
    double2 pj = res[j];
    double hg_j = 0.0;
    if constexpr (PARTNER_HG) hg_j = hgs[j];
    double vrel = 0.0;
    double r = pair_numer(e, hg_i, hg_j, size_i, pj.x, pj.y, lambda_0, vrel) / m;
    double mean = 0.0, second = 0.0, maximum = 0.0;
    r *= outcome_moments(size_i, pj.x, vrel >= V_FRAG, mean, second, maximum);
    t.rate = r;
    t.j1 = r*mean;
    t.j2 = r*second;
    t.jmax = r > 0.0 ? maximum : 0.0;
    t.vrel = vrel;
    t.size_j = pj.x;
    return t;
}

__global__ __launch_bounds__(BLOCK) void k_bath_rate(gdx_due due, const int *nbr, const double2 *st,
                                                    const double2 *res, const double *measure, const gdx_env *env,
                                                    const double *hgs, double lambda_0,
                                                    const unsigned long long *scr_min,
                                                    const unsigned long long *scr_max, double *lower, double *upper,
                                                    double4 *cached, gdx_raw_rate *raw)
{
    __shared__ double accd[COL_BIN_S][4];
    __shared__ int acci[COL_BIN_S][3];
    int gi, s0, s1;
    if (!due_block(due, RATE_OWNERS, gi, s0, s1)) return;
    const int g = due.gid[gi];
    // moving bounds from this wave's extrema (col_size_bnds.cu:12-21)
    double lo = lower[g], up = upper[g];
    unsigned long long mx = scr_max[g];
    if (mx != 0ull)
    {
        lo = 0.5*bitsd(scr_min[g]);
        up = 8.0*bitsd(mx);
        if (threadIdx.x == 0 && blockIdx.x == due.bstart[gi]) { lower[g] = lo; upper[g] = up; }
    }
    if (threadIdx.x < COL_BIN_S*4) accd[threadIdx.x / 4][threadIdx.x % 4] = 0.0;
    if (threadIdx.x < COL_BIN_S*3) acci[threadIdx.x / 3][threadIdx.x % 3] = 0;
    __syncthreads();
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const int seg = lane / SEG, kl = lane % SEG;
    for (int base = s0 + wave*NSEG; base < s1; base += WAVES*NSEG)
    {
        const int s = base + seg;
        const bool valid = s < s1;
        double2 me = valid ? st[s] : make_double2(0.0, 0.0);
        double m = valid ? measure[s] : 0.0;
        int j = (valid && m > 0.0) ? nbr[(long long)s*N_K + kl] : -1;
        gdx_env e{};
        double hg_i = 0.0;
        if constexpr (COAG_KERNEL == 3) { if (valid) e = env[s]; }
        if constexpr (PARTNER_HG) { if (valid) hg_i = hgs[s]; }
        pair_term t = eval_pair(j, e, hg_i, hgs, res, me.x, m, lambda_0);
        int ok = (isfinite(t.rate) && t.rate >= 0.0) ? 1 : 0;
        double R = seg_sum<SEG>(t.rate);
        double J1 = seg_sum<SEG>(t.j1);
        double J2 = seg_sum<SEG>(t.j2);
        double JM = seg_max<SEG>(t.jmax);
        int V = seg_and<SEG>(ok);
        if (valid && kl == 0)
        {
            cached[s] = make_double4(V ? R : -1.0, J1, J2, JM);
            int sb = size_bin(me.x, lo, up);
            atomicAdd(&acci[sb][0], 1);
            atomicAdd(&acci[sb][1], 1);
            double weight = me.y*grain_mass(me.x);
            if (!isfinite(weight) || !(weight > 0.0) || !isfinite(R) || R < 0.0 || !isfinite(J1) || J1 < 0.0
                || !isfinite(J2) || J2 < 0.0)
            {
                atomicAdd(&acci[sb][2], 1);
            }
            else
            {
                atomicAdd(&accd[sb][0], weight);
                atomicAdd(&accd[sb][1], weight*R);
                atomicAdd(&accd[sb][2], weight*J1);
                atomicAdd(&accd[sb][3], weight*J2);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < COL_BIN_S*4)
    {
        int b = threadIdx.x / 4, f = threadIdx.x % 4;
        double v = accd[b][f];
        if (v != 0.0) atomicAdd(&(&raw[g*COL_BIN_S + b].mass)[f], v);
    }
    else if (threadIdx.x >= 64 && threadIdx.x < 64 + COL_BIN_S*3)
    {
        int b = (threadIdx.x - 64) / 3, f = (threadIdx.x - 64) % 3;
        int v = acci[b][f];
        if (v != 0) atomicAdd(&(&raw[g*COL_BIN_S + b].count)[f], v);
    }
}

// per-block audit accumulators of col_audit_bin (col_audit_bin.cu:31-59)
struct audit_lds
{
    double d[COL_BIN_S][11];   // mass, pf, pe, pvf, pve, pg, pvg, touched, event_weight, growth, end_mass
    unsigned long long mx[COL_BIN_S][2];
    int cnt[COL_BIN_S][2];     // owner_count, invalid_count
};

// one owner's audit terms: start bin from the published size, end mass to the end bin
__device__ __forceinline__ void audit_owner(audit_lds &a, double2 old, double size_end, double rate0, double dur,
                                            double hazard, double j1, double j2, double jmax, int events, double lo,
                                            double up)
{
    int bs = size_bin(old.x, lo, up);
    int be = size_bin(size_end, lo, up);
    double weight = old.y*grain_mass(old.x);
    double probability = -expm1(-rate0*dur);
    double ev = static_cast<double>(events);
    atomicAdd(&a.cnt[bs][0], 1);
    if (!isfinite(weight) || !(weight > 0.0) || !isfinite(probability) || !isfinite(hazard) || hazard < 0.0
        || !isfinite(ev) || !isfinite(j1) || j1 < 0.0 || !isfinite(j2) || j2 < 0.0 || !isfinite(jmax) || jmax < 0.0
        || !isfinite(size_end) || !(size_end > 0.0))
    {
        atomicAdd(&a.cnt[bs][1], 1);
        return;
    }
    atomicAdd(&a.d[bs][0], weight);
    atomicAdd(&a.d[bs][1], weight*probability);
    atomicAdd(&a.d[bs][2], weight*hazard);
    atomicAdd(&a.d[bs][3], weight*weight*probability*(1.0 - probability));
    atomicAdd(&a.d[bs][4], weight*weight*hazard);
    atomicAdd(&a.d[bs][5], weight*j1);
    atomicAdd(&a.d[bs][6], weight*weight*j2);
    atomicMax(&a.mx[bs][0], dbits(weight));
    atomicMax(&a.mx[bs][1], dbits(weight*jmax));
    if (events > 0) atomicAdd(&a.d[bs][7], weight);
    atomicAdd(&a.d[bs][8], weight*ev);
    atomicAdd(&a.d[bs][9], weight*fabs(log(size_end / old.x)));
    atomicAdd(&a.d[be][10], weight);
}

__global__ __launch_bounds__(BLOCK) void k_chain(gdx_due due, const int *nbr, double2 *st, const double2 *res,
                                                const double4 *cached, const double *measure, const gdx_env *env,
                                                const double *hgs, const int *slot_pidx, uint64_t *rng,
                                                uint64_t seed, double lambda_0, const double *lower,
                                                const double *upper, unsigned long long *scr_min,
                                                unsigned long long *scr_max, gdx_raw_audit *raw, int *err,
                                                unsigned long long *counters)
{
    __shared__ audit_lds a;
    __shared__ int queue[WAVES][WAVE];
    __shared__ unsigned long long stat[3];
    int gi, s0, s1;
    if (!due_block(due, CHAIN_OWNERS, gi, s0, s1)) return;
    const int g = due.gid[gi];
    const double end = due.dt[gi];
    const double lo = lower[g], up = upper[g];
    for (int i = threadIdx.x; i < COL_BIN_S*11; i += BLOCK) (&a.d[0][0])[i] = 0.0;
    if (threadIdx.x < COL_BIN_S*2) { (&a.mx[0][0])[threadIdx.x] = 0ull; (&a.cnt[0][0])[threadIdx.x] = 0; }
    if (threadIdx.x < 3) stat[threadIdx.x] = 0ull;
    // this group's extrema scratch was consumed by k_bath_rate; clear it for the group's next refresh
    if (threadIdx.x == 0 && blockIdx.x == due.bstart[gi]) { scr_min[g] = dbits(__builtin_inf()); scr_max[g] = 0ull; }
    __syncthreads();

    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    unsigned long long n_events = 0, n_chain = 0, n_pairs = 0;
    int err_code = 0;

    // ---- screen: one lane per owner (col_skip_scan.cu:19-58) ----
    int nq;
    {
        const int s = s0 + wave*WAVE + lane;
        const bool valid = s < s1;
        bool chain = false;
        if (valid)
        {
            double4 c = cached[s];
            double m = measure[s];
            double hazard = 0.0, j1 = 0.0, j2 = 0.0, jmax = 0.0;
            bool done = false;
            if (!(m > 0.0))
            {
                done = true;
            }
            else if (isfinite(c.x) && c.x >= 0.0)
            {
                if (c.x == 0.0)
                {
                    done = true;
                }
                else
                {
                    uint64_t trial = rng[s];
                    double wait = -log(col_uniform(uint32_t(slot_pidx[s]), seed, trial)) / c.x;
                    double remaining = end;
                    if (isfinite(wait) && wait > 0.0 && wait >= remaining)
                    {
                        hazard = c.x*remaining;
                        j1 = c.y*remaining;
                        j2 = c.z*remaining;
                        jmax = fmax(0.0, c.w);
                        done = true;
                        rng[s] = trial;
                    }
                    // otherwise the draw stays uncommitted: the chain repeats the same deviate
                }
            }
            if (done)
            {
                double2 old = res[s];
                double rate0 = (isfinite(c.x) && c.x >= 0.0) ? c.x : 0.0;
                audit_owner(a, old, old.x, rate0, end, hazard, j1, j2, jmax, 0, lo, up);
            }
            else
            {
                chain = true;
            }
        }
        unsigned long long ball = __ballot(chain);
        if (chain) queue[wave][__popcll(ball & ((1ull << lane) - 1ull))] = s;
        nq = __popcll(ball);
        __builtin_amdgcn_wave_barrier();
    }

    // ---- event chains: one owner per segment, each run to its bath end (col_chain_run.cu:127-382) ----
    const int seg = lane / SEG, kl = lane % SEG, segbase = seg*SEG;
    int item = seg;
    int next = NSEG;
    bool loaded = false;
    int s = 0, events = 0;
    double size_i = 0, numr_i = 0, time = 0, hazard = 0, j1 = 0, j2 = 0, jmax = 0, m = 0, rate0 = 0, hg_i = 0;
    double2 old = make_double2(0.0, 0.0);
    uint64_t ctr = 0;
    uint32_t pidx = 0;
    gdx_env e{};
    while (true)
    {
        const bool has = item < nq;
        if (__ballot(has) == 0ull) break;
        bool fin = false;
        if (has)
        {
            if (!loaded)
            {
                s = queue[wave][item];
                double2 cur = st[s];
                size_i = cur.x;
                numr_i = cur.y;
                old = res[s];
                m = measure[s];
                ctr = rng[s];
                pidx = uint32_t(slot_pidx[s]);
                rate0 = cached[s].x;
                if constexpr (COAG_KERNEL == 3) e = env[s];
                if constexpr (PARTNER_HG) hg_i = hgs[s];
                time = hazard = j1 = j2 = jmax = 0.0;
                events = 0;
                loaded = true;
            }
            int j = nbr[(long long)s*N_K + kl];
            pair_term t = eval_pair(j, e, hg_i, hgs, res, size_i, m, lambda_0);
            int last = (!isfinite(t.rate) || t.rate < 0.0) ? -2 : (t.rate > 0.0 ? kl : -1);
            double R = seg_sum<SEG>(t.rate);
            double J1 = seg_sum<SEG>(t.j1);
            double J2 = seg_sum<SEG>(t.j2);
            double JM = seg_max<SEG>(t.jmax);
            int LAST = seg_last<SEG>(last);
            n_pairs += (kl == 0) ? N_K : 0;
            int code = 0;
            if (LAST == -2) { code = 1; fin = true; }
            else if (!isfinite(R) || R < 0.0) { code = 2; fin = true; }
            else if (R == 0.0) { time = end; fin = true; }
            else
            {
                double wait = -log(col_uniform(pidx, seed, ctr)) / R;
                double remaining = end - time;
                if (!isfinite(wait) || !(wait > 0.0)) { code = 3; fin = true; }
                else if (wait >= remaining)
                {
                    hazard += R*remaining;
                    j1 += J1*remaining;
                    j2 += J2*remaining;
                    jmax = fmax(jmax, JM);
                    time = end;
                    fin = true;
                }
                else
                {
                    double event_time = time + wait;
                    if (!(event_time > time)) { code = 4; fin = true; }
                    else
                    {
                        hazard += R*wait;
                        j1 += J1*wait;
                        j2 += J2*wait;
                        jmax = fmax(jmax, JM);
                        // partner by the cumulative walk in slot order, last positive slot on roundoff
                        double target = col_uniform(pidx, seed, ctr)*R;
                        double cumulative = 0.0;
                        int slot = -1;
                        for (int k = 0; k < N_K; ++k)
                        {
                            double rk = __shfl(t.rate, segbase + k, WAVE);
                            if (rk > 0.0)
                            {
                                cumulative += rk;
                                if (slot < 0 && cumulative > target) slot = k;
                            }
                        }
                        if (slot < 0) slot = LAST;
                        if (slot < 0) { code = 5; fin = true; }
                        else
                        {
                            double size_j = __shfl(t.size_j, segbase + slot, WAVE);
                            double vrel = __shfl(t.vrel, segbase + slot, WAVE);
                            double mass_before = numr_i*size_i*size_i*size_i;
                            bool high_speed = vrel >= V_FRAG;
                            double sample = high_speed ? col_uniform(pidx, seed, ctr) : 0.0;
                            double size_new = sample_outcome(size_i, size_j, high_speed, sample);
                            numr_i = mass_before / (size_new*size_new*size_new);
                            size_i = size_new;
                            time = event_time;
                            ++events;
                            double mass_after = numr_i*size_i*size_i*size_i;
                            double mass_scale = fmax(fabs(mass_before), fabs(mass_after));
                            if (!isfinite(size_i) || !isfinite(numr_i) || !(size_i > 0.0) || !(numr_i > 0.0)
                                || fabs(mass_after - mass_before) > 2.0e-12*mass_scale)
                            { code = 6; fin = true; }
                            else if (events >= EVENT_LIMIT) { code = 7; fin = true; }
                        }
                    }
                }
            }
            if (fin)
            {
                if (kl == 0)
                {
                    if (events > 0) st[s] = make_double2(size_i, numr_i);
                    rng[s] = ctr;
                    err_code = max(err_code, code);
                    if (!code) audit_owner(a, old, size_i, rate0, end, hazard, j1, j2, jmax, events, lo, up);
                    n_events += events;
                    n_chain += 1;
                }
                loaded = false;
            }
        }
        // finished segments take the next queued owners in order
        unsigned long long fb = __ballot(has && fin && kl == 0);
        if (has && fin) item = next + __popcll(fb & ((1ull << segbase) - 1ull));
        next += __popcll(fb);
    }

    // ---- flush block sums ----
    if (err_code) atomicMax(err, err_code);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
        n_events += __shfl_xor(n_events, o, WAVE);
        n_chain += __shfl_xor(n_chain, o, WAVE);
        n_pairs += __shfl_xor(n_pairs, o, WAVE);
    }
    if (lane == 0)
    {
        atomicAdd(&stat[0], n_events);
        atomicAdd(&stat[1], n_chain);
        atomicAdd(&stat[2], n_pairs);
    }
    __syncthreads();
    gdx_raw_audit *dst = raw + g*COL_BIN_S;
    for (int i = threadIdx.x; i < COL_BIN_S*11; i += BLOCK)
    {
        int b = i / 11, f = i % 11;
        double v = a.d[b][f];
        if (v == 0.0) continue;
        double *base = (f < 7) ? &dst[b].mass + f : &dst[b].touched + (f - 7);
        atomicAdd(base, v);
    }
    if (threadIdx.x < COL_BIN_S*2)
    {
        int b = threadIdx.x / 2, f = threadIdx.x % 2;
        unsigned long long v = a.mx[b][f];
        if (v) atomicMax(f == 0 ? &dst[b].maximum_weight_bits : &dst[b].maximum_g_jump_bits, v);
        int cnt = a.cnt[b][f];
        if (cnt) atomicAdd(f == 0 ? &dst[b].owner_count : &dst[b].invalid_count, cnt);
    }
    if (threadIdx.x < 3 && stat[threadIdx.x]) atomicAdd(&counters[threadIdx.x], stat[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_cached_to_particles(const double4 *cached, const int *slot_pidx, int n_active,
                                                            double *out)
{
    long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (s >= n_active) return;
    double4 c = cached[s];
    double *o = out + 4ll*slot_pidx[s];
    o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = c.w;
}

inline int blocks_for_due(gdx_due &due, int per_block)
{
    int b = 0;
    for (int i = 0; i < due.n; ++i)
    {
        due.bstart[i] = b;
        b += (due.cnt[i] + per_block - 1) / per_block;
    }
    due.bstart[due.n] = b;
    return b;
}

}  // namespace

void launch_reset_scratch(context &c)
{
    k_reset_scratch<<<1, 64, 0, c.stream>>>(c.scr_min.p, c.scr_max.p);
    hip_check(hipGetLastError(), "launch reset scratch");
}

void launch_publish(context &c, const gdx_due &due_in)
{
    gdx_due due = due_in;
    int nb = blocks_for_due(due, PUBLISH_OWNERS);
    if (!nb) return;
    k_publish<<<nb, BLOCK, 0, c.stream>>>(due, c.st.p, c.res.p, c.scr_min.p, c.scr_max.p);
    hip_check(hipGetLastError(), "launch publish");
}

void launch_bath_rate(context &c, const gdx_due &due_in, double lambda_0)
{
    gdx_due due = due_in;
    int nb = blocks_for_due(due, RATE_OWNERS);
    if (!nb) return;
    k_bath_rate<<<nb, BLOCK, 0, c.stream>>>(due, c.nbr.p, c.st.p, c.res.p, c.measure.p, c.env.p, c.hgs.p, lambda_0,
                                           c.scr_min.p, c.scr_max.p, c.lower.p, c.upper.p, c.cached.p,
                                           c.raw_rate.p);
    hip_check(hipGetLastError(), "launch bath rate");
}

void launch_chain(context &c, const gdx_due &due_in, double lambda_0)
{
    gdx_due due = due_in;
    int nb = blocks_for_due(due, CHAIN_OWNERS);
    if (!nb) return;
    k_chain<<<nb, BLOCK, 0, c.stream>>>(due, c.nbr.p, c.st.p, c.res.p, c.cached.p, c.measure.p, c.env.p, c.hgs.p,
                                       c.slot_pidx.p, c.rng.p, c.seed, lambda_0, c.lower.p, c.upper.p, c.scr_min.p,
                                       c.scr_max.p, c.raw_audit.p, c.err.p, c.counters.p);
    hip_check(hipGetLastError(), "launch chain");
}

void launch_gather_cached(context &c, double *out)
{
    if (c.n_active <= 0) return;
    k_cached_to_particles<<<(c.n_active + 255) / 256, 256, 0, c.stream>>>(c.cached.p, c.slot_pidx.p, c.n_active, out);
    hip_check(hipGetLastError(), "launch gather cached");
}

}  // namespace gdx
