// Frozen-bath collision operator kernels on gfx950.
//
// One refresh wave of evolve_local_collisions (inc/swarm/_col_chain.cuh:291-485) runs in two phases
// over the due groups' contiguous owner slots. Rate phase:
//   k_publish      col_bath_init + col_comp_zero + col_size_scan (src/swarm/col_bath_init.cu,
//                  col_comp_zero.cu, col_size_scan.cu): freeze the reservoir, record size extrema
//   k_bath_rate    col_size_bnds + col_bath_rate + col_count_bin + col_rate_bins (col_size_bnds.cu,
//                  col_bath_rate.cu, col_count_bin.cu, col_rate_bins.cu): bath-start rate moments and
//                  per-raw-bin controller sums
// Chain phase:
//   k_screen       col_skip_scan (col_skip_scan.cu): no-event owners finish and enter the audit sums,
//                  the others are queued per group
//   k_chain        col_chain_run (col_chain_run.cu): every queued owner's chain to its bath end
//   k_chain_audit  col_audit_bin (col_audit_bin.cu) for the chained owners
// Merged-bin sums are sums of raw-bin sums, so the host merges after the binmap is known.
// Every kernel reads its groups through a table of per-group array pointers, so one launch serves
// the due groups of several independent swarms (realizations) as well as those of one.
#include "gdx_internal.h"
#include "gdx_device.h"
#include "gdx_mlp.h"

#include <algorithm>
#include <cstdint>
#include <cstring>

#pragma clang fp contract(off)

namespace gdx {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / WAVE;
constexpr int PUBLISH_OWNERS = 1024;
// owners expected to take at least this many events in the bath (rate at the bath start times the
// bath) join the heavy end of their group's queue region, which the chain takes first: long chains
// then start with the launch instead of holding it open after the other owners are done
constexpr double HEAVY_EVENTS = 4.0;
// events one owner may take in one bath: the reference allows COL_EVENT_CAP events per launch and
// 1e6 continuation launches (inc/swarm/_col_chain.cuh:381-384); exceeding this bound is status 7
constexpr int EVENT_LIMIT = 1 << 22;

// the due group of this block: last entry with bstart <= block (binary search, block-uniform)
__device__ __forceinline__ bool due_block(const gdx_wave &w, int per_block, int &gi, int &s0, int &s1)
{
    const int b = xcd_block();
    if (b >= w.bstart[w.n]) return false;
    int lo = 0, hi = w.n - 1;
    while (lo < hi)
    {
        int mid = (lo + hi + 1) >> 1;
        if (w.bstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    gi = lo;
    int local = b - w.bstart[gi];
    s0 = w.beg[gi] + local*per_block;
    s1 = min(w.beg[gi] + w.cnt[gi], s0 + per_block);
    return s0 < s1;
}

// owners of due group gi rated with the exact K-sum in this bath: all of a plain swarm's, one run of
// THIN_RUN slots in the swarm's thin_stride runs of a learned swarm's (run sub, sub + stride, ...)
__device__ __forceinline__ bool exact_owner(const gdx_wave &w, int gi, int s)
{
    if (!w.learned) return true;
    const int sub = w.sub[gi];
    return sub < 0 || ((s - w.beg[gi]) / THIN_RUN) % w.table[w.ref[gi]].thin_stride == sub;
}

// wavefront reductions of a double through DPP lane moves (no LDS): quad exchanges, row shifts by 4
// and 8 leave each 16-lane row's total in its lane 15, row broadcasts 15 and 31 carry them to lane
// 63. Disabled or out-of-row sources contribute 0, so the inputs must be >= 0 for wave_max. Every
// lane of the wavefront must be active; the fixed tree makes the result deterministic.
template <int CTRL, int ROWS, bool BOUND>
__device__ __forceinline__ double dpp_move(double v)
{
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, BOUND);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, BOUND);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63(double v)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0xb1, 0xf, true>(v);      // quad_perm [1,0,3,2]
    v += dpp_move<0x4e, 0xf, true>(v);      // quad_perm [2,3,0,1]
    v += dpp_move<0x114, 0xf, true>(v);     // row_shr:4
    v += dpp_move<0x118, 0xf, true>(v);     // row_shr:8
    v += dpp_move<0x142, 0xa, false>(v);    // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc, false>(v);    // row_bcast:31 into rows 2 and 3
    return lane63(v);
}
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_move<0xb1, 0xf, true>(v));
    v = fmax(v, dpp_move<0x4e, 0xf, true>(v));
    v = fmax(v, dpp_move<0x114, 0xf, true>(v));
    v = fmax(v, dpp_move<0x118, 0xf, true>(v));
    v = fmax(v, dpp_move<0x142, 0xa, false>(v));
    v = fmax(v, dpp_move<0x143, 0xc, false>(v));
    return lane63(v);
}

// one K-lane segment per owner whose lane t holds row position t (slot bitrev(t)): the reduction
// levels pair positions differing in bit 0, 1, 2, 3 (quad exchanges, row half-mirror, row mirror),
// then whole rows; in slots that is the reference's upper-half fold (slot j with j + K/2 first).
// Segments are whole 16-lane rows, and every lane of a segment that reduces must be active.
__device__ __forceinline__ double lane_value(double v, int l)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
template <bool MAX>
__device__ __forceinline__ double seg_fold(double v)
{
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dpp_move<0xb1, 0xf, true>(v));       // quad_perm [1,0,3,2]
    v = op(v, dpp_move<0x4e, 0xf, true>(v));       // quad_perm [2,3,0,1]
    v = op(v, dpp_move<0x141, 0xf, true>(v));      // row_half_mirror
    v = op(v, dpp_move<0x140, 0xf, true>(v));      // row_mirror
    if constexpr (SEG == 16)
    {
        return v;
    }
    else if constexpr (SEG == 32)
    {
        v = op(v, dpp_move<0x142, 0xa, false>(v)); // rows 1 and 3 take rows 0 and 2
        double a = lane_value(v, 31), b = lane_value(v, 63);
        return (threadIdx.x % WAVE) < 32 ? a : b;
    }
    else
    {
        v = op(v, dpp_move<0x142, 0xa, false>(v));
        v = op(v, dpp_move<0x143, 0xc, false>(v)); // rows 2 and 3 take lane 31
        return lane_value(v, 63);
    }
}
// inclusive prefix sums over a segment's lanes in lane order (lane t: lanes 0..t of its segment), as a
// tree: shifts by 1, 2, 4 and 8 within 16-lane rows (zeros shifted in), then row broadcasts 15 and 31
// carry whole rows; every lane of the segment must be active
__device__ __forceinline__ double seg_scan(double v)
{
    v += dpp_move<0x111, 0xf, true>(v);            // row_shr:1
    v += dpp_move<0x112, 0xf, true>(v);            // row_shr:2
    v += dpp_move<0x114, 0xf, true>(v);            // row_shr:4
    v += dpp_move<0x118, 0xf, true>(v);            // row_shr:8
    if constexpr (SEG >= 32) v += dpp_move<0x142, 0xa, false>(v);   // rows 1 and 3 take lanes 15 and 47
    if constexpr (SEG == 64) v += dpp_move<0x143, 0xc, false>(v);   // rows 2 and 3 take lane 31
    return v;
}
// segment maximum of a non-negative integer (0 fills disabled sources)
__device__ __forceinline__ int seg_max_u(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, true));
    if constexpr (SEG == 16)
    {
        return v;
    }
    else if constexpr (SEG == 32)
    {
        v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
        int a = __builtin_amdgcn_readlane(v, 31), b = __builtin_amdgcn_readlane(v, 63);
        return (threadIdx.x % WAVE) < 32 ? a : b;
    }
    else
    {
        v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
        v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
        return __builtin_amdgcn_readlane(v, 63);
    }
}

__global__ void k_reset_scratch(const gdx_gref *table, int ng)
{
    int g = blockIdx.x*blockDim.x + threadIdx.x;
    if (g < ng)
    {
        *table[g].scr_min = dbits(__builtin_inf());
        *table[g].scr_max = 0ull;
    }
}

// col_bath_init + col_comp_zero + col_size_scan (col_bath_init.cu:11-23, col_size_scan.cu:13-25)
__global__ __launch_bounds__(BLOCK) void k_publish(gdx_wave w, gdx_changed *changed, int cap, int *nchanged)
{
    __shared__ unsigned long long red[2][WAVES];
    __shared__ int list[PUBLISH_OWNERS];
    __shared__ double2 replaced[PUBLISH_OWNERS];
    __shared__ int count, base;
    int gi, s0, s1;
    if (!due_block(w, PUBLISH_OWNERS, gi, s0, s1)) return;
    const gdx_gref &G = w.table[w.ref[gi]];
    // the group's arrays, read from the table once (the loop's stores could otherwise alias them)
    const double2 *__restrict__ st = G.st;
    double2 *__restrict__ res = G.res;
    float2 *__restrict__ res32 = G.res32;
    unsigned char *__restrict__ dirty = G.dirty;
    const bool track = G.stale != nullptr;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    unsigned long long lo = ~0ull, hi = 0ull;
    // Only slots whose state changed since their last publish are written: a clean slot's published
    // records already equal its state (the state load writes both, and every state change marks the
    // slot dirty). Every round's record and flag in one round of loads, then the records replaced by the
    // dirty slots (clean slots read the block's first slot instead, one cache line), all before any
    // store (a flat pointer could alias the LDS list, which would keep each round's loads behind the
    // last store).
    constexpr int ROUNDS = PUBLISH_OWNERS / BLOCK;
    double2 vs[ROUNDS], rs[ROUNDS];
    unsigned char ds[ROUNDS];
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        const int sv = s < s1 ? s : s0;
        vs[k] = st[sv];
        ds[k] = dirty[sv];
    }
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        rs[k] = res[(s < s1 && ds[k]) ? s : s0];
    }
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        if (s >= s1) break;
        const double2 v = vs[k];
        const bool changed_here = ds[k];
        // positive doubles order like their bit patterns (col_size_scan.cu:20-24)
        if (v.x > 0 && isfinite(v.x)) { lo = min(lo, dbits(v.x)); hi = max(hi, dbits(v.x)); }
        // a changed record: the slot's own bound goes stale and every owner whose row holds it takes the
        // exact change of that pair rate (k_mark_stale, from the lists of changed slots and the records
        // they replaced)
        if (changed_here)
        {
            const double was_x = rs[k].x, was_y = rs[k].y;
            res[s] = v;
            res32[s] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
            dirty[s] = 0;
            if (track)
            {
                const int e = atomicAdd(&count, 1);
                list[e] = s;
                replaced[e] = make_double2(was_x, was_y);
            }
        }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
    {
        lo = min(lo, __shfl_xor(lo, m, WAVE));
        hi = max(hi, __shfl_xor(hi, m, WAVE));
    }
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (lane == 0) { red[0][wave] = lo; red[1][wave] = hi; }
    __syncthreads();
    // the block's changed-slot list, shared with blocks of the same XCD
    const int li = int(blockIdx.x) % CHANGED_LISTS;
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < WAVES; ++w) { lo = min(lo, red[0][w]); hi = max(hi, red[1][w]); }
        if (hi != 0ull)
        {
            atomicMin(G.scr_min, lo);
            atomicMax(G.scr_max, hi);
        }
        base = count ? atomicAdd(nchanged + li, count) : 0;
    }
    __syncthreads();
    gdx_changed *out = changed + (long long)li*cap + base;
    for (int i = threadIdx.x; i < count; i += BLOCK)
        out[i] = gdx_changed{w.ref[gi], list[i], replaced[i].x, replaced[i].y};
}

// bound upkeep after a publish: the changed slot's own bound goes stale (its owner is evaluated again),
// and every owner whose row holds the slot and whose bound is kept adds the increase of that one pair
// rate, old record to new (decreases are dropped). The pair rate is the exact FP64 term of the rate pass
// for the owner's current state, so a kept bound that held at its evaluation still holds: R rises by at
// most the sum of its increments. A wavefront takes MARK_SLOTS changed slots at a time and spreads the
// (slot, reader) pairs of all of them over its lanes, so that reader lists of uneven length keep the
// lanes busy.
constexpr int MARK_SLOTS = 16;

// a listed slot as the pair loop reads it: its swarm's arrays (the same in every group entry of the
// swarm) and its fresh and replaced records, gathered once by the slot's lane
struct mark_slot
{
    const int *rev;
    const double *measure;
    const unsigned char *stale;
    const gdx_env *env;
    const double2 *st;
    const double *hgs;
    double *lam_add;
    double lambda_0;
    double size_new, numr_new, hg, size_old, numr_old;
};

__global__ __launch_bounds__(BLOCK) void k_mark_stale(const gdx_gref *table, const gdx_changed *changed, int cap,
                                                     const int *nchanged)
{
    __shared__ int incl[WAVES][MARK_SLOTS];
    __shared__ unsigned first[WAVES][MARK_SLOTS];
    __shared__ mark_slot info[WAVES][MARK_SLOTS];
    // the publish's lists, taken one after the other: entry v of the whole is entry v - start[l] of list l
    int start[CHANGED_LISTS + 1];
    start[0] = 0;
#pragma unroll
    for (int l = 0; l < CHANGED_LISTS; ++l) start[l + 1] = start[l] + nchanged[l];
    const int n = start[CHANGED_LISTS];
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const int gw = blockIdx.x*WAVES + wv, nw = gridDim.x*WAVES;
    for (int v0 = gw*MARK_SLOTS; v0 < n; v0 += nw*MARK_SLOTS)
    {
        int cnt = 0;
        if (lane < MARK_SLOTS)
        {
            unsigned b = 0;
            mark_slot ms{};
            if (v0 + lane < n)
            {
                const int v = v0 + lane;
                int l = 0;
#pragma unroll
                for (int k = 1; k < CHANGED_LISTS; ++k) l = v >= start[k] ? k : l;
                int at = v;
#pragma unroll
                for (int k = 0; k < CHANGED_LISTS; ++k) at = l == k ? v - start[k] : at;
                const gdx_changed c = changed[(long long)l*cap + at];
                const gdx_gref &G = table[c.ref];
                G.stale[c.slot] = 1;
                if (G.bound_reuse)
                {
                    b = G.rev_beg[c.slot];
                    cnt = int(G.rev_end[c.slot] - b);
                }
                const partner fresh = load_partner(c.slot, G.res, G.hgs);
                ms = mark_slot{G.rev, G.measure, G.stale, G.env, G.st, G.hgs, G.lam_add, G.lambda_0,
                               fresh.size_j, fresh.numr_j, fresh.hg_j, c.size, c.numr};
            }
            info[wv][lane] = ms;
            first[wv][lane] = b;
        }
        // inclusive prefix of the reader counts over the MARK_SLOTS lanes (every lane shuffles)
#pragma unroll
        for (int d = 1; d < MARK_SLOTS; d <<= 1)
        {
            const int t = __shfl_up(cnt, d, WAVE);
            if (lane >= d && lane < MARK_SLOTS) cnt += t;
        }
        if (lane < MARK_SLOTS) incl[wv][lane] = cnt;
        const int total = __shfl(cnt, MARK_SLOTS - 1, WAVE);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int p = lane; p < total; p += WAVE)
        {
            int e = 0;
#pragma unroll
            for (int k = 0; k < MARK_SLOTS - 1; ++k) e += incl[wv][k] <= p;
            const mark_slot &I = info[wv][e];
            const int o = I.rev[first[wv][e] + unsigned(p - (e ? incl[wv][e - 1] : 0))];
            // the reader's record and flags in one round of loads
            const double m = I.measure[o];
            const bool stale = I.stale[o];
            const double size_o = I.st[o].x;
            gdx_env env{};
            double hg_o = 0.0;
            if constexpr (COAG_KERNEL == 3) env = I.env[o];
            if constexpr (PARTNER_HG) hg_o = I.hgs[o];
            if (!(m > 0.0) || stale) continue;      // a stale bound is evaluated again anyway
            const partner fresh{I.size_new, I.numr_new, I.hg, true};
            const partner before{I.size_old, I.numr_old, I.hg, true};
            const owner_terms t = make_owner(env, size_o, m, hg_o);
            const double up = eval_pair_rate(t, fresh, I.lambda_0).rate - eval_pair_rate(t, before, I.lambda_0).rate;
            if (up > 0.0 && isfinite(up)) atomicAdd(I.lam_add + o, up);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------------------------
// bath-start rates, one lane per owner (col_bath_rate.cu:13-129, col_count_bin.cu, col_rate_bins.cu)
// The lane streams its K slots from the bit-reversed row, four pair terms per chunk: positions
// (4c, 4c+1) and (4c+2, 4c+3) are slot pairs (j, j + K/2), the chunk sum is a level-2 partial of the
// reference's upper-half fold, and higher levels combine through a carry stack whose branches
// depend only on the chunk index (wavefront-uniform). The sums equal the fold bit for bit.
// ---------------------------------------------------------------------------------------------
struct fold3
{
    double r, a, b;
};
__device__ __forceinline__ fold3 add3(const fold3 &x, const fold3 &y) { return {x.r + y.r, x.a + y.a, x.b + y.b}; }

// exact bath-start sums of one owner per lane (col_bath_rate.cu:13-129): the row streamed in
// bit-reversed slot order as int4 chunks of four slots, indices one chunk ahead of the partner loads;
// chunk sums are level-2 partials of the reference's upper-half fold and higher levels combine through
// a carry stack whose branches depend only on the chunk index, so the sums equal the fold bit for bit.
// cat3 is the lane's LDS slot for the owner's catastrophic moments.
struct rate_sums
{
    fold3 total;
    double JM;
    bool V;
};
// MOMENTS false: the rates alone (eval_pair_rate, the same rate bit for bit), no jump moments
// chunks (optional): the rate partial of every four consecutive row positions, (t0 + t1) + (t2 + t3)
template <bool MOMENTS = true>
__device__ __forceinline__ rate_sums owner_rate_sums(const gdx_gref *GP, int s, bool own, double2 me, double m,
                                                     double *cat3, double *chunks = nullptr)
{
    const int *nbr = GP->nbr;
    const double2 *res = GP->res;
    const gdx_env *env = GP->env;
    const double *hgs = GP->hgs;
    const double lambda_0 = GP->lambda_0;
    gdx_env e{};
    double hg_i = 0.0;
    if constexpr (COAG_KERNEL == 3) { if (own) e = env[s]; }
    if constexpr (PARTNER_HG) { if (own) hg_i = hgs[s]; }
    const owner_terms t = make_owner(e, me.x, m, hg_i);
    // read only by lanes with a catastrophic pair: kept in LDS rather than in registers
    if constexpr (MOMENTS)
    {
        cat_moments c = catastrophic_moments(me.x);
        cat3[0] = c.mean;
        cat3[1] = c.second;
        cat3[2] = c.maximum;
    }
    const cat_moments_lds cm{cat3};
    auto eval = [&](const partner &q) {
        if constexpr (MOMENTS) return eval_pair_terms(t, cm, q, lambda_0);
        else return eval_pair_rate(t, q, lambda_0);
    };
    const bool live = own && m > 0.0;
    const int4 *row = reinterpret_cast<const int4 *>(nbr + (long long)(own ? s : 0)*N_K);

    fold3 stack[LOG2K];
    fold3 total{0.0, 0.0, 0.0};
    double JM = 0.0;
    bool V = true;
    // the neighbour indices run one chunk ahead of the partner loads
    int4 jj = live ? row[0] : make_int4(-1, -1, -1, -1);
    for (int c = 0; c < N_K / 4; ++c)
    {
        int4 jn = (live && c + 1 < N_K / 4) ? row[c + 1] : make_int4(-1, -1, -1, -1);
        // each term is folded as soon as it is formed: (t0 + t1) + (t2 + t3), the level-2 partial
        const partner q0 = load_partner(jj.x, res, hgs), q1 = load_partner(jj.y, res, hgs);
        const partner q2 = load_partner(jj.z, res, hgs), q3 = load_partner(jj.w, res, hgs);
        pair_term p = eval(q0);
        fold3 a{p.rate, p.j1, p.j2};
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q1);
        a = add3(a, {p.rate, p.j1, p.j2});
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q2);
        fold3 b{p.rate, p.j1, p.j2};
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q3);
        b = add3(b, {p.rate, p.j1, p.j2});
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        fold3 part = add3(a, b);
        if (chunks) chunks[c] = part.r;
        bool placed = false;
#pragma unroll
        for (int lv = 2; lv < LOG2K; ++lv)
        {
            if (!placed)
            {
                if (((c >> (lv - 2)) & 1) == 0) { stack[lv] = part; placed = true; }
                else part = add3(stack[lv], part);
            }
        }
        if (!placed) total = part;
        jj = jn;
    }
    return rate_sums{total, JM, V};
}

// owner_rate_sums over a group of G consecutive lanes: lane g of the group streams the row positions
// [g*N_K/G, (g+1)*N_K/G) with the same chunk partials and carry stack, and the lanes' partials combine
// by xor shuffles (lane g with g ^ 1, then g ^ 2, ...), which are the upper levels of the same fold: the
// sums equal owner_rate_sums bit for bit, for a G-th of the serial pair evaluations per lane. The
// group's lanes share `own`, and all of them must be active (the shuffles read the group's lanes);
// chunks (optional) receives the rate partials of the lane's own chunks.
template <int G, bool MOMENTS = true>
__device__ __forceinline__ rate_sums group_rate_sums(const gdx_gref *GP, int s, bool own, double2 me, double m, int g,
                                                     double *cat3, double *chunks = nullptr)
{
    constexpr int P = N_K / G;                 // positions per lane
    constexpr int CP = P / 4;                  // chunks per lane
    constexpr int LP = LOG2K - ((G == 1) ? 0 : (G == 2) ? 1 : (G == 4) ? 2 : 3);   // log2(P)
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "lane groups of 1, 2, 4 or 8");
    static_assert(P >= 4 && (1 << LP) == P, "at least one chunk of four positions per lane");
    const int *nbr = GP->nbr;
    const double2 *res = GP->res;
    const gdx_env *env = GP->env;
    const double *hgs = GP->hgs;
    const double lambda_0 = GP->lambda_0;
    gdx_env e{};
    double hg_i = 0.0;
    if constexpr (COAG_KERNEL == 3) { if (own) e = env[s]; }
    if constexpr (PARTNER_HG) { if (own) hg_i = hgs[s]; }
    const owner_terms t = make_owner(e, me.x, m, hg_i);
    if constexpr (MOMENTS)
    {
        cat_moments c = catastrophic_moments(me.x);
        cat3[0] = c.mean;
        cat3[1] = c.second;
        cat3[2] = c.maximum;
    }
    const cat_moments_lds cm{cat3};
    auto eval = [&](const partner &q) {
        if constexpr (MOMENTS) return eval_pair_terms(t, cm, q, lambda_0);
        else return eval_pair_rate(t, q, lambda_0);
    };
    const bool live = own && m > 0.0;
    const int4 *row = reinterpret_cast<const int4 *>(nbr + (long long)(own ? s : 0)*N_K) + g*CP;

    fold3 stack[LP > 2 ? LP : 3];
    fold3 total{0.0, 0.0, 0.0};
    double JM = 0.0;
    bool V = true;
    int4 jj = live ? row[0] : make_int4(-1, -1, -1, -1);
    for (int c = 0; c < CP; ++c)
    {
        int4 jn = (live && c + 1 < CP) ? row[c + 1] : make_int4(-1, -1, -1, -1);
        const partner q0 = load_partner(jj.x, res, hgs), q1 = load_partner(jj.y, res, hgs);
        const partner q2 = load_partner(jj.z, res, hgs), q3 = load_partner(jj.w, res, hgs);
        pair_term p = eval(q0);
        fold3 a{p.rate, p.j1, p.j2};
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q1);
        a = add3(a, {p.rate, p.j1, p.j2});
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q2);
        fold3 b{p.rate, p.j1, p.j2};
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        p = eval(q3);
        b = add3(b, {p.rate, p.j1, p.j2});
        V = V && isfinite(p.rate) && p.rate >= 0.0;
        JM = fmax(JM, p.jmax);
        fold3 part = add3(a, b);
        if (chunks) chunks[c] = part.r;
        bool placed = false;
#pragma unroll
        for (int lv = 2; lv < LP; ++lv)
        {
            if (!placed)
            {
                if (((c >> (lv - 2)) & 1) == 0) { stack[lv] = part; placed = true; }
                else part = add3(stack[lv], part);
            }
        }
        if (!placed) total = part;
        jj = jn;
    }
    // the levels above a lane's positions: partner lanes within the group
#pragma unroll
    for (int mask = 1; mask < G; mask <<= 1)
    {
        total.r += __shfl_xor(total.r, mask, WAVE);
        if constexpr (MOMENTS)
        {
            total.a += __shfl_xor(total.a, mask, WAVE);
            total.b += __shfl_xor(total.b, mask, WAVE);
            JM = fmax(JM, __shfl_xor(JM, mask, WAVE));
        }
        V = V && __shfl_xor(int(V), mask, WAVE);
    }
    (void)g;
    return rate_sums{total, JM, V};
}

// LANES = 1: the owners of plain swarms' groups, one lane each. LANES = RATE_LANES: a lane group per
// owner, for the exact subsample of learned swarms' groups and for plain groups in waves with few plain
// owners (in both cases too few wavefronts to hide the serial pair loop of one lane per owner; the
// group's lanes fold their positions in the same order, so the sums are the same bits). Each launch has
// its own block starts, so that a group is rated by exactly one of the two.
constexpr int RATE_LANES = 4;
static_assert(BLOCK / RATE_LANES == RATE_SUB_OWNERS, "host block starts of the lane-group launch");

template <int LANES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_bath_rate(gdx_wave w)
{
    constexpr bool GROUPED = LANES > 1;
    constexpr int OWNERS = BLOCK / LANES;
    __shared__ double accd[COL_BIN_S][4];     // mass, weighted rate, change, second
    __shared__ int acci[COL_BIN_S][3];        // count, owner count, invalid count
    // blocks cover the group's rated owners (every owner, or a learned swarm's subsample)
    const int b = xcd_block();
    if (b >= w.bstart[w.n]) return;
    int gi = 0;
    {
        int lo = 0, hi = w.n - 1;
        while (lo < hi)
        {
            int mid = (lo + hi + 1) >> 1;
            if (w.bstart[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        gi = lo;
    }
    const int lane = threadIdx.x % WAVE;
    const int g = lane % LANES;                         // lane within the owner's group
    const int i = (b - w.bstart[gi])*OWNERS + int(threadIdx.x) / LANES;
    const int sub = GROUPED ? w.sub[gi] : -1;
    const gdx_gref *GP = w.table + w.ref[gi];
    const double2 *st = GP->st;
    const double *measure = GP->measure;
    // the exact subsample of a learned swarm checks the bound of the same bath: block totals
    __shared__ unsigned checked[2];
    if (threadIdx.x < COL_BIN_S*4) (&accd[0][0])[threadIdx.x] = 0.0;
    if (threadIdx.x < COL_BIN_S*3) (&acci[0][0])[threadIdx.x] = 0;
    if (threadIdx.x < 2) checked[threadIdx.x] = 0u;
    __syncthreads();
    const int s = w.beg[gi] + (sub < 0 ? i : ((i / THIN_RUN)*GP->thin_stride + sub)*THIN_RUN + i % THIN_RUN);
    const bool own = s < w.beg[gi] + w.cnt[gi];
    double2 me = own ? st[s] : make_double2(1.0, 0.0);
    const double m = own ? measure[s] : 0.0;
    __shared__ double cats[BLOCK][3];
    const bool live = own && m > 0.0;
    rate_sums rs;
    if constexpr (LANES == 1) rs = owner_rate_sums(GP, s, own, me, m, &cats[threadIdx.x][0]);
    else rs = group_rate_sums<LANES>(GP, s, own, me, m, g, &cats[threadIdx.x][0]);
    // the owner's sums are on every lane of its group: its first lane writes and counts
    const bool lead = own && g == 0;
    const fold3 total = rs.total;
    const double JM = rs.JM;
    const bool V = rs.V;
    const double R = total.r, J1 = total.a, J2 = total.b;
    if (lead) GP->cached[s] = make_double4(V ? R : -1.0, J1, J2, JM);

    // moving bounds from this wave's extrema (col_size_bnds.cu:12-21); read after the pair loop so
    // they hold no registers during it
    const gdx_gref &G = *GP;
    double lo = *G.lower, up = *G.upper;
    unsigned long long mx = *G.scr_max;
    if (mx != 0ull)
    {
        lo = 0.5*bitsd(*G.scr_min);
        up = 8.0*bitsd(mx);
        if (threadIdx.x == 0 && xcd_block() == w.bstart[gi]) { *G.lower = lo; *G.upper = up; }
    }

    if constexpr (GROUPED)
    {
        // only a learned group's subsample has a bound to check
        const bool check = sub >= 0 && lead && live && V;
        const bool viol = check && !(R <= GP->safety*GP->lam[s] + GP->lam_add[s]);
        const unsigned nc = __popcll(__ballot(check)), nv = __popcll(__ballot(viol));
        if (lane == 0 && nc) atomicAdd(&checked[0], nc);
        if (lane == 0 && nv) atomicAdd(&checked[1], nv);
    }
    // controller-bin sums: each owner adds to the block's LDS bins (native FP64 LDS atomics)
    if (lead)
    {
        const int sb = size_bin(me.x, lo, up);
        const double weight = me.y*grain_mass(me.x);
        const bool bad = !isfinite(weight) || !(weight > 0.0) || !isfinite(R) || R < 0.0 || !isfinite(J1)
            || J1 < 0.0 || !isfinite(J2) || J2 < 0.0;
        atomicAdd(&acci[sb][0], 1);
        atomicAdd(&acci[sb][1], 1);
        if (bad)
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
    __syncthreads();
    if constexpr (GROUPED)
    {
        if (threadIdx.x == 255 && checked[0])
        {
            atomicAdd(G.counters + GDX_CTR_CHECKS, (unsigned long long)checked[0]);
            atomicAdd(G.wave_ctr + GDX_WAVE_RATE_CHECKS, (unsigned long long)checked[0]);
            if (checked[1])
            {
                atomicAdd(G.counters + GDX_CTR_VIOLATIONS, (unsigned long long)checked[1]);
                atomicAdd(G.wave_ctr + GDX_WAVE_RATE_VIOLATIONS, (unsigned long long)checked[1]);
            }
        }
    }
    if (threadIdx.x < COL_BIN_S*4)
    {
        int b = threadIdx.x / 4, f = threadIdx.x % 4;
        double v = accd[b][f];
        if (v != 0.0) atomicAdd(&(&G.raw_rate[b].mass)[f], v);
    }
    else if (threadIdx.x >= 64 && threadIdx.x < 64 + COL_BIN_S*3)
    {
        int b = (threadIdx.x - 64) / 3, f = (threadIdx.x - 64) % 3;
        int v = acci[b][f];
        if (v != 0) atomicAdd(&(&G.raw_rate[b].count)[f], v);
    }
}

// ---------------------------------------------------------------------------------------------
// learned majorant: features, bound per owner (no reference counterpart)
// ---------------------------------------------------------------------------------------------
// Kernel weight of one partner from its FP32 published record (the factors of _get_col_chain_rate
// that need no transcendental, _col_rates.cuh:167-223): N_j for the constant kernel, N_j (m_i + m_j)
// and N_j m_i m_j for the sum and product kernels, N_j (s_i + s_j)^2 for the physical one (over the
// vertical-overlap root in a vertically integrated disk). The scale S = sum over partners / measure
// (times lambda_0 for the synthetic kernels) leaves the network the relative speed and outcome factors.
__device__ __forceinline__ float partner_weight(float size_i, float mass_i, float2 q, float hg_i, float hg_j)
{
    const float sj = q.x, nj = q.y;
    constexpr float MASS = static_cast<float>(PI*RHO_0/6.0);
    if constexpr (COAG_KERNEL == 0) { (void)size_i; (void)mass_i; (void)hg_i; (void)hg_j; (void)sj; return nj; }
    else if constexpr (COAG_KERNEL == 1) { (void)hg_i; (void)hg_j; return nj*(mass_i + MASS*sj*sj*sj); }
    else if constexpr (COAG_KERNEL == 2) { (void)hg_i; (void)hg_j; return nj*mass_i*(MASS*sj*sj*sj); }
    else
    {
        (void)mass_i;
        const float a = size_i + sj;
        float w = nj*a*a;
        if constexpr (PARTNER_HG) w /= sqrtf(hg_i*hg_i + hg_j*hg_j);
        else { (void)hg_i; (void)hg_j; }
        return w;
    }
}

// features of owner s against its published partners (FP32):
//   0 ln s_i, 1 ln St_i (physical kernel), 2-5 partner-weighted mean, spread, min and max of
//   ln(s_j / s_i), 6 ln of the owner's radial distance and 7 its height over the gas scale height
//   (physical kernel); returns S (0 when no partner has weight)
__device__ double owner_features(const gdx_gref &G, int s, float *f)
{
    const double2 me = G.st[s];
    const double m = G.measure[s];
    const double size_i = me.x;
    const float sf = static_cast<float>(size_i), mass_i = static_cast<float>(grain_mass(size_i));
    // size ratios in log2 (the hardware logarithm; sizes are normal floats), converted to ln at the end
    const float ls = __builtin_amdgcn_logf(sf);
    float hg_i = 0.0f;
    if constexpr (PARTNER_HG) hg_i = static_cast<float>(G.hgs[s]);
    const int4 *row = reinterpret_cast<const int4 *>(G.nbr + (long long)s*N_K);
    float W = 0.0f, m1 = 0.0f, m2 = 0.0f;
    float dmin = __builtin_inff(), dmax = -__builtin_inff();
    // sixteen partners per step: their slots in four vector loads, then sixteen independent 8-byte
    // gathers of the FP32 published records
#pragma unroll 1
    for (int c = 0; c < N_K / 16; ++c)
    {
        int jj[16];
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int4 j4 = row[4*c + v];
            jj[4*v] = j4.x; jj[4*v + 1] = j4.y; jj[4*v + 2] = j4.z; jj[4*v + 3] = j4.w;
        }
        float2 q[16];
        float hg[16];
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            // empty slots read slot 0 and are masked below: sixteen loads without branches
            const int j = max(jj[t], 0);
            q[t] = G.res32[j];
            hg[t] = 0.0f;
            if constexpr (PARTNER_HG) hg[t] = static_cast<float>(G.hgs[j]);
        }
#pragma unroll
        for (int t = 0; t < 16; ++t)
        {
            const float w0 = partner_weight(sf, mass_i, q[t], hg_i, hg[t]);
            const bool ok = jj[t] >= 0 && w0 > 0.0f && isfinite(w0);
            const float w = ok ? w0 : 0.0f;
            const float d = ok ? __builtin_amdgcn_logf(q[t].x) - ls : 0.0f;
            W += w;
            m1 += w*d;
            m2 += w*d*d;
            dmin = ok ? fminf(dmin, d) : dmin;
            dmax = ok ? fmaxf(dmax, d) : dmax;
        }
    }
    constexpr float LN2 = 0.69314718f;
    for (int k = 0; k < mlp::F; ++k) f[k] = 0.0f;
    f[0] = ls*LN2;
    if (W > 0.0f)
    {
        const float mean = m1 / W;
        f[2] = mean*LN2;
        f[3] = sqrtf(fmaxf(0.0f, m2 / W - mean*mean))*LN2;
        f[4] = dmin*LN2;
        f[5] = dmax*LN2;
    }
    if constexpr (COAG_KERNEL == 3)
    {
        const gdx_env e = G.env[s];
        f[1] = logf(static_cast<float>(stokes_scaled(size_i, e.radial, e.strat)));
        f[6] = logf(static_cast<float>(e.radial));
        f[7] = static_cast<float>(e.Z*e.omega / e.cs);
    }
    if (!(m > 0.0) || !(W > 0.0f)) return 0.0;
    return (COAG_KERNEL == 3 ? double(W) : G.lambda_0*double(W)) / m;
}

// bounds of the due owners of learned swarms. An owner whose own state is unchanged since its last
// evaluation keeps its bound (stale flags: set by the publish that changes the owner's record, cleared
// here; partner changes raise kept bounds exactly, k_mark_stale). A block takes MAJ_SLOTS due slots of one
// group, lists its stale ones in LDS (a few percent of them once bounds are kept, so a block of 256 slots
// would leave its one evaluating wavefront mostly idle), and evaluates the list 64 owners per wavefront:
// features per lane, then the network over the wavefront's owners as four 16-owner tiles on the matrix
// cores.
__global__ __launch_bounds__(BLOCK) void k_majorant(gdx_wave w)
{
    __shared__ float feat[BLOCK][mlp::F + 1];
    __shared__ float ys[BLOCK];
    __shared__ int list[MAJ_SLOTS];
    __shared__ int wave_tot[WAVES];
    int gi, s0, s1;
    if (!due_block(w, MAJ_SLOTS, gi, s0, s1)) return;
    const gdx_gref &G = w.table[w.ref[gi]];
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    // the stale slots in slot order: each thread reads the flags of SPAN consecutive slots in one round
    // of loads (no LDS store in between, which a flat pointer could alias), then a block-wide prefix
    // of the per-thread counts places them in the list
    constexpr int SPAN = MAJ_SLOTS / BLOCK;
    const int t0 = s0 + int(threadIdx.x)*SPAN;
    const unsigned char *stale_flags = (G.bound_reuse && G.stale) ? G.stale : nullptr;
    unsigned char flag[SPAN];
#pragma unroll
    for (int k = 0; k < SPAN; ++k) flag[k] = (t0 + k < s1) ? (stale_flags ? stale_flags[t0 + k] : 1) : 0;
    int mine = 0;
#pragma unroll
    for (int k = 0; k < SPAN; ++k) mine += flag[k] ? 1 : 0;
    int incl = mine;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1)
    {
        const int v = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += v;
    }
    if (lane == WAVE - 1) wave_tot[wv] = incl;
    __syncthreads();
    int at = incl - mine, n = 0;
    for (int k = 0; k < WAVES; ++k)
    {
        if (k < wv) at += wave_tot[k];
        n += wave_tot[k];
    }
#pragma unroll
    for (int k = 0; k < SPAN; ++k)
        if (flag[k]) list[at++] = t0 + k;
    __syncthreads();
    if (threadIdx.x == 0 && n > 0)
        atomicAdd(G.counters + GDX_CTR_BOUND_EVALS + blockIdx.x % GDX_CTR_SPREAD, (unsigned long long)n);
    const mlp::pack *P = static_cast<const mlp::pack *>(G.mlp);
    // wavefront-uniform loop, no block barrier inside
    for (int base = wv*WAVE; base < n; base += BLOCK)
    {
        const bool has = base + lane < n;
        const int so = has ? list[base + lane] : s0;
        float f[mlp::F];
        double S = 0.0;
        if (has) S = owner_features(G, so, f);
        else for (int k = 0; k < mlp::F; ++k) f[k] = 0.0f;
        for (int k = 0; k < mlp::F; ++k) feat[threadIdx.x][k] = f[k];
        // the tiles read only this wavefront's rows
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int tt = 0; tt < WAVE / 16; ++tt)
        {
            const int row = wv*WAVE + tt*16 + (lane & 15);
            mlp::forward_state fs;
            float z = mlp::tile_forward(P, mlp::input_fragment(P, &feat[row][0], true), fs);
            if (lane < 16) ys[row] = mlp::softplus(z);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (has)
        {
            G.lam[so] = S*static_cast<double>(ys[threadIdx.x]);
            G.lam_add[so] = 0.0;
            if (G.stale) G.stale[so] = 0;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// training samples: features, scale and the cached exact rate of every active owner, scattered to
// particle order
__global__ __launch_bounds__(256) void k_sample(gdx_gref G, int n, float *feat, double *scale, double *rate)
{
    const int s = blockIdx.x*blockDim.x + threadIdx.x;
    if (s >= n) return;
    float f[mlp::F];
    double S = owner_features(G, s, f);
    const long long p = G.slot_pidx[s];
    for (int k = 0; k < mlp::F; ++k) feat[p*mlp::F + k] = f[k];
    scale[p] = S;
    rate[p] = G.cached[s].x;
}

// ---------------------------------------------------------------------------------------------
// screen, event chains and audit (col_skip_scan.cu:13-59, col_chain_run.cu:21-404, col_audit_bin.cu:13-60)
// ---------------------------------------------------------------------------------------------
struct audit_lds
{
    double d[COL_BIN_S][11];   // mass, pf, pe, pvf, pve, pg, pvg, touched, event_weight, growth, end_mass
    unsigned long long mx[COL_BIN_S][2];
    int cnt[COL_BIN_S][2];     // owner_count, invalid_count
};

__device__ __forceinline__ void clear_audit(audit_lds &a)
{
    for (int i = threadIdx.x; i < COL_BIN_S*11; i += BLOCK) (&a.d[0][0])[i] = 0.0;
    if (threadIdx.x < COL_BIN_S*2) { (&a.mx[0][0])[threadIdx.x] = 0ull; (&a.cnt[0][0])[threadIdx.x] = 0; }
}

__device__ __forceinline__ void flush_audit(const audit_lds &a, gdx_raw_audit *dst)
{
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
}

// audit terms of one owner per lane (col_audit_bin.cu:31-59): every field but the end mass goes to the
// start bin bs, the end mass to the end bin be. Each lane adds its terms to the block's LDS bins with
// native FP64 LDS atomics (lanes of one bin meet on one address and the LDS serialises them). Without
// events (EVENTS false) the end bin is the start bin and the event fields are zero.
template <bool EVENTS>
__device__ __forceinline__ void audit_wave(audit_lds &a, bool take, int bs, int be, double weight,
                                           double probability, double hazard, double j1, double j2, double jmax,
                                           double ev, double growth, bool bad)
{
    if (!take) return;
    atomicAdd(&a.cnt[bs][0], 1);
    if (bad)
    {
        atomicAdd(&a.cnt[bs][1], 1);
        return;
    }
    const double w = weight;
    atomicAdd(&a.d[bs][0], w);
    atomicAdd(&a.d[bs][1], w*probability);
    atomicAdd(&a.d[bs][2], w*hazard);
    atomicAdd(&a.d[bs][3], w*w*probability*(1.0 - probability));
    atomicAdd(&a.d[bs][4], w*w*hazard);
    atomicAdd(&a.d[bs][5], w*j1);
    atomicAdd(&a.d[bs][6], w*w*j2);
    if (w > 0.0) atomicMax(&a.mx[bs][0], dbits(w));
    const double x1 = w*jmax;
    if (x1 > 0.0) atomicMax(&a.mx[bs][1], dbits(x1));
    if constexpr (EVENTS)
    {
        if (ev > 0.0) atomicAdd(&a.d[bs][7], w);
        atomicAdd(&a.d[bs][8], w*ev);
        atomicAdd(&a.d[bs][9], w*growth);
        atomicAdd(&a.d[be][10], w);
    }
    else
    {
        (void)be; (void)ev; (void)growth;
        atomicAdd(&a.d[bs][10], w);
    }
}

// screen (col_skip_scan.cu:13-59): one lane per owner, ROUNDS owners per lane. In large waves a block
// covers SCREEN_OWNERS slots (SCREEN_ROUNDS per lane), so that the block's fixed latencies - group
// search, counter atomics, audit flush - are paid once per 1024 owners; in waves with few owners one
// round per lane spreads them over four times the blocks. Owners with no event in the bath finish here
// and enter the audit sums; the rest are appended to their group's queue region [beg, beg + cnt).
constexpr int SCREEN_ROUNDS = 4;
constexpr int SCREEN_OWNERS = BLOCK*SCREEN_ROUNDS;
static_assert(CHANGED_LISTS % XCDS == 0, "every changed-slot list belongs to the blocks of one XCD");
static_assert(SCREEN_OWNERS == PUBLISH_OWNERS, "the screen and the publish share the 1024-owner block starts");

template <int ROUNDS>
__global__ __launch_bounds__(BLOCK) void k_screen(gdx_wave w)
{
    __shared__ audit_lds a;
    __shared__ int block_tot[3], block_base[3];
    int gi, s0, s1;
    if (!due_block(w, BLOCK*ROUNDS, gi, s0, s1)) return;
    const gdx_gref &G = w.table[w.ref[gi]];
    // the group's arrays as distinct pointers, so that the rounds' loads can run ahead of the earlier
    // rounds' counter stores
    const double4 *__restrict__ cached = G.cached;
    const double *__restrict__ measure = G.measure;
    const double2 *__restrict__ res = G.res;
    const int *__restrict__ slot_pidx = G.slot_pidx;
    uint64_t *__restrict__ rng = G.rng;
    const double *__restrict__ lam = G.lam;
    const double *__restrict__ lam_add = G.lam_add;
    const double safety = G.safety;
    const uint64_t seed = G.seed;
    const double end = w.dt[gi];
    const double lo = *G.lower, up = *G.upper;
    clear_audit(a);
    if (threadIdx.x < 3) block_tot[threadIdx.x] = 0;
    // this group's extrema scratch was consumed by k_bath_rate; clear it for the group's next refresh
    if (threadIdx.x == 0 && xcd_block() == w.bstart[gi]) { *G.scr_min = dbits(__builtin_inf()); *G.scr_max = 0ull; }
    __syncthreads();

    const int lane = threadIdx.x % WAVE;
    const unsigned long long below = (1ull << lane) - 1ull;
    const bool learned = w.learned && w.sub[gi] >= 0;
    // per round: the owner's list (0 light queue, 1 candidates, 2 heavy queue, -1 none) and its offset
    // within the block's share of that list, and a candidate's {time, bound}
    int list[ROUNDS], off[ROUNDS];
    double2 cand[ROUNDS];
    unsigned n_viol = 0;
    // every round's inputs first, in one round of loads without branches (a branch around a load makes
    // the compiler wait for it there): an owner of the exact subsample (or of a plain swarm) reads its
    // cached sums and its record, a thinned owner its bound (lam, lam_add); the loads an owner does not
    // need read the block's first slot instead, one cache line for the whole block
    double4 in[ROUNDS];
    double2 bnd[ROUNDS], rec[ROUNDS];
    double mes[ROUNDS];
    uint64_t ctr0[ROUNDS];
    int pid[ROUNDS];
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        const int sv = s < s1 ? s : s0;
        const bool thinned = learned && !exact_owner(w, gi, sv);
        mes[k] = measure[sv];
        ctr0[k] = rng[sv];
        pid[k] = slot_pidx[sv];
        in[k] = cached[thinned ? s0 : sv];
        rec[k] = res[thinned ? s0 : sv];
        bnd[k] = make_double2(lam[thinned ? sv : s0], lam_add[thinned ? sv : s0]);
    }
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        const bool valid = s < s1;
        bool chain = false, done = false, heavy = false;
        // learned swarms: owners outside the exact subsample draw a candidate time from their bound
        // (thinning); the subsample follows the plain screen (its bound was checked by the rate pass)
        bool violation = false, candidate = false;
        double2 cs = make_double2(0.0, 0.0);
        double hazard = 0.0, j1 = 0.0, j2 = 0.0, jmax = 0.0, rate0 = 0.0;
        double2 old = make_double2(1.0, 0.0);
        if (valid && learned && !exact_owner(w, gi, s))
        {
            const bool thin = mes[k] > 0.0;
            const double L = thin ? safety*bnd[k].x + bnd[k].y : 0.0;
            if (thin && (!isfinite(L) || L < 0.0))
            {
                // unusable bound: the owner takes the plain chain from the bath start
                violation = chain = true;
            }
            else if (thin && L > 0.0)
            {
                uint64_t trial = ctr0[k];
                const double t1 = -log(col_uniform(uint32_t(pid[k]), seed, trial)) / L;
                rng[s] = trial;
                if (t1 < end)
                {
                    candidate = true;
                    cs = make_double2(t1, L);
                }
            }
        }
        else if (valid)
        {
            const double4 c = in[k];
            const double m = mes[k];
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
                    uint64_t trial = ctr0[k];
                    double wait = -log(col_uniform(uint32_t(pid[k]), seed, trial)) / c.x;
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
            chain = !done;
            heavy = chain && isfinite(c.x) && c.x*end >= HEAVY_EVENTS;
            if (done)
            {
                old = rec[k];
                rate0 = (isfinite(c.x) && c.x >= 0.0) ? c.x : 0.0;
            }
        }
        // audit terms of a finished owner only: in a learned swarm most wavefronts hold no owner of the
        // exact subsample and skip them (the bin, the mass and expm1 cost more than the draw)
        if (done)
        {
            const int bs = size_bin(old.x, lo, up);
            const double weight = old.y*grain_mass(old.x);
            const double probability = -expm1(-rate0*end);
            const bool bad = !isfinite(weight) || !(weight > 0.0) || !isfinite(probability) || !isfinite(hazard)
                || hazard < 0.0 || !isfinite(j1) || j1 < 0.0 || !isfinite(j2) || j2 < 0.0 || !isfinite(jmax)
                || jmax < 0.0 || !isfinite(old.x) || !(old.x > 0.0);
            audit_wave<false>(a, true, bs, bs, weight, probability, hazard, j1, j2, jmax, 0.0, 0.0, bad);
        }
        // owners with an event join the group's queue (heavy ones from the region's start, the others
        // from its end) and thinning candidates its candidate list: offsets within the block's share
        // from one LDS atomic per wavefront and list
        const unsigned long long bl = __ballot(chain && !heavy), bh = __ballot(heavy), cb = __ballot(candidate);
        int ol = 0, oc = 0, oh = 0;
        if (lane == 0)
        {
            ol = bl ? atomicAdd(&block_tot[0], __popcll(bl)) : 0;
            oc = cb ? atomicAdd(&block_tot[1], __popcll(cb)) : 0;
            oh = bh ? atomicAdd(&block_tot[2], __popcll(bh)) : 0;
        }
        ol = __shfl(ol, 0, WAVE);
        oc = __shfl(oc, 0, WAVE);
        oh = __shfl(oh, 0, WAVE);
        list[k] = heavy ? 2 : chain ? 0 : candidate ? 1 : -1;
        off[k] = heavy ? oh + __popcll(bh & below) : chain ? ol + __popcll(bl & below)
                                                           : oc + __popcll(cb & below);
        cand[k] = cs;
        n_viol += __popcll(__ballot(violation));
    }
    // one returning global atomic per block and list (the group counters are shared by hundreds of
    // blocks on every XCD)
    __syncthreads();
    if (threadIdx.x == 0) block_base[0] = block_tot[0] ? atomicAdd(G.qcount, block_tot[0]) : 0;
    if (threadIdx.x == WAVE) block_base[1] = block_tot[1] ? atomicAdd(G.ccount, block_tot[1]) : 0;
    if (threadIdx.x == 2*WAVE) block_base[2] = block_tot[2] ? atomicAdd(G.hcount, block_tot[2]) : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k)
    {
        const int s = s0 + k*BLOCK + int(threadIdx.x);
        if (list[k] == 0 || list[k] == 2)
        {
            const int pos = list[k] == 2 ? w.beg[gi] + block_base[2] + off[k]
                                         : w.beg[gi] + w.cnt[gi] - 1 - (block_base[0] + off[k]);
            G.queue[pos] = s;
            if (learned) G.qstart[pos] = make_double2(0.0, 0.0);
        }
        else if (list[k] == 1)
        {
            const int pos = w.beg[gi] + block_base[1] + off[k];
            G.cqueue[pos] = s;
            G.cstart[pos] = cand[k];
        }
    }
    // unusable bounds (rare) are counted here; candidates are counted by the chain, screened owners by
    // the host
    if (learned && n_viol && lane == 0)
    {
        atomicAdd(G.counters + GDX_CTR_CHECKS, (unsigned long long)n_viol);
        atomicAdd(G.wave_ctr + GDX_WAVE_CHECKS, (unsigned long long)n_viol);
        atomicAdd(G.counters + GDX_CTR_VIOLATIONS, (unsigned long long)n_viol);
        atomicAdd(G.wave_ctr + GDX_WAVE_VIOLATIONS, (unsigned long long)n_viol);
    }
    __syncthreads();
    flush_audit(a, G.raw_audit);
}

// thinning candidates of learned swarms, a group of CAND_LANES lanes per candidate (the K-sums spread
// over the group's lanes, the chosen chunk's four pair rates shared): the exact K-sum at the
// candidate time (the owner and its frozen partners are unchanged since the bath start, so this is the
// bath-start rate). Below the bound the candidate is an event with probability R / bound, and a
// rejected candidate continues on R from the candidate time. Above the bound the exact process is the
// bound's process plus an independent one of rate R - bound: the first event is the earlier of the
// candidate time and that process's first point, so the owner's event time is still exactly
// distributed. The owner's first event (at the candidate time when accepted, else at the end of its next
// wait if that falls inside the bath) is taken here, with the draws and operations of the chain: the
// partner walk over the same rates, the outcome, then the K-sum at the new size and the next wait.
// Owners whose next event still falls inside the bath enter the chain after that first event, and the
// chain repeats the pending draw; any invalid rate or state sends the owner to the chain from the
// candidate time with its draws uncommitted, and the chain reports it.
// lanes per candidate. One: groups of 2 or 4 lanes (the K-sums spread over the group, the chosen
// chunk's four rates shared) measured slower on the fiducial disk, since every lane of a group repeats
// the candidate's draws and decisions and the extra wavefronts do not fit on the CUs at once, while one
// lane per candidate already puts every wavefront of a launch in flight together.
constexpr int CAND_LANES = 1;
constexpr int CAND_OWNERS = BLOCK / CAND_LANES;
static_assert(CAND_LANES == 1 || CAND_LANES == 2 || CAND_LANES == 4, "the lanes share a chunk's four rates");

// per due group of a learned swarm: its candidate blocks (ceil(candidates / CAND_OWNERS)), exclusive
// prefix in vstart[0..n], the total in vstart[n]
__global__ __launch_bounds__(256) void k_cand_prefix(gdx_wave w, int *vstart)
{
    __shared__ int part[256];
    const int t = threadIdx.x;
    const int per = (w.n + 255) / 256;
    auto blocks_of = [&](int i) {
        return w.sub[i] < 0 ? 0 : (*w.table[w.ref[i]].ccount + CAND_OWNERS - 1) / CAND_OWNERS;
    };
    int sum = 0;
    for (int i = t*per; i < min(w.n, (t + 1)*per); ++i) sum += blocks_of(i);
    part[t] = sum;
    __syncthreads();
    if (t == 0)
    {
        int acc = 0;
        for (int k = 0; k < 256; ++k) { const int v = part[k]; part[k] = acc; acc += v; }
    }
    __syncthreads();
    int a = part[t];
    for (int i = t*per; i < min(w.n, (t + 1)*per); ++i)
    {
        vstart[i] = a;
        a += blocks_of(i);
    }
    if (t == 255) vstart[w.n] = a;
}

// a grid of CAND_GRID blocks walks the candidate blocks of every group (vstart from k_cand_prefix),
// each XCD's blocks over one contiguous range of them, so that no block is launched for the many due
// slots without a candidate
constexpr int CAND_GRID = 2048;

__global__ __launch_bounds__(BLOCK) void k_candidates(gdx_wave w, const int *vstart)
{
    constexpr int CHUNKS = N_K / 4;                    // chunk partials of a row
    constexpr int CP = CHUNKS / CAND_LANES;            // chunks per lane
    // candidates, exact checks, violations, accepted, pair evaluations beyond the first K-sum, events
    // (= owners) finished here
    __shared__ unsigned tally[6];
    __shared__ int block_off[2][WAVES], block_tot[2], block_base[2];
    __shared__ double chunk_rate[CHUNKS][CAND_OWNERS];
    const int total = vstart[w.n];
    const int per = (total + XCDS - 1) / XCDS, xcd = int(blockIdx.x) % XCDS;
    const int local = int(blockIdx.x) / XCDS, stride = int(gridDim.x) / XCDS;
    const int lane = threadIdx.x % WAVE;
    const int g = lane % CAND_LANES, cand = int(threadIdx.x) / CAND_LANES, gbase = lane - g;
    for (int v = local; v < per && xcd*per + v < total; v += stride)
    {
    const int vb = xcd*per + v;
    int gi = 0;
    {
        int lo = 0, hi = w.n - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (vstart[mid] <= vb) lo = mid;
            else hi = mid - 1;
        }
        gi = lo;
    }
    const gdx_gref *GP = w.table + w.ref[gi];
    const int beg = w.beg[gi];
    const int n_cand = *GP->ccount;
    const int s0 = beg + (vb - vstart[gi])*CAND_OWNERS;
    if (threadIdx.x < 6) tally[threadIdx.x] = 0u;
    if (threadIdx.x < 2) block_tot[threadIdx.x] = 0;
    __syncthreads();
    const int pos = s0 + cand;
    const bool own = pos - beg < n_cand;
    const int s = own ? GP->cqueue[pos] : beg;
    const double2 me = own ? GP->st[s] : make_double2(1.0, 0.0);
    const double m = own ? GP->measure[s] : 0.0;
    double chunks[CP];
    const rate_sums rs = group_rate_sums<CAND_LANES, false>(GP, s, own, me, m, g, nullptr, chunks);
    const double R = rs.total.r;
#pragma unroll
    for (int c = 0; c < CP; ++c) chunk_rate[g*CP + c][cand] = chunks[c];
    // the walk below reads the chunk partials of the group's other lanes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // every lane of a candidate's group runs the same draws and decisions; its first lane writes
    bool chain = false, check = false, viol = false, accepted = false, heavy = false, finished = false;
    unsigned extra_pairs = 0;
    double2 qs = make_double2(0.0, 0.0);
    if (own)
    {
        const double2 c = GP->cstart[pos];
        const double end = w.dt[gi];
        const uint32_t pidx = uint32_t(GP->slot_pidx[s]);
        const uint64_t seed = GP->seed;
        uint64_t ctr = GP->rng[s];
        if (!rs.V || !isfinite(R) || R < 0.0)
        {
            chain = true;                                  // the chain reports the invalid rate
        }
        else
        {
            check = true;
            bool event;
            double t_event = c.x;
            if (R > c.y)
            {
                // the bound is below the exact rate: the exact process is the bound's process plus an
                // independent one of rate R - bound, so the owner's first event is the earlier of the
                // candidate time and that process's first point (both inside the bath)
                viol = true;
                const double wx = -log(col_uniform(pidx, seed, ctr)) / (R - c.y);
                if (wx < c.x) t_event = wx;
                event = true;
            }
            else
            {
                accepted = col_uniform(pidx, seed, ctr)*c.y < R;
                event = accepted;
            }
            // the first event: at t_event, or after the next wait on R from the candidate time
            uint64_t k = ctr;                              // draws taken here, committed at the end
            if (event)
            {
                qs = make_double2(t_event, 1.0);
            }
            else if (R > 0.0)
            {
                const double wait = -log(col_uniform(pidx, seed, k)) / R;
                if (isfinite(wait) && wait > 0.0 && wait >= end - c.x) ctr = k;      // no event in the bath
                else
                {
                    qs = make_double2(c.x, 0.0);           // the chain repeats the draw
                    t_event = c.x + wait;
                    event = isfinite(wait) && wait > 0.0 && t_event > c.x;
                    chain = !event;
                }
            }
            if (event)
            {
                chain = true;                              // unless finished below
                // partner with probability rate / R: the chunk whose cumulative chunk partials (row
                // order) pass the target, then the cumulative walk over the chunk's four rates; the
                // last positive chunk and slot on roundoff. The chain walks the slots in order
                // instead; both pick a partner with probability rate / R (no reference counterpart).
                const double target = col_uniform(pidx, seed, k)*R;
                double before = 0.0;
                int cc = -1, last_c = -1;
                double last_before = 0.0;
                for (int c = 0; c < CHUNKS; ++c)
                {
                    const double v = chunk_rate[c][cand];
                    if (v > 0.0)
                    {
                        if (before + v > target) { cc = c; break; }
                        last_c = c;
                        last_before = before;
                        before += v;
                    }
                }
                if (cc < 0) { cc = last_c; before = last_before; }
                double size_j = 0.0, vrel = 0.0;
                int slot = -1;
                if (cc >= 0)
                {
                    // the chunk's four pair rates: positions 4 cc + g, 4 cc + g + CAND_LANES, ... on lane g
                    gdx_env e{};
                    double hg_i = 0.0;
                    if constexpr (COAG_KERNEL == 3) e = GP->env[s];
                    if constexpr (PARTNER_HG) hg_i = GP->hgs[s];
                    const owner_terms ot = make_owner(e, me.x, m, hg_i);
                    pair_term mine[4 / CAND_LANES];
#pragma unroll
                    for (int j = 0; j < 4 / CAND_LANES; ++j)
                    {
                        const partner q = load_partner(GP->nbr[(long long)s*N_K + 4*cc + g + j*CAND_LANES], GP->res,
                                                       GP->hgs);
                        mine[j] = eval_pair_rate(ot, q, GP->lambda_0);
                    }
                    double cumulative = before, last_size = 0.0, last_vrel = 0.0;
                    int last = -1;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                    {
                        const pair_term &src = mine[r / CAND_LANES];
                        const double rate = __shfl(src.rate, gbase + r % CAND_LANES, WAVE);
                        const double sj = __shfl(src.size_j, gbase + r % CAND_LANES, WAVE);
                        const double vr = __shfl(src.vrel, gbase + r % CAND_LANES, WAVE);
                        if (slot < 0 && rate > 0.0)
                        {
                            cumulative += rate;
                            last = r;
                            last_size = sj;
                            last_vrel = vr;
                            if (cumulative > target) { slot = r; size_j = sj; vrel = vr; }
                        }
                    }
                    extra_pairs += 4;
                    if (slot < 0) { slot = last; size_j = last_size; vrel = last_vrel; }
                }
                if (slot >= 0)
                {
                    const double size_i = me.x, numr_i = me.y;
                    const double mass_before = numr_i*size_i*size_i*size_i;
                    const bool high_speed = vrel >= V_FRAG;
                    const double sample = high_speed ? col_uniform(pidx, seed, k) : 0.0;
                    const double size_new = sample_outcome(size_i, size_j, high_speed, sample);
                    const double numr_new = mass_before / (size_new*size_new*size_new);
                    const double mass_after = numr_new*size_new*size_new*size_new;
                    const double mass_scale = fmax(fabs(mass_before), fabs(mass_after));
                    const bool ok = isfinite(size_new) && isfinite(numr_new) && size_new > 0.0 && numr_new > 0.0
                        && !(fabs(mass_after - mass_before) > 2.0e-12*mass_scale);
                    if (ok)
                    {
                        // the K-sum at the new size and the next wait from the event
                        const rate_sums r2 = group_rate_sums<CAND_LANES, false>(
                            GP, s, true, make_double2(size_new, numr_new), m, g, nullptr);
                        extra_pairs += N_K;
                        const double R2 = r2.total.r;
                        const double remaining = end - t_event;
                        bool done = false;
                        uint64_t k2 = k;
                        if (r2.V && isfinite(R2) && R2 >= 0.0)
                        {
                            if (R2 == 0.0) done = true;
                            else
                            {
                                const double wait2 = -log(col_uniform(pidx, seed, k2)) / R2;
                                done = isfinite(wait2) && wait2 > 0.0 && wait2 >= remaining;
                                if (!done) heavy = R2*remaining >= HEAVY_EVENTS;
                            }
                        }
                        if (g == 0)
                        {
                            if (GP->evw)
                            {
                                double log_mass;
                                const int cat = outcome_category(size_i, size_j, high_speed, sample, size_new, log_mass);
                                GP->evw[s].count[cat] += 1ull;
                                GP->evw[s].log_mass[cat] += log_mass;
                            }
                            GP->st[s] = make_double2(size_new, numr_new);
                            GP->dirty[s] = 1;
                        }
                        // finished: the wait draw is committed; else the chain continues after the event
                        // and repeats it
                        ctr = done ? k2 : k;
                        finished = done;
                        chain = !done;
                        qs = make_double2(t_event, 2.0);
                    }
                }
            }
        }
        if (g == 0) GP->rng[s] = ctr;
    }
    // queue appends (heavy owners from the region's start, the others from its end) by each candidate's
    // first lane: wavefront offsets from LDS, one returning global atomic per block and list
    const bool lead = g == 0;
    const unsigned long long bl = __ballot(lead && chain && !heavy), bh = __ballot(lead && chain && heavy);
    const int wv = threadIdx.x / WAVE;
    if (lane == 0)
    {
        block_off[0][wv] = bl ? atomicAdd(&block_tot[0], __popcll(bl)) : 0;
        block_off[1][wv] = bh ? atomicAdd(&block_tot[1], __popcll(bh)) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) block_base[0] = block_tot[0] ? atomicAdd(GP->qcount, block_tot[0]) : 0;
    if (threadIdx.x == WAVE) block_base[1] = block_tot[1] ? atomicAdd(GP->hcount, block_tot[1]) : 0;
    __syncthreads();
    if (lead && chain)
    {
        const unsigned long long below = (1ull << lane) - 1ull;
        const int q = heavy ? beg + block_base[1] + block_off[1][wv] + __popcll(bh & below)
                            : beg + w.cnt[gi] - 1 - (block_base[0] + block_off[0][wv] + __popcll(bl & below));
        GP->queue[q] = s;
        GP->qstart[q] = qs;
    }
    const unsigned nc = __popcll(__ballot(lead && own)), nk = __popcll(__ballot(lead && check));
    const unsigned nv = __popcll(__ballot(lead && viol)), na = __popcll(__ballot(lead && accepted));
    const unsigned nf = __popcll(__ballot(lead && finished));
    unsigned np = lead ? extra_pairs : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) np += __shfl_xor(np, o, WAVE);
    if (lane == 0)
    {
        if (nc) atomicAdd(&tally[0], nc);
        if (nk) atomicAdd(&tally[1], nk);
        if (nv) atomicAdd(&tally[2], nv);
        if (na) atomicAdd(&tally[3], na);
        if (np) atomicAdd(&tally[4], np);
        if (nf) atomicAdd(&tally[5], nf);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        if (tally[0] || tally[4])
        {
            atomicAdd(GP->counters + GDX_CTR_PAIRS, (unsigned long long)tally[0]*N_K + tally[4]);
            atomicAdd(GP->counters + GDX_CTR_CAND_PAIRS, (unsigned long long)tally[0]*N_K + tally[4]);
        }
        if (tally[0]) atomicAdd(GP->counters + GDX_CTR_CANDIDATES, (unsigned long long)tally[0]);
        if (tally[1])
        {
            atomicAdd(GP->counters + GDX_CTR_CHECKS, (unsigned long long)tally[1]);
            atomicAdd(GP->wave_ctr + GDX_WAVE_CHECKS, (unsigned long long)tally[1]);
        }
        if (tally[2])
        {
            atomicAdd(GP->counters + GDX_CTR_VIOLATIONS, (unsigned long long)tally[2]);
            atomicAdd(GP->wave_ctr + GDX_WAVE_VIOLATIONS, (unsigned long long)tally[2]);
        }
        if (tally[3]) atomicAdd(GP->counters + GDX_CTR_ACCEPTED, (unsigned long long)tally[3]);
        if (tally[5])
        {
            // owners finished here took one event each: counted as the chain counts its owners (one
            // launch block each, longest chain at least one event)
            atomicAdd(GP->counters + GDX_CTR_EVENTS, (unsigned long long)tally[5]);
            atomicAdd(GP->counters + GDX_CTR_CHAINED, (unsigned long long)tally[5]);
            atomicMax(GP->counters + GDX_CTR_LONGEST, 1ull);
            atomicMax(GP->wave_ctr + GDX_WAVE_LONGEST, 1ull);
            atomicAdd(GP->wave_ctr + GDX_WAVE_BLOCKS, (unsigned long long)tally[5]);
            atomicAdd(GP->wave_ctr + GDX_WAVE_CHAINED, (unsigned long long)tally[5]);
        }
    }
    // the next candidate block resets the block's shared counters and chunk partials
    __syncthreads();
    }
}

// exclusive prefixes of the due groups' heavy queue lengths (qprefix[0..n]) and light queue lengths
// (qprefix[n+1..2n+1]), one block
__global__ __launch_bounds__(256) void k_queue_prefix(gdx_wave w, int *qprefix)
{
    __shared__ int part[2][256];
    const int t = threadIdx.x;
    const int per = (w.n + 255) / 256;
    int sh = 0, sl = 0;
    for (int i = t*per; i < min(w.n, (t + 1)*per); ++i)
    {
        sh += *w.table[w.ref[i]].hcount;
        sl += *w.table[w.ref[i]].qcount;
    }
    part[0][t] = sh;
    part[1][t] = sl;
    __syncthreads();
    if (t < 2)
    {
        int acc = 0;
        for (int k = 0; k < 256; ++k) { int v = part[t][k]; part[t][k] = acc; acc += v; }
    }
    __syncthreads();
    int ah = part[0][t], al = part[1][t];
    int *lprefix = qprefix + w.n + 1;
    for (int i = t*per; i < min(w.n, (t + 1)*per); ++i)
    {
        qprefix[i] = ah;
        lprefix[i] = al;
        ah += *w.table[w.ref[i]].hcount;
        al += *w.table[w.ref[i]].qcount;
    }
    if (t == 255)
    {
        qprefix[w.n] = ah;
        lprefix[w.n] = al;
    }
}

// event chains (col_chain_run.cu:127-382): a persistent grid over the queued owners. Each K-lane segment
// of a wavefront runs one owner to its bath end, then takes the wavefront's next owner; the chain sums go
// to the owner's queue position for the audit pass. The owners are dealt out without atomics: heavy
// positions round-robin over the wavefronts (the long chains start together, one per wavefront), light
// positions as one contiguous range per wavefront. Work counters shared by the whole grid cost one
// same-address device atomic per refill, and those serialise: with thousands of wavefronts the counter,
// not the chains, set the launch's length.
constexpr int CHAIN_PREFIX_LDS = 4096;     // due groups whose queue prefix a chain block keeps in LDS

__global__ __launch_bounds__(BLOCK) void k_chain(gdx_wave w)
{
    __shared__ double walk[WAVES][NSEG][N_K];
    __shared__ int qp[2*(CHAIN_PREFIX_LDS + 1)];
    __shared__ unsigned long long end_key, end_sum[5];
    __shared__ unsigned long long *end_wave;
    __shared__ int lpool;                     // the block's next light position
    if (threadIdx.x == 0) end_key = 0ull;
    if (threadIdx.x < 5) end_sum[threadIdx.x] = 0ull;
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    // the queue prefixes of the due groups (heavy, then light), searched at every refill: LDS when
    // they fit. Positions [0, heavy) of the run are the heavy queues, group after group; the rest are
    // the light queues.
    const bool in_lds = w.n <= CHAIN_PREFIX_LDS;
    if (in_lds)
        for (int i = threadIdx.x; i < 2*(w.n + 1); i += BLOCK) qp[i] = w.qprefix[i];
    __syncthreads();
    const int *hprefix = in_lds ? qp : w.qprefix;
    const int *lprefix = hprefix + w.n + 1;
    const int heavy = hprefix[w.n];
    const int total = heavy + lprefix[w.n];

    // lane kl of a segment holds row position kl, i.e. slot bitrev(kl)
    const int seg = lane / SEG, kl = lane % SEG, segbase = seg*SEG, slot_k = bitrev_slot(kl);
    const unsigned long long seg_mask = (SEG == 64) ? ~0ull : ((1ull << SEG) - 1ull) << segbase;
    const unsigned long long seg_lead = (N_K == 64) ? 1ull : (N_K == 32) ? 0x0000000100000001ull
                                                                         : 0x0001000100010001ull;
    unsigned long long n_events = 0, n_chain = 0, n_pairs = 0, n_blocks = 0;
    int longest = 0;
    // accepted thinning candidate of a learned swarm: its first event happens at its start time
    bool first_event = false;
    bool heavy_owner = false;
    bool walk_exact = false;                  // the owner's swarm takes the sequential partner walk only
    gdx_event_work *evw = nullptr;            // the owner's swarm's event statistics (null: none kept)
    bool moments = true;
    bool active = false;
    int s = 0, pos = 0, gi = 0, events = 0;
    double end = 0.0, numr_i = 0.0, time = 0.0, hazard = 0.0, j1 = 0.0, j2 = 0.0, jmax = 0.0;
    uint64_t ctr = 0, seed = 0;
    uint32_t pidx = 0;
    double lambda_0 = 0.0;
    const gdx_gref *G = nullptr;              // the current owner's group
    unsigned long long *ctr_of = nullptr;     // the swarm whose counters the accumulators hold
    unsigned long long *wave_of = nullptr;
    owner_terms t{};
    partner q{0.0, 0.0, 0.0, false};
    partner_terms pt{0.0, 0.0, 0.0, 0.0};
    // the owner's next SEG draws, one per lane of the segment (lane kl: draw pool_base + kl, and its
    // -log for waits), computed side by side and taken in stream order: the same numbers as drawing
    // one at a time, for one Philox block and one logarithm per SEG draws
    double pool_u = 0.0, pool_nl = 0.0;
    uint64_t pool_base = 0;
    int pool_used = 0;
    auto refill = [&](uint64_t base) {
        pool_base = base;
        pool_used = 0;
        pool_u = col_uniform_at(pidx, seed, base + uint64_t(kl));
        pool_nl = -log(pool_u);
    };
    // counters are flushed when a segment moves to an owner of another swarm, and at the end
    auto flush = [&]() {
        if (kl == 0 && ctr_of)
        {
            if (n_events) atomicAdd(ctr_of + 0, n_events);
            if (n_chain) atomicAdd(ctr_of + 1, n_chain);
            if (n_pairs) atomicAdd(ctr_of + 2, n_pairs);
            if (longest) atomicMax(ctr_of + 3, (unsigned long long)longest);
            if (longest) atomicMax(wave_of + 0, (unsigned long long)longest);
            if (n_blocks) atomicAdd(wave_of + 1, n_blocks);
            if (n_chain) atomicAdd(wave_of + 2, n_chain);
        }
        n_events = n_chain = n_pairs = n_blocks = 0;
        longest = 0;
    };
    // Owners are dealt without global atomics. Heavy positions: while they are at most half the
    // wavefronts, heavy owner h runs alone on the first segment of wavefront h * (nwaves / heavy) (a light
    // owner sharing its wavefront would stall the long chain with its loads and branches at every
    // refill), else wavefront gw takes gw, gw + nwaves, ... on any free segment. Light positions: one
    // contiguous range per block, pulled by the block's free segments through an LDS counter, so that
    // the wavefronts of a block balance their light work among themselves (a wavefront that ran a solo
    // heavy owner joins in once its chain is done).
    const int nwaves = int(gridDim.x)*WAVES;
    const int gw = int(blockIdx.x)*WAVES + wave;
    const int light = total - heavy;
    const bool solo = heavy <= nwaves / 2;
    const int hstride = solo && heavy > 0 ? nwaves / heavy : nwaves;
    int hn = solo ? ((gw % hstride == 0 && gw / hstride < heavy) ? gw / hstride : heavy) : gw;
    const int hstep = solo ? heavy : nwaves;  // a solo wavefront's one heavy position, then none
    const int lb1 = int((long long)light*(blockIdx.x + 1) / gridDim.x);
    if (threadIdx.x == 0) lpool = int((long long)light*blockIdx.x / gridDim.x);
    __syncthreads();
    while (true)
    {
        unsigned long long need = __ballot(!active) & seg_lead;
        // a solo heavy owner runs alone: its wavefront takes nothing while it runs, and only that owner
        // (on the first segment) at its first refill
        if (solo && __ballot(active && heavy_owner)) need = 0ull;
        if (solo && hn < heavy) need &= 1ull;
        if (need)
        {
            // free segments, in lane order, take the next heavy positions, then the next light ones
            const int k = __popcll(need);
            const int th = min(k, hn < heavy ? (heavy - 1 - hn) / hstep + 1 : 0);
            int tl = 0, lc = 0;
            if (k > th)
            {
                if (lane == 0) lc = atomicAdd(&lpool, k - th);
                lc = __builtin_amdgcn_readfirstlane(lc);
                tl = max(0, min(k - th, lb1 - lc));
            }
            const int rank = __popcll(need & ((1ull << segbase) - 1ull));
            if (!active && rank < th + tl)
            {
                const bool hv = rank < th;
                const int v = hv ? hn + rank*hstep : lc + (rank - th);
                const int *prefix = hv ? hprefix : lprefix;
                // due group of run position v: last entry with prefix <= v
                int lo = 0, hi = w.n - 1;
                while (lo < hi)
                {
                    int mid = (lo + hi + 1) >> 1;
                    if (prefix[mid] <= v) lo = mid;
                    else hi = mid - 1;
                }
                gi = lo;
                G = w.table + w.ref[gi];
                if (G->counters != ctr_of) { flush(); ctr_of = G->counters; wave_of = G->wave_ctr; }
                // heavy queues grow from the group region's start, light ones from its end
                pos = hv ? w.beg[gi] + (v - prefix[gi]) : w.beg[gi] + w.cnt[gi] - 1 - (v - prefix[gi]);
                s = G->queue[pos];
                end = w.dt[gi];
                seed = G->seed;
                lambda_0 = G->lambda_0;
                double2 cur = G->st[s];
                numr_i = cur.y;
                ctr = G->rng[s];
                pidx = uint32_t(G->slot_pidx[s]);
                double hg_i = 0.0;
                gdx_env e{};
                if constexpr (COAG_KERNEL == 3) e = G->env[s];
                if constexpr (PARTNER_HG) hg_i = G->hgs[s];
                t = make_owner(e, cur.x, G->measure[s], hg_i);
                refill(ctr);
                moments = exact_owner(w, gi, s);
                // the partners are the published states: fixed for the whole chain, and so are their
                // partner-only factors
                q = load_partner(G->nbr[(long long)s*N_K + kl], G->res, G->hgs);
                if constexpr (HOLD_PARTNER) pt = make_partner_terms(t, q.size_j);
                time = hazard = j1 = j2 = jmax = 0.0;
                first_event = false;
                // a learned group's queue entries carry their start (plain groups of a mixed wave do not)
                events = 0;
                if (w.learned && w.sub[gi] >= 0)
                {
                    // 1: an accepted candidate, its event at the start time; 2: its first event already
                    // taken at the start time (the state read above is the state after it)
                    const double2 qs = G->qstart[pos];
                    time = qs.x;
                    first_event = qs.y == 1.0;
                    if (qs.y == 2.0) events = 1;
                }
                heavy_owner = hv;
                walk_exact = G->exact_walk != 0;
                // read once per owner: a load from the group table at every event would hold the
                // event's outcome until it returned
                evw = G->evw;
                active = true;
            }
            hn += th*hstep;
        }
        // no segment running after the refill: the wavefront's positions are done
        if (!__ballot(active)) break;
        // a wavefront running a heavy owner's chain issues ahead of its SIMD's other wavefronts: the
        // longest chains set the launch's length
        if (__ballot(active && heavy_owner)) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
        if (!active) continue;
        // an event takes at most three draws
        if (pool_used > SEG - 3) refill(pool_base + uint64_t(pool_used));
        // the event's next draws (a wait, the partner's target, a fragmentation sample), fetched from
        // their lanes of the segment while the pair terms are evaluated and taken below in stream order
        const double nx_nl = __shfl(pool_nl, segbase + pool_used, WAVE);
        const double nx_u0 = __shfl(pool_u, segbase + pool_used, WAVE);
        const double nx_u1 = __shfl(pool_u, segbase + pool_used + 1, WAVE);
        const double nx_u2 = __shfl(pool_u, segbase + pool_used + 2, WAVE);
        // owners outside a learned swarm's audited subsample keep no compensators: rates alone
        pair_term p = moments ? eval_pair_terms<cat_moments_deferred, HOLD_PARTNER>(t, cat_moments_deferred{}, q,
                                                                                    lambda_0, pt)
                              : eval_pair_rate<HOLD_PARTNER>(t, q, lambda_0, pt);
        // the catastrophic moments (owner-size logarithm and exponentials) only at events where some
        // partner of the owner is in the catastrophic regime, filled in as the pair evaluation forms them
        if constexpr (CAN_FRAGMENT)
        {
            const bool cat = moments && q.valid && catastrophic_pair(t, q.size_j, p.vrel >= V_FRAG);
            if (__ballot(cat) & seg_mask)
            {
                const cat_moments c = catastrophic_moments(t.size_i);
                if (cat)
                {
                    p.j1 = p.rate*c.mean;
                    p.j2 = p.rate*c.second;
                    p.jmax = p.rate > 0.0 ? c.maximum : 0.0;
                }
            }
        }
        const bool bad_rate = !isfinite(p.rate) || p.rate < 0.0;
        double R = seg_fold<false>(p.rate);
        // the jump moments only where a segment of the wavefront keeps compensators
        double J1 = 0.0, J2 = 0.0, JM = 0.0;
        if (__ballot(moments))
        {
            J1 = seg_fold<false>(p.j1);
            J2 = seg_fold<false>(p.j2);
            JM = seg_fold<true>(p.jmax);
        }
        // an invalid rate anywhere in the segment ends the chain (the last positive slot, the walk's
        // fallback on roundoff, is found only where the walk needs it)
        const bool any_bad = (__ballot(bad_rate) & seg_mask) != 0ull;
        n_pairs += (kl == 0) ? N_K : 0;
        bool fin = false;
        int code = 0;
        // accepted thinning candidate: the event happens at the start time, without a wait draw
        const bool now = first_event;
        first_event = false;
        if (any_bad) { code = 1; fin = true; }
        else if (!isfinite(R) || R < 0.0) { code = 2; fin = true; }
        else if (R == 0.0) { time = end; fin = true; }
        else
        {
            // partner by the cumulative walk in slot order, last positive slot on roundoff: the first
            // positive slot whose partial sum exceeds the target. Lane kl takes slot kl. Decided before
            // the wait is known (it needs only the rates, R and the target's draw), so that its lane
            // sums overlap the wait's division; it is used only at an event.
            const double target = (now ? nx_u0 : nx_u1)*R;
            walk[wave][seg][slot_k] = p.rate;
            __builtin_amdgcn_wave_barrier();
            const double own_rate = walk[wave][seg][kl];
            // The partials as a tree (seg_scan). Every rate here is finite and >= 0 (an invalid
            // one ended the chain above), and any order of adding at most 64 non-negative terms
            // lies within 64 u/(1 - 64 u) < 2^-46 relative of the exact sum, so the tree partial
            // and the walk's sequential one differ by less than 2^-45 of the partial: a tree
            // partial farther than 2^-44 (plus an absolute 2^-1000 for subnormal rounding) from
            // the target decides its slot's test exactly as the sequential partial does.
            const double a = seg_scan(own_rate);
            const bool above = a*(1.0 - 0x1p-44) - 0x1p-1000 > target;
            const bool below = a*(1.0 + 0x1p-44) + 0x1p-1000 <= target;
            unsigned long long hit = __ballot(own_rate > 0.0 && above) & seg_mask;
            if (walk_exact || (__ballot(own_rate > 0.0 && !above && !below) & seg_mask))
            {
                // a positive slot's partial within the margin (or the swarm asks for it): the
                // walk's own sequential partials. Lane kl adds the rates of slots 0..kl in slot
                // order and zeros for the later slots (adding a zero to the non-negative
                // cumulative changes no bit).
                const double2 *wr = reinterpret_cast<const double2 *>(&walk[wave][seg][0]);
                // the lane's slot, opaque here: each slot's lane test stays one compare inside
                // the sum instead of a mask kept across the event loop
                int upto = kl;
                asm volatile("" : "+v"(upto));
                double cumulative = 0.0;
#pragma unroll
                for (int k = 0; k < N_K / 2; ++k)
                {
                    const double2 rk = wr[k];
                    cumulative += 2*k <= upto ? rk.x : 0.0;
                    cumulative += 2*k + 1 <= upto ? rk.y : 0.0;
                }
                hit = __ballot(own_rate > 0.0 && cumulative > target) & seg_mask;
            }
            __builtin_amdgcn_wave_barrier();
            double wait = 0.0;
            if (!now)
            {
                wait = nx_nl / R;
                ++pool_used;
            }
            double remaining = end - time;
            if (!now && (!isfinite(wait) || !(wait > 0.0))) { code = 3; fin = true; }
            else if (!now && wait >= remaining)
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
                if (!now && !(event_time > time)) { code = 4; fin = true; }
                else
                {
                    hazard += R*wait;
                    j1 += J1*wait;
                    j2 += J2*wait;
                    jmax = fmax(jmax, JM);
                    // the target's draw (the partner walk above was decided before the wait)
                    ++pool_used;
                    int slot = hit ? int(__builtin_ctzll(hit)) - segbase : -1;
                    if (slot < 0) slot = seg_max_u(p.rate > 0.0 ? slot_k + 1 : 0) - 1;
                    if (slot < 0) { code = 5; fin = true; }
                    else
                    {
                        double size_i = t.size_i;
                        // slot k sits in lane bitrev(k) of the segment
                        double size_j = __shfl(p.size_j, segbase + bitrev_slot(slot), WAVE);
                        double vrel = __shfl(p.vrel, segbase + bitrev_slot(slot), WAVE);
                        double mass_before = numr_i*size_i*size_i*size_i;
                        bool high_speed = vrel >= V_FRAG;
                        double sample = 0.0;
                        if (high_speed)
                        {
                            sample = now ? nx_u1 : nx_u2;
                            ++pool_used;
                        }
                        double size_new = sample_outcome_owner(t, size_j, high_speed, sample);
                        if (kl == 0 && evw)
                        {
                            double log_mass;
                            int cat = outcome_category(size_i, size_j, high_speed, sample, size_new, log_mass);
                            evw[s].count[cat] += 1ull;
                            evw[s].log_mass[cat] += log_mass;
                        }
                        numr_i = mass_before / (size_new*size_new*size_new);
                        time = event_time;
                        ++events;
                        double mass_after = numr_i*size_new*size_new*size_new;
                        double mass_scale = fmax(fabs(mass_before), fabs(mass_after));
                        owner_resize(t, size_new);
                        if (!isfinite(size_new) || !isfinite(numr_i) || !(size_new > 0.0) || !(numr_i > 0.0)
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
                if (events > 0)
                {
                    G->st[s] = make_double2(t.size_i, numr_i);
                    G->dirty[s] = 1;
                }
                G->rng[s] = pool_base + uint64_t(pool_used);
                if (code) atomicMax(G->err, code);
                G->qsum[pos] = make_double4(hazard, j1, j2, jmax);
                G->qevents[pos] = code ? -1 : events;
                longest = max(longest, events);
                // the reference relaunches an owner after every COL_EVENT_CAP accepted events
                n_blocks += (unsigned long long)(events / COL_EVENT_CAP + 1);
                n_events += events;
                n_chain += 1;
            }
            active = false;
        }
    }
    // end of the kernel: segments of the block's first swarm add their counters in LDS and one lane
    // flushes them (every segment flushing at once would queue thousands of global atomics on a few
    // addresses); segments of another swarm flush directly
    if (kl == 0 && ctr_of)
    {
        const unsigned long long key = reinterpret_cast<unsigned long long>(ctr_of);
        const unsigned long long prev = atomicCAS(&end_key, 0ull, key);
        if (prev == 0ull || prev == key)
        {
            atomicAdd(&end_sum[0], n_events);
            atomicAdd(&end_sum[1], n_chain);
            atomicAdd(&end_sum[2], n_pairs);
            atomicMax(&end_sum[3], (unsigned long long)longest);
            atomicAdd(&end_sum[4], n_blocks);
            n_events = n_chain = n_pairs = n_blocks = 0;
            longest = 0;
            if (prev == 0ull) end_wave = wave_of;
        }
    }
    flush();
    __syncthreads();
    if (threadIdx.x == 0 && end_key)
    {
        unsigned long long *c = reinterpret_cast<unsigned long long *>(end_key);
        if (end_sum[0]) atomicAdd(c + 0, end_sum[0]);
        if (end_sum[1]) atomicAdd(c + 1, end_sum[1]);
        if (end_sum[2]) atomicAdd(c + 2, end_sum[2]);
        if (end_sum[3])
        {
            atomicMax(c + 3, end_sum[3]);
            atomicMax(end_wave + 0, end_sum[3]);
        }
        if (end_sum[4]) atomicAdd(end_wave + 1, end_sum[4]);
        if (end_sum[1]) atomicAdd(end_wave + 2, end_sum[1]);
    }
}

// audit of the chained owners, one lane per queue position of each due group (col_audit_bin.cu:13-60),
// AUDIT_ROUNDS positions per lane: a block covers 1024 slots of its group, and blocks without a queued
// position (most of them, the queues being short) leave after reading two counters
constexpr int AUDIT_ROUNDS = 4;
static_assert(BLOCK*AUDIT_ROUNDS == PUBLISH_OWNERS, "the audit uses the 1024-owner block starts");

__global__ __launch_bounds__(BLOCK) void k_chain_audit(gdx_wave w)
{
    __shared__ audit_lds a;
    int gi, s0, s1;
    if (!due_block(w, BLOCK*AUDIT_ROUNDS, gi, s0, s1)) return;
    const gdx_gref &G = w.table[w.ref[gi]];
    // queued positions: the heavy queue [beg, h1) and the light queue [l0, beg + cnt)
    const int h1 = w.beg[gi] + *G.hcount;
    const int l0 = w.beg[gi] + w.cnt[gi] - *G.qcount;
    if (s0 >= h1 && s1 <= l0) return;
    const double end = w.dt[gi];
    const double lo = *G.lower, up = *G.upper;
    const int *__restrict__ queue = G.queue;
    const double2 *__restrict__ st = G.st;
    const double2 *__restrict__ res = G.res;
    const double4 *__restrict__ cached = G.cached;
    const double4 *__restrict__ qsum = G.qsum;
    const int *__restrict__ qevents = G.qevents;
    clear_audit(a);
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < AUDIT_ROUNDS; ++k)
    {
        const int pos = s0 + k*BLOCK + int(threadIdx.x);
        // owners that ended in an error are left out: the operator call fails on them
        bool take = pos < s1 && (pos < h1 || pos >= l0) && qevents[pos] >= 0;
        // learned swarms audit their exact subsample only
        if (take && w.learned) take = exact_owner(w, gi, queue[pos]);
        double2 old = make_double2(1.0, 0.0);
        double size_end = 1.0, rate0 = 0.0;
        double4 sum = make_double4(0.0, 0.0, 0.0, 0.0);
        int events = 0;
        if (take)
        {
            int s = queue[pos];
            old = res[s];
            size_end = st[s].x;
            rate0 = cached[s].x;
            sum = qsum[pos];
            events = qevents[pos];
        }
        if (take)
        {
            const int bs = size_bin(old.x, lo, up);
            const int be = size_bin(size_end, lo, up);
            const double weight = old.y*grain_mass(old.x);
            const double probability = -expm1(-rate0*end);
            const double ev = static_cast<double>(events);
            const bool bad = !isfinite(weight) || !(weight > 0.0) || !isfinite(probability) || !isfinite(sum.x)
                || sum.x < 0.0 || !isfinite(ev) || !isfinite(sum.y) || sum.y < 0.0 || !isfinite(sum.z)
                || sum.z < 0.0 || !isfinite(sum.w) || sum.w < 0.0 || !isfinite(size_end) || !(size_end > 0.0);
            const double growth = fabs(log(size_end / old.x));
            audit_wave<true>(a, true, bs, be, weight, probability, sum.x, sum.y, sum.z, sum.w, ev, growth, bad);
        }
    }
    __syncthreads();
    flush_audit(a, G.raw_audit);
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

// every pair term of every owner (test access): the rate k_bath_rate folds for slot k of owner s
// (_get_col_chain_rate / measure x outcome factor, col_bath_rate.cu:13-129), written in particle order and
// the reference's slot order; empty slots and owners without a measure read 0
__global__ __launch_bounds__(256) void k_pair_rates(gdx_gref G, int n_active, double *out)
{
    const long long e = (long long)blockIdx.x*256 + threadIdx.x;
    if (e >= (long long)n_active*N_K) return;
    const int s = int(e / N_K), pos = int(e % N_K);
    const double2 me = G.st[s];
    const double m = G.measure[s];
    double r = 0.0;
    if (m > 0.0)
    {
        gdx_env env{};
        double hg_i = 0.0;
        if constexpr (COAG_KERNEL == 3) env = G.env[s];
        if constexpr (PARTNER_HG) hg_i = G.hgs[s];
        const owner_terms t = make_owner(env, me.x, m, hg_i);
        r = eval_pair_rate(t, load_partner(G.nbr[e], G.res, G.hgs), G.lambda_0).rate;
    }
    out[(long long)G.slot_pidx[s]*N_K + bitrev_slot(pos)] = r;
}

}  // namespace

void launch_pair_rates(context &c, const gdx_gref &G, double *out)
{
    if (c.n_active <= 0) return;
    const long long total = (long long)c.n_active*N_K;
    k_pair_rates<<<unsigned((total + 255) / 256), 256, 0, c.stream>>>(G, c.n_active, out);
    hip_check(hipGetLastError(), "launch pair rates");
}

void launch_reset_scratch(hipStream_t st, const gdx_gref *table, int ng)
{
    if (ng <= 0) return;
    k_reset_scratch<<<(ng + 255) / 256, 256, 0, st>>>(table, ng);
    hip_check(hipGetLastError(), "launch reset scratch");
}

// mark: some swarm of the wave tracks stale bounds (has a majorant); the changed-list counts
// (rb.nchanged) are cleared by the caller
void launch_publish(hipStream_t st, const gdx_wave &w, int blocks, run_buffers &rb)
{
    if (!blocks) return;
    // each list holds the owners of its blocks: at most ceil(grid / CHANGED_LISTS) blocks
    const int grid = xcd_grid(blocks);
    rb.changed_cap = (grid + CHANGED_LISTS - 1) / CHANGED_LISTS*PUBLISH_OWNERS;
    rb.changed.alloc(size_t(CHANGED_LISTS)*size_t(rb.changed_cap));
    k_publish<<<grid, BLOCK, 0, st>>>(w, rb.changed.p, rb.changed_cap, rb.nchanged.p);
    hip_check(hipGetLastError(), "launch publish");
}

// after launch_publish of the same wave: the slots it listed mark their readers stale
void launch_mark_stale(hipStream_t st, const gdx_wave &w, run_buffers &rb)
{
    if (!w.total) return;
    constexpr int PER_BLOCK = WAVES*MARK_SLOTS;
    const int grid = std::max(1, std::min(4096, (w.total + PER_BLOCK - 1) / PER_BLOCK));
    k_mark_stale<<<grid, BLOCK, 0, st>>>(w.table, rb.changed.p, rb.changed_cap, rb.nchanged.p);
    hip_check(hipGetLastError(), "launch mark stale");
}

void launch_bath_rate(hipStream_t st, const gdx_wave &w, int blocks)
{
    if (!blocks) return;
    k_bath_rate<1><<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch bath rate");
}

void launch_bath_rate_grouped(hipStream_t st, const gdx_wave &w, int blocks)
{
    if (!blocks) return;
    k_bath_rate<RATE_LANES><<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch lane-group rate");
}

__global__ __launch_bounds__(256) void k_event_sum(const gdx_event_work *w, int n, gdx_event_work *sum)
{
    __shared__ unsigned long long cnt[256];
    __shared__ double lm[256];
    const int k = blockIdx.x, t = threadIdx.x;
    unsigned long long a = 0;
    double b = 0.0;
    for (int i = t; i < n; i += 256)
    {
        a += w[i].count[k];
        b += w[i].log_mass[k];
    }
    cnt[t] = a;
    lm[t] = b;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1)
    {
        if (t < stride) { cnt[t] += cnt[t + stride]; lm[t] += lm[t + stride]; }
        __syncthreads();
    }
    if (t == 0) { sum->count[k] = cnt[0]; sum->log_mass[k] = lm[0]; }
}

void launch_event_sum(context &c, hipStream_t st)
{
    k_event_sum<<<GDX_EVENT_CATEGORIES, 256, 0, st>>>(c.evw.p, c.n_active, c.ev_sum.p);
    hip_check(hipGetLastError(), "launch event sum");
}

__global__ __launch_bounds__(256) void k_zero(zero_ranges z)
{
    for (int r = 0; r < z.n; ++r)
    {
        unsigned int *p = static_cast<unsigned int *>(z.p[r]);
        const unsigned long long words = z.bytes[r] / 4;
        for (unsigned long long i = (unsigned long long)blockIdx.x*blockDim.x + threadIdx.x; i < words;
             i += (unsigned long long)gridDim.x*blockDim.x)
            p[i] = 0u;
    }
}

void launch_zero(hipStream_t st, const zero_ranges &z)
{
    unsigned long long words = 0;
    for (int r = 0; r < z.n; ++r) words = std::max(words, z.bytes[r] / 4);
    if (!words) return;
    const int blocks = int(std::min<unsigned long long>(256, (words + 255) / 256));
    k_zero<<<blocks, 256, 0, st>>>(z);
    hip_check(hipGetLastError(), "launch zero");
}

__global__ __launch_bounds__(256) void k_to_host(uint4 *dst, const uint4 *src, int n)
{
    for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < n; i += gridDim.x*blockDim.x) dst[i] = src[i];
}

void launch_to_host(hipStream_t st, void *host_dst, const void *src, size_t bytes)
{
    const int n = int(bytes / 16);
    if (!n) return;
    k_to_host<<<std::min(16, (n + 255) / 256), 256, 0, st>>>(static_cast<uint4 *>(host_dst),
                                                             static_cast<const uint4 *>(src), n);
    hip_check(hipGetLastError(), "launch read-back");
}

void run_buffers::read_back(size_t from, size_t to, hipStream_t st)
{
    launch_to_host(st, back_h.p + from, back_d.p + from, to - from);
}

// a phase's due list (ints and bath ends) in the launch's arguments, written by the first block
struct stage_data
{
    unsigned int *dst;
    int words;
    unsigned int data[STAGE_INLINE_BYTES / 4];
};

__global__ __launch_bounds__(256) void k_stage(zero_ranges z, stage_data d)
{
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < d.words; i += blockDim.x) d.dst[i] = d.data[i];
    for (int r = 0; r < z.n; ++r)
    {
        unsigned int *p = static_cast<unsigned int *>(z.p[r]);
        const unsigned long long words = z.bytes[r] / 4;
        for (unsigned long long i = (unsigned long long)blockIdx.x*blockDim.x + threadIdx.x; i < words;
             i += (unsigned long long)gridDim.x*blockDim.x)
            p[i] = 0u;
    }
}

void launch_stage(hipStream_t st, const zero_ranges &z, void *dst, const void *src, size_t bytes)
{
    if (bytes > size_t(STAGE_INLINE_BYTES) || bytes % 4)
    {
        launch_zero(st, z);
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "due list");
        return;
    }
    stage_data d;
    d.dst = static_cast<unsigned int *>(dst);
    d.words = int(bytes / 4);
    std::memcpy(d.data, src, bytes);
    unsigned long long words = 0;
    for (int r = 0; r < z.n; ++r) words = std::max(words, z.bytes[r] / 4);
    const int blocks = int(std::max<unsigned long long>(1, std::min<unsigned long long>(256, (words + 255) / 256)));
    k_stage<<<blocks, 256, 0, st>>>(z, d);
    hip_check(hipGetLastError(), "launch stage");
}

void launch_majorant(hipStream_t st, const gdx_wave &w, int blocks)
{
    if (!blocks) return;
    k_majorant<<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch majorant");
}

void launch_sample(context &c, const gdx_gref &G, float *feat, double *scale, double *rate)
{
    if (c.n_active <= 0) return;
    k_sample<<<(c.n_active + 255) / 256, 256, 0, c.stream>>>(G, c.n_active, feat, scale, rate);
    hip_check(hipGetLastError(), "launch sample");
}

void launch_screen(hipStream_t st, const gdx_wave &w, int blocks, bool narrow)
{
    if (!blocks) return;
    if (narrow) k_screen<1><<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    else k_screen<SCREEN_ROUNDS><<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch screen");
}

// learned waves only, after the screen of the same wave
void launch_candidates(hipStream_t st, const gdx_wave &w, int blocks, int *vstart)
{
    if (!blocks || !w.learned) return;
    k_cand_prefix<<<1, 256, 0, st>>>(w, vstart);
    hip_check(hipGetLastError(), "launch candidate prefix");
    const int grid = std::min(blocks, CAND_GRID);
    k_candidates<<<std::max(XCDS, grid / XCDS*XCDS), BLOCK, 0, st>>>(w, vstart);
    hip_check(hipGetLastError(), "launch candidates");
}

void launch_queue_prefix(hipStream_t st, const gdx_wave &w, int *qprefix)
{
    k_queue_prefix<<<1, 256, 0, st>>>(w, qprefix);
    hip_check(hipGetLastError(), "launch queue prefix");
}

int chain_resident_blocks()
{
    int per_cu = 0, cus = 0, dev = 0;
    hip_check(hipGetDevice(&dev), "get device");
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chain, BLOCK, 0), "chain occupancy");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    return std::max(1, per_cu*cus);
}

// a persistent grid of at most `resident` blocks, fewer for small waves
void launch_chain(hipStream_t st, const gdx_wave &w, int resident)
{
    if (!w.total) return;
    int grid = std::max(1, std::min(resident, (w.total + 4*NSEG - 1) / (4*NSEG)));
    k_chain<<<grid, BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch chain");
}

void launch_chain_audit(hipStream_t st, const gdx_wave &w, int blocks)
{
    if (!blocks) return;
    k_chain_audit<<<xcd_grid(blocks), BLOCK, 0, st>>>(w);
    hip_check(hipGetLastError(), "launch chain audit");
}

void launch_gather_cached(context &c, double *out)
{
    if (c.n_active <= 0) return;
    k_cached_to_particles<<<(c.n_active + 255) / 256, 256, 0, c.stream>>>(c.cached.p, c.slot_pidx.p, c.n_active, out);
    hip_check(hipGetLastError(), "launch gather cached");
}

}  // namespace gdx
