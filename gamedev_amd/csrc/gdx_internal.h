// Host-side internals of the collision library: device buffers, the operator context and the
// launch wrappers implemented by geometry.hip and operator.hip.
#ifndef GDX_INTERNAL_H
#define GDX_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gdx_col.h"
#include "gdx_model.h"

namespace gdx {

// library status, mapped 1:1 onto the reference's failure paths (see include/gdx_col.h)
struct status_error : std::runtime_error
{
    int code;
    status_error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

void hip_check(hipError_t e, const char *what);

// owning device allocation
template <typename T>
struct dbuf
{
    T *p = nullptr;
    size_t n = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        if (count == 0) count = 1;
        hip_check(hipMalloc(reinterpret_cast<void **>(&p), sizeof(T)*count), "device allocation");
        n = count;
    }
    // grow preserving the first `keep` entries
    void grow(size_t count, size_t keep, hipStream_t s)
    {
        if (count <= n && p) return;
        T *q = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&q), sizeof(T)*count), "device allocation");
        if (p && keep) hip_check(hipMemcpyAsync(q, p, sizeof(T)*keep, hipMemcpyDeviceToDevice, s), "grow copy");
        hip_check(hipStreamSynchronize(s), "grow sync");
        release();
        p = q;
        n = count;
    }
};

// page-locked host array: device-to-host copies land without a staging copy
template <typename T>
struct pinned
{
    T *p = nullptr;
    size_t n = 0;
    pinned() = default;
    pinned(const pinned &) = delete;
    pinned &operator=(const pinned &) = delete;
    ~pinned() { if (p) (void)hipHostFree(p); }
    void resize(size_t count)
    {
        if (count <= n && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&p), sizeof(T)*std::max<size_t>(count, 1)), "pinned host buffer");
        n = count;
    }
    T *data() { return p; }
    const T *data() const { return p; }
};

// up to eight device ranges cleared by one launch (byte counts are multiples of 4)
struct zero_ranges
{
    void *p[8];
    unsigned long long bytes[8];
    int n = 0;
    void add(void *ptr, size_t b)
    {
        if (n == 8) throw std::logic_error("zero_ranges: more than 8 ranges");
        p[n] = ptr;
        bytes[n] = b;
        ++n;
    }
};

// particle-order view of one bound state field: element i lives at base + i*stride bytes
struct field
{
    const char *base = nullptr;
    long long stride = 8;
};
struct bound_state
{
    field x, y, z, vx, vy, vz, size, numr;   // vx..vz optional (finiteness screen only)
    bool bound = false;
};

template <typename T>
void swap_buf(dbuf<T> &a, dbuf<T> &b)
{
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

// left-balanced k-d tree over all collision sites, node order: {x, y, z, split dim bits}
struct kd_tree
{
    int n = 0;
    dbuf<float4> node;
    dbuf<int> idx;             // particle index per node
    dbuf<float> world;         // data bounds {lo xyz, hi xyz}
};

// a typed view into one of run_buffers' shared allocations
template <typename T>
struct view
{
    T *p = nullptr;
    T *data() { return p; }
    const T *data() const { return p; }
};

// buffers of operator runs: the group table, pinned staging and device copies of the due lists,
// the chain's queue prefix
struct run_buffers
{
    // due-list layout per phase: ref, beg, cnt, sub (n each), then five block-start arrays (n + 1 each),
    // then the bath ends (doubles): one page-locked block, one device block, one copy per phase
    static constexpr int DUE_INTS = 9, DUE_EXTRA = 5;
    dbuf<gdx_gref> table;
    std::vector<gdx_gref> h_table;
    pinned<unsigned char> due_h;
    dbuf<unsigned char> due_d;
    int *h_int = nullptr;
    double *h_dt = nullptr;
    view<int> d_int;
    view<double> d_dt;
    size_t due_ints_bytes = 0;
    size_t cap = 0;
    dbuf<int> qprefix;
    dbuf<int> cprefix;                // candidate blocks per due group (exclusive prefix), learned chain phases
    dbuf<gdx_changed> changed;        // a publish's changed slots and their replaced records (bound upkeep)
    dbuf<int> nchanged;               // entries of each changed-slot list
    int changed_cap = 0;              // entries each changed-slot list can hold
    int resident = 0;                 // resident blocks of the persistent chain kernel
    // controller sums, queue lengths, counters and errors of every group and swarm of the run. The
    // read-backs share one device block and one page-locked mirror laid out [raw_rate][wave_ctr][err]
    // [raw_audit], so that the rate phase reads [raw_rate, wave_ctr] and the chain phase [wave_ctr,
    // raw_audit] with one copy each
    dbuf<unsigned char> back_d;
    pinned<unsigned char> back_h;
    size_t off_rate = 0, off_wave = 0, off_err = 0, off_audit = 0, back_bytes = 0;
    view<gdx_raw_rate> raw_rate, h_rate;
    view<gdx_raw_audit> raw_audit, h_audit;
    view<int> err, h_err;
    view<unsigned long long> wave_ctr, h_wave;
    dbuf<int> qcount, hcount, ccount;
    dbuf<unsigned long long> counters;
    pinned<unsigned long long> h_counters;
    void size_run(int groups, int swarms)
    {
        auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
        off_rate = 0;
        off_wave = align(off_rate + sizeof(gdx_raw_rate)*size_t(groups)*COL_BIN_S);
        off_err = align(off_wave + sizeof(unsigned long long)*size_t(swarms)*GDX_WAVE_CTRS);
        off_audit = align(off_err + sizeof(int)*size_t(swarms));
        back_bytes = align(off_audit + sizeof(gdx_raw_audit)*size_t(groups)*COL_BIN_S);
        back_d.alloc(back_bytes);
        back_h.resize(back_bytes);
        raw_rate.p = reinterpret_cast<gdx_raw_rate *>(back_d.p + off_rate);
        wave_ctr.p = reinterpret_cast<unsigned long long *>(back_d.p + off_wave);
        err.p = reinterpret_cast<int *>(back_d.p + off_err);
        raw_audit.p = reinterpret_cast<gdx_raw_audit *>(back_d.p + off_audit);
        h_rate.p = reinterpret_cast<gdx_raw_rate *>(back_h.p + off_rate);
        h_wave.p = reinterpret_cast<unsigned long long *>(back_h.p + off_wave);
        h_err.p = reinterpret_cast<int *>(back_h.p + off_err);
        h_audit.p = reinterpret_cast<gdx_raw_audit *>(back_h.p + off_audit);
        qcount.alloc(groups);
        hcount.alloc(groups);
        ccount.alloc(groups);
        counters.alloc(size_t(swarms)*GDX_CTRS);
        h_counters.resize(size_t(swarms)*GDX_CTRS);
    }
    // the read-back bytes [from, to) (offsets above, multiples of 256) written into the page-locked mirror
    // by a kernel on the stream (no copy engine to synchronise with)
    void read_back(size_t from, size_t to, hipStream_t st);
    run_buffers() = default;
    run_buffers(const run_buffers &) = delete;
    run_buffers &operator=(const run_buffers &) = delete;
    // room for n due entries
    void reserve(size_t n)
    {
        if (n <= cap && h_int) return;
        size_t m = std::max<size_t>(n, 64);
        due_ints_bytes = (sizeof(int)*(DUE_INTS*m + DUE_EXTRA) + 15) & ~size_t(15);
        const size_t bytes = due_ints_bytes + sizeof(double)*m;
        due_h.resize(bytes);
        due_d.alloc(bytes);
        h_int = reinterpret_cast<int *>(due_h.p);
        h_dt = reinterpret_cast<double *>(due_h.p + due_ints_bytes);
        d_int.p = reinterpret_cast<int *>(due_d.p);
        d_dt.p = reinterpret_cast<double *>(due_d.p + due_ints_bytes);
        qprefix.alloc(2*(m + 1));
        cprefix.alloc(m + 1);
        nchanged.alloc(CHANGED_LISTS);
        cap = m;
    }
};

// reusable scratch for radix sorts and scans
struct sort_scratch
{
    dbuf<uint64_t> keys_alt;
    dbuf<int> vals_alt;
    dbuf<unsigned int> hist;
    dbuf<unsigned int> partial;
};

struct morton_index
{
    int n = 0, dim = 3, max_level = 20, leaf_target = 128;
    float ox = 0, oy = 0, oz = 0, width = 1;
    dbuf<uint64_t> keys;       // sorted keys
    dbuf<int> idx;             // sorted record -> source index
    dbuf<float4> rec;          // sorted records {x, y, z, code}; code sign bit marks inactive
    dbuf<gdx_node> nodes;
    int node_count = 0;
    int leaf_count = 0;
};

struct context
{
    // configuration
    int n = 0;
    double total_mass = 0.0;
    uint64_t seed = 0;
    int flags = 0;
    int search = GDX_COL_SEARCH_MORTON;
    int leaf_target = MORTON_LEAF_TARGET;   // Morton refinement threshold (gdx_col_set_leaf_target, autotuner)
    int last_waves = 0;        // refresh waves of the last operator call (ensemble load balance)
    int topk_tile = 16;        // top-K merge tile: 0 insertion, 8/16/32 keys per merge (gdx_col_set_topk_tile, autotuner)
    int topk_waves = 1;        // top-K block shape: 1, 2 or 4 wavefronts per block (gdx_col_set_topk_block, autotuner)
    int narrow_owners = GDX_COL_NARROW_OWNERS;   // waves of at most this many owners take the narrow
                                                 // rate and screen launches (gdx_col_set_narrow_waves)
    hipStream_t stream = nullptr;
    bound_state user;

    // geometry epoch (particle order)
    dbuf<float4> pts;          // {x, y, z, search cap}
    dbuf<unsigned char> active;
    dbuf<unsigned char> group;
    dbuf<int> bad;             // first non-finite particle + 1
    morton_index index;
    kd_tree kd;
    sort_scratch scratch;
    dbuf<int> pidx_slot;       // particle -> slot (-1 inactive)
    dbuf<int> slot_pidx;       // slot -> particle
    dbuf<unsigned long long> graph;   // per group: 64-bit dependency mask
    dbuf<unsigned int> overflow;
    // azimuthal wedge: physical records then seam ghosts, their codes 3*i + image, ghost counts and offsets
    dbuf<float4> wedge_rec;
    dbuf<int> wedge_code;
    dbuf<unsigned int> ghost_count, ghost_off;
    dbuf<unsigned char> nbr_image;  // n_active x N_K image of each neighbour (slot order)
    int n_records = 0;              // indexed records: n, plus ghosts in a wedge
    float q_max = 0.0f;             // largest search cap of the epoch (wedge seam test)
    bool unique_ids = true;         // images of one particle can never share a search ball
    int n_active = 0;
    std::vector<int> group_beg, group_cnt, group_all;   // slot ranges of active owners; all particles per group
    std::vector<uint32_t> edges;                        // LOCAL_GROUPS x LOCAL_WORDS
    bool geometry_valid = false;
    bool state_loaded = false;

    // per-slot operator state (slot order)
    dbuf<int> nbr;             // n_active x N_K partner slots, -1 empty
    dbuf<double> measure;
    dbuf<double2> st;          // {size, numr}
    dbuf<double2> res;         // frozen reservoir {size_old, numr_old}
    dbuf<double4> cached;      // bath-start {rate (-1 invalid), first, second, maximum}
    dbuf<gdx_env> env;
    dbuf<double> hgs;          // partner gas scale height (vertically integrated physical kernel)
    dbuf<uint64_t> rng;        // per-slot draw counters
    dbuf<uint64_t> rng_p;      // per-particle draw counters (particle order, persistent)

    // controller scratch
    dbuf<unsigned long long> scr_min, scr_max;
    dbuf<double> lower, upper;
    dbuf<int> queue;           // event owners of one wave, in their group's slot range
    dbuf<double4> qsum;        // chain sums (hazard, j1, j2, jmax) per queue position
    dbuf<int> qevents;         // chain events per queue position, -1 after an error
    dbuf<double2> qstart;      // learned: {start time, event at the start} per queue position
    dbuf<int> cqueue;          // learned: thinning candidates per group slot range
    dbuf<double2> cstart;      // learned: {candidate time, bound} per candidate
    dbuf<double> lam;          // learned: per-slot S * softplus(net) of the last evaluation
    dbuf<double> lam_add;      // learned: exact pair-rate increases from partner changes since then
    dbuf<unsigned char> dirty; // state changed since the slot's last publish
    dbuf<unsigned char> stale; // learned: bound inputs changed since the slot's last evaluation
    dbuf<unsigned> rev_beg, rev_end;   // learned: reverse neighbour lists of the geometry epoch
    dbuf<int> rev;
    bool rev_valid = false;
    dbuf<float2> res32;        // published {size, number} in FP32 for the majorant's partner gathers
    const void *mlp = nullptr; // learned: packed network of the attached majorant (null: plain)
    gdx_col_majorant model = nullptr;
    unsigned long long model_version = 0;   // the model's version the cached bounds were computed with
    int bound_reuse = 1;       // learned: keep the bounds of owners whose inputs are unchanged
    double safety = 1.2;
    double safety_scale = 1.0; // learned: adaptive factor on the bound, kept across operator calls
    int thin_stride = THIN_STRIDE;   // learned: exact subsample stride (THIN_STRIDE_LARGE for large swarms)
    int majorant_min_owners = GDX_COL_MAJORANT_MIN_OWNERS;   // smaller swarms keep the plain method
    double fixed_bath = 0.0;   // > 0: every bath lasts this long (test access; the controller still audits)
    int exact_walk = 0;        // 1: the chain's partner walk from sequential partials only (test access)
    bool diag = false;         // GDX_COL_DIAGNOSTICS
    dbuf<gdx_event_work> evw;  // per-slot event statistics of the current call
    dbuf<gdx_event_work> ev_sum;
    std::vector<double> h_lower, h_upper;

    // host controller memory (inc/swarm/_col_chain.cuh:46; persists across operator calls)
    struct bath_state { double limit_scale = 1.0; int activity_streak = 0, quiet_streak = 0, floor_streak = 0; };
    std::vector<bath_state> state;

    // optional per-kernel timing (HIP events on the launch stream)
    struct timing_t
    {
        double ms[GDX_COL_KERNELS] = {};
        long long launches[GDX_COL_KERNELS] = {}, owners[GDX_COL_KERNELS] = {};
    } timing;
    std::vector<hipEvent_t> events;

    std::string last_error;
};

// ---- geometry.hip ---------------------------------------------------------------------------
// the reverse neighbour lists of a swarm with a majorant address n_active * N_K entries with 32-bit offsets
constexpr long long MAJORANT_MAX_OWNERS = 4294967295LL / N_K;
void scan_exclusive_u32(const unsigned int *in, unsigned int *out, int n, sort_scratch &s, hipStream_t st);
void radix_sort_pairs(uint64_t *keys, int *vals, int n, int begin_bit, int end_bit, sort_scratch &s, hipStream_t st);
void morton_build(morton_index &ix, const float4 *points, const unsigned char *active, int n, float ox, float oy,
                  float oz, float width, int dim, int leaf_target, int max_level, sort_scratch &s, hipStream_t st,
                  const int *codes = nullptr);
void geometry_epoch(context &c);
void load_state(context &c);
void store_state(context &c);
// test access: exact top-K over a built index for arbitrary queries and a runtime K
void knn_query(const morton_index &ix, const float4 *queries, int nq, int k, int *codes, float *d2,
               unsigned int *overflow, hipStream_t st);

// ---- operator.hip ---------------------------------------------------------------------------
void kd_build(context &c);
void kd_query(context &c);
void launch_reset_scratch(hipStream_t st, const gdx_gref *table, int ng);
void launch_publish(hipStream_t st, const gdx_wave &w, int blocks, run_buffers &rb);
void launch_mark_stale(hipStream_t st, const gdx_wave &w, run_buffers &rb);
// rate pass: one lane per owner (block starts of 256 owners) for plain swarms' groups in large waves; a
// group of lanes per owner (block starts of RATE_SUB_OWNERS owners) for learned swarms' exact subsamples
// and for plain groups in waves of at most context::narrow_owners plain owners
void launch_bath_rate(hipStream_t st, const gdx_wave &w, int blocks);
constexpr int RATE_SUB_OWNERS = 64;
void launch_bath_rate_grouped(hipStream_t st, const gdx_wave &w, int blocks);
// narrow: one owner per lane over block starts of 256 owners (waves of at most context::narrow_owners
// owners), else four per lane over block starts of 1024
void launch_screen(hipStream_t st, const gdx_wave &w, int blocks, bool narrow);
void launch_candidates(hipStream_t st, const gdx_wave &w, int blocks, int *vstart);
void launch_queue_prefix(hipStream_t st, const gdx_wave &w, int *qprefix);
void launch_majorant(hipStream_t st, const gdx_wave &w, int blocks);
void launch_zero(hipStream_t st, const zero_ranges &z);
// bytes (a multiple of 16) from device memory into page-locked host memory by a kernel on the stream
void launch_to_host(hipStream_t st, void *host_dst, const void *src, size_t bytes);
// the ranges cleared and `bytes` of host data (multiple of 4) written to dst by one launch: the data ride
// in the kernel's arguments when they fit (STAGE_INLINE_BYTES), else a zero launch and a copy
constexpr int STAGE_INLINE_BYTES = 3072;
void launch_stage(hipStream_t st, const zero_ranges &z, void *dst, const void *src, size_t bytes);
void launch_sample(context &c, const gdx_gref &G, float *feat, double *scale, double *rate);
// ---- majorant.hip ---------------------------------------------------------------------------
const void *majorant_pack(gdx_col_majorant m);
unsigned long long majorant_version(gdx_col_majorant m);
void invalidate_bounds(context &c);
void build_reverse(context &c);
int chain_resident_blocks();
void launch_chain(hipStream_t st, const gdx_wave &w, int resident);
void launch_chain_audit(hipStream_t st, const gdx_wave &w, int blocks);
void launch_event_sum(context &c, hipStream_t st);
void launch_gather_cached(context &c, double *out_particle_order);
void launch_pair_rates(context &c, const gdx_gref &G, double *out_particle_order);

}  // namespace gdx

#endif  // GDX_INTERNAL_H
