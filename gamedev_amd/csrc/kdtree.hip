// k-d tree collision search on gfx950: a left-balanced k-d tree over every collision site and an
// exact top-K query per active owner, the reference's COLLISION_SEARCH=kdtree path.
//
// Reference behaviour followed: the tag-update build of kdtree::buildTree_thrust
// (inc/swarm/kdtree/builder_thrust.h:263-318, 418-526; layout helpers helpers.h:95-300; subtree
// domains builder_common.h:205-240; widest dimension box.h:111-125) and col_cache_get (KD)
// (inc/swarm/_col_cache.cuh:15-71) with the bounded max-heap of idx_old_heap
// (inc/swarm/kdtree/index_heap.cuh:11-177) and the closest-corner-tracking traversal
// (inc/swarm/kdtree/traverse-cct.h:26-100).
//
// Build: every level sorts (tag, coordinate along the point's split dimension) with the library's
// stable LSD radix sort (the reference's thrust::sort is not stable, so its layout for equal
// coordinates is unspecified; this layout is the stable one), then one thread per array position
// moves its point into the left or right child subtree, or settles it as the subtree's pivot, and
// sets its split dimension to the widest extent of the child subtree's domain.
// Query: one lane per active owner, owners taken in Morton order so that neighbouring lanes walk
// similar paths; heap and stack are per-lane private arrays. Rows are written in heap order.
#include "gdx_internal.h"
#include "gdx_device.h"

#include <algorithm>
#include <climits>
#include <cmath>

#pragma clang fp contract(off)

namespace gdx {
namespace {

constexpr int KD_BLOCK = 256;
constexpr int KD_QUERY_BLOCK = 64;
constexpr int KD_STACK = 30;                 // depth of a left-balanced tree of up to 2^30 points
constexpr int WAVE = 64;

int blocks(long long n, int per) { return static_cast<int>((n + per - 1) / per); }

__device__ __forceinline__ double ld(const field &f, long long i)
{
    return *reinterpret_cast<const double *>(f.base + i*f.stride);
}

// float coordinate -> unsigned key with the same order (-0 and +0 equal, as the comparison is)
__device__ __forceinline__ uint32_t ordered(float x)
{
    uint32_t u = __float_as_uint(x == 0.0f ? 0.0f : x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float coord(const float4 &p, int d) { return d == 0 ? p.x : (d == 1 ? p.y : p.z); }

struct kd_box
{
    float lo[3], hi[3];
};

// widest extent; a later dimension wins ties
__device__ __forceinline__ int widest(const kd_box &b)
{
    int best = 0;
    float w_best = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        float w = b.hi[d] - b.lo[d];
        if (w < w_best) continue;
        w_best = w;
        best = d;
    }
    return best;
}

__device__ __forceinline__ int level_of(int node) { return 63 - __clzll(static_cast<long long>(node) + 1ll); }
__device__ __forceinline__ int levels_for(int n) { return level_of(n - 1) + 1; }
__device__ __forceinline__ int full_nodes(int levels) { return (1 << levels) - 1; }

__device__ __forceinline__ int subtree_size(int n, int root)
{
    int levels_sub = levels_for(n) - level_of(root);
    int first = (root + 1) << (levels_sub - 1);
    int on_last = 1 << (levels_sub - 1);
    int last_on_last = first + on_last - 1;
    int missing = min(max(last_on_last - n, 0), on_last);
    return full_nodes(levels_sub) - missing;
}

// array position of a level-L subtree's pivot during step L
__device__ __forceinline__ int pivot_pos(int L, int n, int subtree)
{
    int settled = full_nodes(L);
    int levels_total = levels_for(n);
    int remaining = levels_total - L;
    int earlier = subtree - settled;
    int left_if_full = earlier*full_nodes(remaining);
    int left_on_last_if_full = earlier*(1 << (remaining - 1));
    int total_on_last = n - full_nodes(levels_total - 1);
    int missing = left_on_last_if_full - min(total_on_last, left_on_last_if_full);
    int left_root = 2*subtree + 1;
    return settled + left_if_full - missing + (left_root >= n ? 0 : subtree_size(n, left_root));
}

// ---------------------------------------------------------------------------------------------
// build
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(KD_BLOCK) void k_kd_init(const float4 *pts, int n, float4 *node, int *idx,
                                                     unsigned int *world_bits)
{
    __shared__ unsigned int red[6][KD_BLOCK / WAVE];
    int i = blockIdx.x*KD_BLOCK + threadIdx.x;
    float4 p = pts[min(i, n - 1)];
    if (i < n)
    {
        node[i] = make_float4(p.x, p.y, p.z, __int_as_float(0));
        idx[i] = i;
    }
    // world box of the data: ordered-key minima and maxima, one atomic per block and bound
    unsigned int k[6] = {ordered(p.x), ordered(p.y), ordered(p.z), ordered(p.x), ordered(p.y), ordered(p.z)};
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
        unsigned int v = k[b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
        {
            unsigned int t = __shfl_xor(v, o, WAVE);
            v = (b < 3) ? min(v, t) : max(v, t);
        }
        if (lane == 0) red[b][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        int b = threadIdx.x;
        unsigned int v = red[b][0];
        for (int w = 1; w < KD_BLOCK / WAVE; ++w) v = (b < 3) ? min(v, red[b][w]) : max(v, red[b][w]);
        if (b < 3) atomicMin(world_bits + b, v);
        else atomicMax(world_bits + b, v);
    }
}

__device__ __forceinline__ float unordered(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_kd_world(const unsigned int *world_bits, float *world, int *dim0)
{
    if (threadIdx.x != 0) return;
    kd_box b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = unordered(world_bits[d]); b.hi[d] = unordered(world_bits[3 + d]); }
    for (int d = 0; d < 3; ++d) { world[d] = b.lo[d]; world[3 + d] = b.hi[d]; }
    *dim0 = widest(b);
}

__global__ __launch_bounds__(KD_BLOCK) void k_kd_dim0(float4 *node, int n, const int *dim0)
{
    int i = blockIdx.x*KD_BLOCK + threadIdx.x;
    if (i < n) node[i].w = __int_as_float(*dim0);
}

__global__ __launch_bounds__(KD_BLOCK) void k_kd_keys(const float4 *node, const uint32_t *tag, int n, uint64_t *key,
                                                     int *val, bool tag_only)
{
    int i = blockIdx.x*KD_BLOCK + threadIdx.x;
    if (i >= n) return;
    float4 p = node[i];
    uint64_t t = tag[i];
    key[i] = tag_only ? t : (t << 32) | ordered(coord(p, __float_as_int(p.w)));
    val[i] = i;
}

__global__ __launch_bounds__(KD_BLOCK) void k_kd_gather(const uint64_t *key, const int *val, const float4 *node,
                                                       const int *idx, int n, bool tag_only, float4 *node_out,
                                                       int *idx_out, uint32_t *tag_out)
{
    int i = blockIdx.x*KD_BLOCK + threadIdx.x;
    if (i >= n) return;
    int from = val[i];
    node_out[i] = node[from];
    idx_out[i] = idx[from];
    tag_out[i] = tag_only ? uint32_t(key[i]) : uint32_t(key[i] >> 32);
}

// step L of the tag update: positions below the settled count are final; every other point moves
// into its subtree's left or right child, or stays as the subtree's pivot (its node)
__global__ __launch_bounds__(KD_BLOCK) void k_kd_update(float4 *node, uint32_t *tag, int n, int L,
                                                       const float *world)
{
    int g = blockIdx.x*KD_BLOCK + threadIdx.x;
    if (g >= n || g < full_nodes(L)) return;
    int subtree = int(tag[g]);
    kd_box b;
#pragma unroll
    for (int d = 0; d < 3; ++d) { b.lo[d] = world[d]; b.hi[d] = world[3 + d]; }
    for (int cur = subtree; cur > 0;)
    {
        int parent = (cur + 1) / 2 - 1;
        float4 pp = node[parent];
        int d = __float_as_int(pp.w);
        float split = coord(pp, d);
        if (cur & 1) b.hi[d] = fminf(split, b.hi[d]);
        else b.lo[d] = fmaxf(split, b.lo[d]);
        cur = parent;
    }
    int pivot = pivot_pos(L, n, subtree);
    float4 pv = node[pivot];
    int pd = __float_as_int(pv.w);
    float pc = coord(pv, pd);
    if (g < pivot) { subtree = 2*subtree + 1; b.hi[pd] = pc; }
    else if (g > pivot) { subtree = 2*subtree + 2; b.lo[pd] = pc; }
    if (g != pivot) node[g].w = __int_as_float(widest(b));
    tag[g] = uint32_t(subtree);
}

// ---------------------------------------------------------------------------------------------
// query
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float d2_of(float ax, float ay, float az, float bx, float by, float bz)
{
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    return dx*dx + dy*dy + dz*dz;
}

__device__ __forceinline__ float key_d2(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }

// col_site_init's wedge images (src/swarm/col_site_init.cu:42-62): physical points at [0, n), every point
// rotated by -width (image 1) at [n, 2n) and by +width (image 2) at [2n, 3n)
__global__ __launch_bounds__(KD_BLOCK) void k_kd_images(const float4 *pts, int n, float cw, float sw, float4 *out)
{
    const long long i = (long long)blockIdx.x*KD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 p = pts[i];
    out[i] = p;
#pragma unroll
    for (int image = 0; image < 2; ++image)
    {
        const float sa = image == 0 ? -sw : sw;
        out[i + (long long)(image + 1)*n] = make_float4(cw*p.x - sa*p.y, sa*p.x + cw*p.y, p.z, p.w);
    }
}

// queries: the nq physical particles in Morton order; the tree holds nt records, record r being particle
// r % n_phys at image r / n_phys when n_phys > 0 (a wedge), else particle r
__global__ __launch_bounds__(KD_QUERY_BLOCK) void k_kd_topk(const float4 *node, const int *node_idx, int nt, int nq,
                                                           int n_phys, float image_dist_min,
                                                           const float *world, const int *sidx, const float4 *pts,
                                                           const unsigned char *active, const int *pidx_slot,
                                                           const unsigned char *group, bound_state u, int *nbr,
                                                           double *measure, unsigned long long *graph,
                                                           unsigned int *overflow, unsigned char *nbr_image)
{
    const int lane = threadIdx.x % WAVE;
    const int m = blockIdx.x*KD_QUERY_BLOCK + threadIdx.x;
    const int n = nt;
    const int p = m < nq ? sidx[m] : 0;
    const bool qactive = m < nq && active[p];
    unsigned long long mask = 0ull;
    int g = qactive ? int(group[p]) : -1;
    if (qactive)
    {
        const float4 qp = pts[p];
        uint64_t heap[N_K];
        const uint64_t empty = (uint64_t(__float_as_uint(qp.w*qp.w)) << 32) | 0xffffffffull;
        // one entry per particle when two of its images can share this ball (_col_cache.cuh:38-41)
        const bool dedup = X_WEDGE && !(image_dist_min < 0.0f || image_dist_min > 2.0f*qp.w);
#pragma unroll
        for (int k = 0; k < N_K; ++k) heap[k] = empty;
        int st_node[KD_STACK];
        float st_c[KD_STACK][3];
        int sp = 0;
        float cull = nextafterf(key_d2(heap[0]), __builtin_inff());
        float cx = fminf(fmaxf(qp.x, world[0]), world[3]);
        float cy = fminf(fmaxf(qp.y, world[1]), world[4]);
        float cz = fminf(fmaxf(qp.z, world[2]), world[5]);
        bool run = !(d2_of(cx, cy, cz, qp.x, qp.y, qp.z) > cull);
        int id = 0;
        while (run)
        {
            if (id >= n)
            {
                bool found = false;
                while (sp > 0)
                {
                    --sp;
                    cx = st_c[sp][0]; cy = st_c[sp][1]; cz = st_c[sp][2];
                    if (d2_of(cx, cy, cz, qp.x, qp.y, qp.z) >= cull) continue;
                    id = st_node[sp];
                    found = true;
                    break;
                }
                if (!found) break;
            }
            const float4 np = node[id];
            const int r = node_idx[id];
            const int image = X_WEDGE ? r / n_phys : 0;
            const int j = X_WEDGE ? r - image*n_phys : r;
            const float d2 = d2_of(np.x, np.y, np.z, qp.x, qp.y, qp.z);
            if (active[j])
            {
                uint64_t cand = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(3*j + image);
                // a particle already in the heap is replaced only by a nearer image of itself
                // (idx_old_heap::processCandidate, inc/swarm/kdtree/index_heap.cuh:112-175)
                int slot = 0;
                bool take = cand < heap[0];
                if (X_WEDGE && dedup)
                {
                    int found = -1;
#pragma unroll
                    for (int k = 0; k < N_K; ++k)
                    {
                        const uint32_t ck = uint32_t(heap[k]);
                        if (found < 0 && ck != 0xffffffffu && int(ck / 3u) == j) found = k;
                    }
                    if (found >= 0)
                    {
                        slot = found;
                        take = cand < heap[found];
                    }
                }
                if (take)
                {
                    while (true)
                    {
                        int child = 2*slot + 1;
                        if (child >= N_K) { heap[slot] = cand; break; }
                        uint64_t cv = heap[child];
                        if (child + 1 < N_K && heap[child + 1] > cv) { cv = heap[child + 1]; ++child; }
                        if (cv < cand) { heap[slot] = cand; break; }
                        heap[slot] = cv;
                        slot = child;
                    }
                }
            }
            cull = nextafterf(key_d2(heap[0]), __builtin_inff());
            const int dim = __float_as_int(np.w);
            const float nd = coord(np, dim);
            const bool left_close = coord(qp, dim) < nd;
            const int lc = 2*id + 1, rc = lc + 1;
            const int far = left_close ? rc : lc;
            float fx = cx, fy = cy, fz = cz;
            if (dim == 0) fx = nd; else if (dim == 1) fy = nd; else fz = nd;
            if (far < n && d2_of(fx, fy, fz, qp.x, qp.y, qp.z) < cull)
            {
                if (sp == KD_STACK) { atomicOr(overflow, 1u); run = false; break; }
                st_node[sp] = far;
                st_c[sp][0] = fx; st_c[sp][1] = fy; st_c[sp][2] = fz;
                ++sp;
            }
            id = left_close ? lc : rc;
        }
        const int s = pidx_slot[p];
        float max_d2 = 0.0f;
        int row[N_K];
#pragma unroll
        for (int k = 0; k < N_K; ++k)
        {
            uint32_t code = uint32_t(heap[k]);
            if (code != 0xffffffffu)
            {
                max_d2 = fmaxf(max_d2, key_d2(heap[k]));
                int jj = int(code) / 3;
                row[k] = pidx_slot[jj];
                mask |= 1ull << group[jj];
            }
            else
            {
                row[k] = -1;
            }
        }
        int4 *dst = reinterpret_cast<int4 *>(nbr + (long long)s*N_K);
#pragma unroll
        for (int t = 0; t < N_K; t += 4)
            dst[t / 4] = make_int4(row[bitrev_slot(t)], row[bitrev_slot(t + 1)], row[bitrev_slot(t + 2)],
                                   row[bitrev_slot(t + 3)]);
        if constexpr (X_WEDGE)
        {
#pragma unroll
            for (int t = 0; t < N_K; ++t)
            {
                const uint32_t code = uint32_t(heap[bitrev_slot(t)]);
                nbr_image[(long long)s*N_K + t] = static_cast<unsigned char>(code != 0xffffffffu ? code % 3u : 0u);
            }
        }
        measure[s] = ball_measure(ld(u.y, p), ld(u.z, p), sqrt(static_cast<double>(max_d2)));
    }
    // one atomic per distinct owner group of this wavefront (col_dep_graph.cu:14-60)
    unsigned long long todo = __ballot(qactive);
    while (todo)
    {
        int leader = __ffsll(static_cast<long long>(todo)) - 1;
        int lg = __builtin_amdgcn_readlane(g, leader);
        unsigned long long mine = (g == lg) ? mask : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mine |= __shfl_xor(mine, o, WAVE);
        if (lane == leader) atomicOr(graph + lg, mine);
        todo &= ~__ballot(g == lg);
    }
}

}  // namespace

void kd_build(context &c)
{
    const float4 *src = c.pts.p;
    int n = c.n;
    hipStream_t st = c.stream;
    if constexpr (X_WEDGE)
    {
        if (3ll*n > INT_MAX / 3) throw status_error(12, "wedge k-d node count exceeds the index limit");
        const float width = static_cast<float>(X_MAX - X_MIN);
        const float sw = static_cast<float>(std::sin(static_cast<double>(width)));
        const float cw = static_cast<float>(std::cos(static_cast<double>(width)));
        c.wedge_rec.alloc(size_t(3)*n);
        k_kd_images<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(c.pts.p, n, cw, sw, c.wedge_rec.p);
        hip_check(hipGetLastError(), "launch k-d images");
        src = c.wedge_rec.p;
        n *= 3;
    }
    kd_tree &t = c.kd;
    t.n = n;
    t.node.alloc(n);
    t.idx.alloc(n);
    t.world.alloc(6);
    dbuf<float4> node_b;
    dbuf<int> idx_b;
    dbuf<uint32_t> tag, tag_b;
    dbuf<uint64_t> key;
    dbuf<int> val;
    dbuf<unsigned int> wbits;
    dbuf<int> dim0;
    node_b.alloc(n);
    idx_b.alloc(n);
    tag.alloc(n);
    tag_b.alloc(n);
    key.alloc(n);
    val.alloc(n);
    wbits.alloc(6);
    dim0.alloc(1);
    const unsigned int init[6] = {~0u, ~0u, ~0u, 0u, 0u, 0u};
    hip_check(hipMemcpyAsync(wbits.p, init, sizeof(init), hipMemcpyHostToDevice, st), "k-d world init");
    hip_check(hipMemsetAsync(tag.p, 0, sizeof(uint32_t)*n, st), "k-d tags");
    k_kd_init<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(src, n, t.node.p, t.idx.p, wbits.p);
    k_kd_world<<<1, 64, 0, st>>>(wbits.p, t.world.p, dim0.p);
    k_kd_dim0<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(t.node.p, n, dim0.p);
    hip_check(hipGetLastError(), "launch k-d init");
    int levels = 0;
    while ((1ll << levels) - 1 < n) ++levels;      // levels of a left-balanced tree of n nodes
    const int deepest = levels - 1;
    auto sort_level = [&](int L, bool tag_only) {
        k_kd_keys<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(t.node.p, tag.p, n, key.p, val.p, tag_only);
        // tags of step L are below 2^(L+1); the final pass sorts tags of up to `levels` bits
        int tag_bits = tag_only ? levels : L + 1;
        radix_sort_pairs(key.p, val.p, n, 0, (tag_only ? 0 : 32) + tag_bits, c.scratch, st);
        k_kd_gather<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(key.p, val.p, t.node.p, t.idx.p, n, tag_only,
                                                             node_b.p, idx_b.p, tag_b.p);
        hip_check(hipGetLastError(), "launch k-d sort");
        swap_buf(t.node, node_b);
        swap_buf(t.idx, idx_b);
        swap_buf(tag, tag_b);
    };
    for (int L = 0; L < deepest; ++L)
    {
        sort_level(L, false);
        k_kd_update<<<blocks(n, KD_BLOCK), KD_BLOCK, 0, st>>>(t.node.p, tag.p, n, L, t.world.p);
        hip_check(hipGetLastError(), "launch k-d update");
    }
    sort_level(deepest, true);
    hip_check(hipStreamSynchronize(st), "k-d build");
}

void kd_query(context &c)
{
    const int n = c.n;
    float image_dist_min = -1.0f;
    if constexpr (X_WEDGE)
    {
        // _get_image_dist_min (inc/swarm/param_grid.cuh:86-89), each sine rounded once from double
        auto sinf1 = [](float a) { return static_cast<float>(std::sin(static_cast<double>(a))); };
        const float x_min = static_cast<float>(X_MIN), x_max = static_cast<float>(X_MAX);
        image_dist_min = 2.0f*static_cast<float>(Y_MIN)
            *std::fmin(sinf1(static_cast<float>(Z_MIN)), sinf1(static_cast<float>(Z_MAX)))
            *std::fabs(sinf1(0.5f*(x_max - x_min)));
        c.nbr_image.alloc(size_t(std::max(1, c.n_active))*N_K);
    }
    k_kd_topk<<<blocks(n, KD_QUERY_BLOCK), KD_QUERY_BLOCK, 0, c.stream>>>(
        c.kd.node.p, c.kd.idx.p, c.kd.n, n, X_WEDGE ? n : 0, image_dist_min, c.kd.world.p, c.index.idx.p, c.pts.p,
        c.active.p, c.pidx_slot.p, c.group.p, c.user, c.nbr.p, c.measure.p, c.graph.p, c.overflow.p,
        c.nbr_image.p);
    hip_check(hipGetLastError(), "launch k-d top-K");
}

}  // namespace gdx
