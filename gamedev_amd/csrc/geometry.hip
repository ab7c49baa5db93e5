// Geometry epoch of the collision operator on gfx950: collision sites, Morton keys, stable radix
// sort, adaptive Morton hierarchy, exact top-K neighbour lists and the KNN ball measure.
//
// Reference behaviour followed: col_site_init (src/swarm/col_site_init.cu:79-118), morton_keygen and
// the hierarchy build (inc/swarm/morton/morton_index.cuh:41-229), the root cube of
// morton_ghost_index::build_index (inc/swarm/morton/morton_ghost.cuh:176-183), col_cache_get
// (inc/swarm/_col_cache.cuh:73-159) and col_space_bin / col_dep_graph (src/swarm/col_space_bin.cu,
// src/swarm/col_dep_graph.cu). The top-K contract is the reference's: the K smallest (d^2, code)
// among active records with d^2 <= q^2, ascending, where code = 3*particle index.
#include "gdx_internal.h"
#include "gdx_device.h"
#include "gdx_topk.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#pragma clang fp contract(off)

namespace gdx {

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw status_error(17, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ double ld(const field &f, long long i)
{
    return *reinterpret_cast<const double *>(f.base + i*f.stride);
}

// ============================================================================================
// device-wide exclusive scan of u32 (reduce, scan partials, downsweep)
// ============================================================================================
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK*SCAN_ITEMS;

__device__ __forceinline__ unsigned int block_exclusive(unsigned int v, unsigned int *lds_waves, unsigned int &total)
{
    int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    unsigned int inc = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1)
    {
        unsigned int t = __shfl_up(inc, o, WAVE);
        if (lane >= o) inc += t;
    }
    if (lane == WAVE - 1) lds_waves[wave] = inc;
    __syncthreads();
    unsigned int base = 0;
    total = 0;
    for (int w = 0; w < SCAN_BLOCK / WAVE; ++w)
    {
        unsigned int x = lds_waves[w];
        if (w < wave) base += x;
        total += x;
    }
    __syncthreads();
    return base + inc - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const unsigned int *in, int n, unsigned int *partial)
{
    __shared__ unsigned int waves[SCAN_BLOCK / WAVE];
    long long base = (long long)blockIdx.x*SCAN_TILE + (long long)threadIdx.x*SCAN_ITEMS;
    unsigned int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) if (base + i < n) s += in[base + i];
    unsigned int total;
    block_exclusive(s, waves, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_partials(unsigned int *partial, int m)
{
    __shared__ unsigned int waves[SCAN_BLOCK / WAVE];
    unsigned int carry = 0;
    for (int chunk = 0; chunk < m; chunk += SCAN_TILE)
    {
        unsigned int v[SCAN_ITEMS];
        unsigned int s = 0;
        int base = chunk + threadIdx.x*SCAN_ITEMS;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) { v[i] = (base + i < m) ? partial[base + i] : 0u; s += v[i]; }
        unsigned int total;
        unsigned int ex = block_exclusive(s, waves, total) + carry;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) { if (base + i < m) partial[base + i] = ex; ex += v[i]; }
        carry += total;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const unsigned int *in, unsigned int *out, int n,
                                                         const unsigned int *partial)
{
    __shared__ unsigned int waves[SCAN_BLOCK / WAVE];
    long long base = (long long)blockIdx.x*SCAN_TILE + (long long)threadIdx.x*SCAN_ITEMS;
    unsigned int v[SCAN_ITEMS];
    unsigned int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) { v[i] = (base + i < n) ? in[base + i] : 0u; s += v[i]; }
    unsigned int total;
    unsigned int ex = block_exclusive(s, waves, total) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) { if (base + i < n) out[base + i] = ex; ex += v[i]; }
}

// ============================================================================================
// stable LSD radix sort of (u64 key, int value), 8-bit digits, 4096-key tiles ranked in LDS
// ============================================================================================
constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_BLOCK*SORT_ITEMS;
constexpr int RADIX = 256;

__global__ __launch_bounds__(SORT_BLOCK) void k_radix_hist(const uint64_t *keys, int n, int shift, unsigned int mask,
                                                          unsigned int *hist, int nblocks)
{
    __shared__ unsigned int cnt[RADIX];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    long long base = (long long)blockIdx.x*SORT_TILE;
#pragma unroll 4
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        long long i = base + r*SORT_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&cnt[unsigned(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[(long long)threadIdx.x*nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(SORT_BLOCK) void k_radix_scatter(const uint64_t *keys_in, const int *vals_in, int n,
                                                             int shift, unsigned int mask, const unsigned int *offsets,
                                                             int nblocks, uint64_t *keys_out, int *vals_out)
{
    __shared__ unsigned int running[RADIX];
    __shared__ unsigned int wcnt[SORT_BLOCK / WAVE][RADIX];
    const int t = threadIdx.x, lane = t % WAVE, wave = t / WAVE;
    running[t] = offsets[(long long)t*nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < SORT_BLOCK / WAVE; ++w) wcnt[w][t] = 0;
    __syncthreads();
    long long base = (long long)blockIdx.x*SORT_TILE;
    for (int r = 0; r < SORT_ITEMS; ++r)
    {
        long long i = base + r*SORT_BLOCK + t;
        bool valid = i < n;
        uint64_t key = valid ? keys_in[i] : 0ull;
        int val = valid ? vals_in[i] : 0;
        unsigned int digit = unsigned(key >> shift) & mask;
        // lanes of this wavefront holding the same digit, in lane (= input) order
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            bool bit = (digit >> b) & 1u;
            unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int rank = __popcll(peers & lanemask_lt(lane));
        if (valid && rank == 0) wcnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid)
        {
            unsigned int pos = running[digit] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][digit];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        unsigned int add = 0;
#pragma unroll
        for (int w = 0; w < SORT_BLOCK / WAVE; ++w) { add += wcnt[w][t]; wcnt[w][t] = 0; }
        running[t] += add;
        __syncthreads();
    }
}

// ============================================================================================
// collision sites (src/swarm/col_site_init.cu:79-118) and spatial groups (col_space_bin.cu:11-35)
// ============================================================================================
__device__ __forceinline__ bool is_active(double y, double z)
{
    if (y < Y_MIN || y >= Y_MAX) return false;
#ifdef HALF_DISK
    if (N_Z > 1 && (z < Z_MIN || z > Z_MAX)) return false;
#else
    if (N_Z > 1 && (z < Z_MIN || z >= Z_MAX)) return false;
#endif
    return true;
}

__device__ __forceinline__ int space_bin(double x, double y, double z)
{
    int bin_x = 0;
    int bin_y = static_cast<int>(floor((y - Y_MIN) / (Y_MAX - Y_MIN)*COL_BIN_Y));
    int bin_z = 0;
    if constexpr (N_X > 1) bin_x = static_cast<int>(floor((x - X_MIN) / (X_MAX - X_MIN)*COL_BIN_X));
    if constexpr (N_Z > 1)
    {
        double extent_z = Z_MAX - Z_MIN;
        bin_z = static_cast<int>(floor((z - Z_MIN) / extent_z*COL_BIN_Z));
    }
    bin_x = (bin_x < 0) ? 0 : ((bin_x >= COL_BIN_X) ? COL_BIN_X - 1 : bin_x);
    bin_y = (bin_y < 0) ? 0 : ((bin_y >= COL_BIN_Y) ? COL_BIN_Y - 1 : bin_y);
    bin_z = (bin_z < 0) ? 0 : ((bin_z >= COL_BIN_Z) ? COL_BIN_Z - 1 : bin_z);
    constexpr int count_x = (N_X > 1) ? COL_BIN_X : 1;
    return bin_x + count_x*(bin_y + COL_BIN_Y*bin_z);
}

__global__ __launch_bounds__(256) void k_site_init(bound_state u, int n, float4 *pts, unsigned char *active,
                                                  unsigned char *group, int *bad)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = ld(u.x, i), y = ld(u.y, i), z = ld(u.z, i);
    bool finite = isfinite(x) && isfinite(y) && isfinite(z) && isfinite(ld(u.size, i)) && isfinite(ld(u.numr, i));
    if (u.vx.base) finite = finite && isfinite(ld(u.vx, i)) && isfinite(ld(u.vy, i)) && isfinite(ld(u.vz, i));
    if (!finite)
    {
        active[i] = 0;
        group[i] = 0;
        pts[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        atomicMin(bad, int(i) + 1);
        return;
    }
    bool act = is_active(y, z);
    double R = cyl_R(y, z);
    double Z = cyl_Z(y, z);
    float4 p;
    if (N_X == 1 && N_Z == 1) p = make_float4(static_cast<float>(R), 0.0f, 0.0f, 0.0f);
    else p = make_float4(static_cast<float>(R*cos(x)), static_cast<float>(R*sin(x)), static_cast<float>(Z), 0.0f);
    p.w = act ? static_cast<float>(H_SEARCH*hg(R)*R) : 0.0f;
    pts[i] = p;
    active[i] = static_cast<unsigned char>(act);
    group[i] = static_cast<unsigned char>(space_bin(x, y, z));
}

// ============================================================================================
// Morton keys and the adaptive hierarchy (inc/swarm/morton/morton_index.cuh:41-145)
// ============================================================================================
__device__ __forceinline__ uint64_t expand3(uint32_t v)
{
    uint64_t b = v & 0x1fffffu;
    b = (b | b << 32) & 0x001f00000000ffffULL;
    b = (b | b << 16) & 0x001f0000ff0000ffULL;
    b = (b | b << 8)  & 0x100f00f00f00f00fULL;
    b = (b | b << 4)  & 0x10c30c30c30c30c3ULL;
    b = (b | b << 2)  & 0x1249249249249249ULL;
    return b;
}

__global__ __launch_bounds__(256) void k_morton_keys(const float4 *pts, int n, float ox, float oy, float oz,
                                                    float width, int max_level, int dim, uint64_t *keys, int *vals)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    int cells = 1 << max_level;
    float scale = static_cast<float>(cells) / width;
    int ix = min(cells - 1, max(0, static_cast<int>((p.x - ox)*scale)));
    int iy = min(cells - 1, max(0, static_cast<int>((p.y - oy)*scale)));
    int iz = (dim == 2) ? 0 : min(cells - 1, max(0, static_cast<int>((p.z - oz)*scale)));
    keys[i] = expand3(uint32_t(ix)) | (expand3(uint32_t(iy)) << 1) | (expand3(uint32_t(iz)) << 2);
    vals[i] = int(i);
}

// sorted record m carries its code 3*i + image; the activity of physical particle i marks it
__global__ __launch_bounds__(256) void k_records(const int *sidx, const float4 *pts, const unsigned char *active,
                                                const int *codes, int n, float4 *rec)
{
    long long m = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= n) return;
    int p = sidx[m];
    float4 q = pts[p];
    int code = codes ? codes[p] : 3*p;
    if (active && !active[code / 3]) code |= int(0x80000000u);
    rec[m] = make_float4(q.x, q.y, q.z, __int_as_float(code));
}

// ============================================================================================
// azimuthal wedge: periodic ghosts (inc/swarm/morton/morton_ghost.cuh:28-86)
// ============================================================================================
// Cartesian distance from a record to the radial half-plane at azimuth offset `off` from it
// (_get_morton_seam_dist, morton_ghost.cuh:28-35)
__device__ __forceinline__ float seam_dist(float px, float py, float off)
{
    float R = hypotf(px, py);
    return (cosf(off) >= 0.0f) ? R*fabsf(sinf(off)) : R;
}

// largest search cap of the epoch (the max_element of swarm_runtime.cu:413-414); caps are >= 0
__global__ __launch_bounds__(256) void k_cap_max(const float4 *pts, int n, unsigned int *out)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    float w = i < n ? pts[i].w : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w = fmaxf(w, __shfl_xor(w, o, WAVE));
    if (threadIdx.x % WAVE == 0) atomicMax(out, __float_as_uint(w));
}

// ghosts per particle: one for each seam whose distance is within the largest cap (morton_gcount)
__global__ __launch_bounds__(256) void k_ghost_count(bound_state u, const float4 *pts, int n, float q_max,
                                                    float x_min, float x_max, unsigned int *count)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    float posx = static_cast<float>(ld(u.x, i));
    unsigned c = 0;
    if (seam_dist(p.x, p.y, posx - x_min) <= q_max) ++c;
    if (seam_dist(p.x, p.y, x_max - posx) <= q_max) ++c;
    count[i] = c;
}

// physical records first (code 3i), then each particle's ghosts at n + offset: image 2 rotated by
// +width near x_min, image 1 rotated by -width near x_max (morton_gwrite); the rotation's sine and
// cosine are computed once on the host
__global__ __launch_bounds__(256) void k_ghost_write(bound_state u, const float4 *pts, int n, const unsigned int *off,
                                                    float q_max, float x_min, float x_max, float cw, float sw,
                                                    float4 *rec, int *code)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    rec[i] = p;
    code[i] = 3*int(i);
    float posx = static_cast<float>(ld(u.x, i));
    long long g = n + off[i];
    if (seam_dist(p.x, p.y, posx - x_min) <= q_max)
    {
        rec[g] = make_float4(cw*p.x - sw*p.y, sw*p.x + cw*p.y, p.z, 0.0f);
        code[g] = 3*int(i) + 2;
        ++g;
    }
    if (seam_dist(p.x, p.y, x_max - posx) <= q_max)
    {
        rec[g] = make_float4(cw*p.x + sw*p.y, -sw*p.x + cw*p.y, p.z, 0.0f);
        code[g] = 3*int(i) + 1;
    }
}

__device__ __forceinline__ int child_end(const uint64_t *keys, int begin, int end, int shift, int code)
{
    while (begin < end)
    {
        int mid = begin + (end - begin) / 2;
        if (int((keys[mid] >> shift) & 7ULL) < code) begin = mid + 1;
        else end = mid;
    }
    return begin;
}

__global__ void k_root(gdx_node *nodes, int n, float ox, float oy, float oz, float width)
{
    nodes[0] = gdx_node{ox, oy, oz, width, 0, n, -1, 0};
}

__global__ __launch_bounds__(256) void k_node_counts(gdx_node *nodes, const uint64_t *keys, int first, int count,
                                                    int level, int max_level, int dim, int leaf_target,
                                                    unsigned int *children)
{
    int idx = blockIdx.x*blockDim.x + threadIdx.x;
    if (idx == 0) children[count] = 0;
    if (idx >= count) return;
    gdx_node &node = nodes[first + idx];
    int cc = 0;
    if (node.count > leaf_target && level < max_level)
    {
        int b = node.begin, e = b + node.count, shift = 3*(max_level - level - 1);
        for (int code = 1; code <= (1 << dim); ++code)
        {
            int next = child_end(keys, b, e, shift, code);
            cc += next > b;
            b = next;
        }
    }
    node.child_count = cc;
    children[idx] = cc;
}

__global__ __launch_bounds__(256) void k_node_expand(gdx_node *nodes, const uint64_t *keys, int first, int count,
                                                    int next_first, const unsigned int *offsets, int level,
                                                    int max_level, int dim)
{
    int idx = blockIdx.x*blockDim.x + threadIdx.x;
    if (idx >= count) return;
    gdx_node parent = nodes[first + idx];
    if (!parent.child_count) return;
    int b = parent.begin, e = b + parent.count;
    int slot = next_first + int(offsets[idx]);
    nodes[first + idx].first_child = slot;
    int shift = 3*(max_level - level - 1);
    float w = 0.5f*parent.width;
    int mask = 0;
    for (int code = 0; code < (1 << dim); ++code)
    {
        int next = child_end(keys, b, e, shift, code + 1);
        if (next > b)
        {
            float lx = parent.lx, ly = parent.ly, lz = parent.lz;
            if (code & 1) lx += w;
            if (code & 2) ly += w;
            if (dim == 3 && (code & 4)) lz += w;
            nodes[slot++] = gdx_node{lx, ly, lz, w, b, next - b, -1, 0};
            mask |= 1 << code;
        }
        b = next;
    }
    // child codes in bits 8..15 beside the child count (bits 0..3)
    nodes[first + idx].child_count = parent.child_count | (mask << 8);
}

// ============================================================================================
// exact top-K: one lane per query, 64 Morton-consecutive queries per wavefront share one
// traversal; leaf records are loaded as 64-record tiles and broadcast lane by lane
// ============================================================================================
constexpr int TOPK_BLOCK = 256;
constexpr int TOPK_STACK = 256;

__device__ __forceinline__ bool nb_less(float da, int ia, float db, int ib)   // morton_types.cuh:68-72
{ return da < db || (da == db && ia < ib); }

// padded point-to-cell lower distance (morton_index.cuh:296-314)
__device__ __forceinline__ float node_d2(float qx, float qy, float qz, const gdx_node &nd, int dim, int max_level)
{
    float ux = nd.lx + nd.width, uy = nd.ly + nd.width, uz = nd.lz + nd.width;
    float scale = fmaxf(1.0f, fmaxf(fmaxf(fabsf(nd.lx), fabsf(ux)),
        fmaxf(fmaxf(fabsf(nd.ly), fabsf(uy)), fmaxf(fabsf(nd.lz), fabsf(uz)))));
    float pad = 2.0f*static_cast<float>(max_level + 2)*FLT_EPSILON*scale;
    float dx = fmaxf(fmaxf(nd.lx - qx - pad, 0.0f), qx - ux - pad);
    float dy = fmaxf(fmaxf(nd.ly - qy - pad, 0.0f), qy - uy - pad);
    float dz = (dim == 2) ? 0.0f : fmaxf(fmaxf(nd.lz - qz - pad, 0.0f), qz - uz - pad);
    return dx*dx + dy*dy + dz*dz;
}

// insert (nd, nc) into the ascending list; precondition: it precedes the current last entry
template <int K>
__device__ __forceinline__ void topk_insert(float (&d)[K], int (&c)[K], float nd, int nc)
{
    bool carry = true;
#pragma unroll
    for (int k = K - 1; k > 0; --k)
    {
        bool lt = nb_less(nd, nc, d[k - 1], c[k - 1]);
        float dk = lt ? d[k - 1] : (carry ? nd : d[k]);
        int ck = lt ? c[k - 1] : (carry ? nc : c[k]);
        d[k] = dk;
        c[k] = ck;
        carry = lt;
    }
    if (carry) { d[0] = nd; c[0] = nc; }
}

// insert (nd, nc) keeping one entry per physical particle (code / 3): an entry of the same particle is
// replaced when the candidate precedes it, else the candidate is dropped. The list stays the K first of
// the candidates seen so far with each particle collapsed to its nearest image, which is what the
// reference's 3K candidates with duplicate removal (morton_query.cuh:62-140) select: every record up
// to the K-th particle's distance belongs to one of those K particles, so at most 3K records precede
// it. Precondition as topk_insert.
template <int K>
__device__ __forceinline__ void topk_insert_dedup(float (&d)[K], int (&c)[K], float nd, int nc)
{
    const unsigned pn = static_cast<unsigned>(nc) / 3u;
    int e = -1;
    float de = 0.0f;
    int ce = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        const bool same = e < 0 && c[k] != INT_MAX && static_cast<unsigned>(c[k]) / 3u == pn;
        e = same ? k : e;
        de = same ? d[k] : de;
        ce = same ? c[k] : ce;
    }
    if (e >= 0)
    {
        if (!nb_less(nd, nc, de, ce)) return;
#pragma unroll
        for (int k = 0; k < K - 1; ++k)
        {
            d[k] = k >= e ? d[k + 1] : d[k];
            c[k] = k >= e ? c[k + 1] : c[k];
        }
        d[K - 1] = __builtin_inff();
        c[K - 1] = INT_MAX;
    }
    topk_insert<K>(d, c, nd, nc);
}

__device__ __forceinline__ float readlane_f(float v, int l)
{ return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// Traversal shared by both selection schemes: the wavefront pops nodes from one LDS stack, children
// nearest-first from the wavefront's centre, and pushes only the nodes some lane needs; every needed leaf
// goes to leaf(begin, end, need) with need the lane's own test. cut() is the lane's current pruning
// distance (the smaller of its cap and its K-th candidate so far).
template <class Cut, class Leaf>
__device__ __forceinline__ void traverse(const gdx_node *nodes, int dim, int max_level, float qx, float qy, float qz,
                                         bool qactive, float cx, float cy, float cz, int *stack,
                                         unsigned int *overflow, Cut cut_of, Leaf leaf)
{
    const int lane = threadIdx.x % WAVE;
    if (__ballot(qactive) == 0ull) return;
    int sp = 1;
    if (lane == 0) stack[0] = 0;
    __builtin_amdgcn_wave_barrier();
    while (sp > 0)
    {
        --sp;
        int id = __builtin_amdgcn_readfirstlane(stack[sp]);
        gdx_node nd = nodes[id];
        float cut = cut_of();
        bool need = qactive && node_d2(qx, qy, qz, nd, dim, max_level) <= cut;
        if (__ballot(need) == 0ull) continue;
        if (nd.child_count == 0)
        {
            leaf(nd.begin, nd.begin + nd.count, need);
            continue;
        }
        // children some lane needs, each child's distance from the wavefront's centre kept by the lane of
        // its index; the needed children go on the stack farthest first (nearest on top, equal distances
        // in child order), each lane writing its own child at its rank
        const int nch = nd.child_count & 15;
        unsigned needed = 0u;
        float kd = __builtin_inff();
        for (int ch = 0; ch < nch; ++ch)
        {
            gdx_node cn = nodes[nd.first_child + ch];
            bool cneed = qactive && node_d2(qx, qy, qz, cn, dim, max_level) <= cut;
            if (__ballot(cneed) == 0ull) continue;
            needed |= 1u << ch;
            const float k = node_d2(cx, cy, cz, cn, dim, max_level);
            if (lane == ch) kd = k;
        }
        int nv = __popc(needed);
        if (sp + nv > TOPK_STACK)
        {
            if (lane == 0) atomicOr(overflow, 1u);
            nv = 0;
        }
        // ranks by whole-wavefront shuffles (every lane active), then each needed child's lane writes it
        int rank = 0;
        for (int m = 0; m < nch; ++m)
        {
            const float km = __shfl(kd, m, WAVE);
            rank += ((needed >> m) & 1u) && (km < kd || (km == kd && m < lane));
        }
        if (nv > 0 && lane < nch && ((needed >> lane) & 1u)) stack[sp + nv - 1 - rank] = nd.first_child + lane;
        sp += nv;
        __builtin_amdgcn_wave_barrier();
    }
}

// insertion scheme: every eligible record goes through the lane's sorted K-entry register list
template <int K, bool DEDUP = false>
__device__ void topk_wave(const float4 *rec, const gdx_node *nodes, int dim, int max_level, float qx, float qy,
                          float qz, float q2, bool qactive, float cx, float cy, float cz, float (&d)[K], int (&c)[K],
                          int *stack, unsigned int *overflow)
{
    const int lane = threadIdx.x % WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) { d[k] = __builtin_inff(); c[k] = INT_MAX; }
    traverse(nodes, dim, max_level, qx, qy, qz, qactive, cx, cy, cz, stack, overflow,
             [&] { return fminf(q2, d[K - 1]); },
             [&](int begin, int end, bool need) {
                 for (int t0 = begin; t0 < end; t0 += WAVE)
                 {
                     int idx = t0 + lane;
                     float4 v = idx < end ? rec[idx] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
                     int cnt = min(WAVE, end - t0);
                     for (int j = 0; j < cnt; ++j)
                     {
                         float px = readlane_f(v.x, j), py = readlane_f(v.y, j), pz = readlane_f(v.z, j);
                         int code = __builtin_amdgcn_readlane(__float_as_int(v.w), j);
                         float dx = qx - px, dy = qy - py, dz = qz - pz;
                         float d2 = dx*dx + dy*dy + dz*dz;
                         bool e = need && code >= 0 && d2 <= q2 && nb_less(d2, code, d[K - 1], c[K - 1]);
                         if (e)
                         {
                             if constexpr (DEDUP) topk_insert_dedup<K>(d, c, d2, code);
                             else topk_insert<K>(d, c, d2, code);
                         }
                     }
                 }
             });
}

// ---- tiled scheme: candidates collected per lane in LDS, merged B at a time (gdx_topk.h) ----
// the lane's `cnt` buffered keys buf[r * WAVE] merged into L
template <int K, int B>
__device__ __forceinline__ void topk_flush(uint64_t (&L)[K], const uint64_t *buf, int cnt)
{
    uint64_t S[B];
#pragma unroll
    for (int r = 0; r < B; ++r)
    {
        const uint64_t v = buf[r*WAVE];
        S[r] = r < cnt ? v : TOPK_EMPTY;
    }
    bitonic_sort<B>(S);
    topk_merge_sorted<K, B>(L, S);
}

// tiled scheme: a record passes the lane's filter when it lies within the cap and precedes the lane's
// K-th key as of its last merge (a stale, hence larger, cutoff: the filter keeps a superset of the
// records the exact list would take); passing keys are appended to the lane's column of an LDS buffer
// and the wavefront merges every lane's buffer once some lane holds B. The K smallest keys of a union do
// not depend on the order of merging, so the result is the insertion scheme's, bit for bit.
// records per overflow check of the tiled scheme: a merge runs before a group when some lane could not
// take all of it
constexpr int TILE_GROUP = 4;

template <int K, int B>
__device__ void topk_wave_tiled(const float4 *__restrict__ rec, const gdx_node *__restrict__ nodes, int dim, int max_level, float qx, float qy,
                                float qz, float q2, bool qactive, float cx, float cy, float cz, float (&d)[K],
                                int (&c)[K], int *stack, uint64_t *buf, unsigned int *overflow)
{
    const int lane = threadIdx.x % WAVE;
    uint64_t L[K];
#pragma unroll
    for (int k = 0; k < K; ++k) L[k] = TOPK_EMPTY;
    uint64_t *col = buf + lane;
    int cnt = 0;
    traverse(nodes, dim, max_level, qx, qy, qz, qactive, cx, cy, cz, stack, overflow,
             [&] { return fminf(q2, __uint_as_float(uint32_t(L[K - 1] >> 32))); },
             [&](int begin, int end, bool need) {
                 // records at uniform addresses (scalar loads, four at a time); the filter is the lane's
                 // cap and the distance of its K-th key as of its last merge (lanes that do not need the
                 // leaf take nothing: d^2 >= 0 > -1); ties at that distance pass and the merge decides them
                 const float cap = need ? q2 : -1.0f;
                 float lim = fminf(cap, __uint_as_float(uint32_t(L[K - 1] >> 32)));
                 for (int t = begin; t < end; t += TILE_GROUP)
                 {
                     if (__ballot(cnt > B - TILE_GROUP))
                     {
                         topk_flush<K, B>(L, col, cnt);
                         cnt = 0;
                         lim = fminf(cap, __uint_as_float(uint32_t(L[K - 1] >> 32)));
                     }
                     // the group's loads first (indices clamped into the leaf), then its tests
                     float4 v[TILE_GROUP];
#pragma unroll
                     for (int u = 0; u < TILE_GROUP; ++u) v[u] = rec[min(t + u, end - 1)];
#pragma unroll
                     for (int u = 0; u < TILE_GROUP; ++u)
                     {
                         const int code = __float_as_int(v[u].w);
                         const float dx = qx - v[u].x, dy = qy - v[u].y, dz = qz - v[u].z;
                         const float d2 = dx*dx + dy*dy + dz*dz;
                         if (t + u < end && code >= 0 && d2 <= lim)
                         {
                             col[cnt*WAVE] = (uint64_t(__float_as_uint(d2)) << 32) | uint32_t(code);
                             ++cnt;
                         }
                     }
                 }
             });
    if (__ballot(cnt > 0)) topk_flush<K, B>(L, col, cnt);
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        d[k] = __uint_as_float(uint32_t(L[k] >> 32));
        c[k] = int(uint32_t(L[k]));
    }
}

__device__ __forceinline__ void wave_centre(float qx, float qy, float qz, bool qactive, float &cx, float &cy, float &cz)
{
    float lo[3] = {qactive ? qx : __builtin_inff(), qactive ? qy : __builtin_inff(), qactive ? qz : __builtin_inff()};
    float hi[3] = {qactive ? qx : -__builtin_inff(), qactive ? qy : -__builtin_inff(), qactive ? qz : -__builtin_inff()};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
#pragma unroll
        for (int m = 32; m > 0; m >>= 1)
        {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], m, WAVE));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], m, WAVE));
        }
    }
    cx = 0.5f*(lo[0] + hi[0]);
    cy = 0.5f*(lo[1] + hi[1]);
    cz = 0.5f*(lo[2] + hi[2]);
}

// col_cache_get (inc/swarm/_col_cache.cuh:73-159) over Morton-ordered queries
// one query per physical record; DEDUP when two images of one particle can share a search ball
// TILE = 0: the insertion scheme; TILE = B > 0: the tiled scheme with B-key merges (no image dedup);
// WPB wavefronts per block (the block shape: every wavefront runs its own queries with its own LDS
// stack and tile, so the shape changes only how the wavefronts are packed onto the CUs)
template <bool DEDUP, int TILE, int WPB>
__global__ __launch_bounds__(WPB*WAVE) void k_col_topk(const float4 *__restrict__ rec,
                                                        const gdx_node *__restrict__ nodes, int n, int dim,
                                                        int max_level, const int *__restrict__ sidx,
                                                        const float4 *__restrict__ pts,
                                                        const int *__restrict__ pidx_slot,
                                                        const unsigned char *__restrict__ group,
                                                        bound_state u, int *__restrict__ nbr,
                                                        double *__restrict__ measure,
                                                        unsigned long long *__restrict__ graph,
                                                        unsigned int *__restrict__ overflow,
                                                        unsigned char *__restrict__ nbr_image)
{
    static_assert(!(DEDUP && TILE), "image dedup runs on the insertion scheme");
    static_assert(TILE == 0 || TILE >= TILE_GROUP, "a merge tile holds at least one record group");
    __shared__ int stacks[WPB][TOPK_STACK];
    __shared__ uint64_t tiles[WPB][TILE > 0 ? TILE : 1][WAVE];
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    long long m = ((long long)xcd_block()*WPB + wave)*WAVE + lane;
    bool valid = m < n;
    float4 me = valid ? rec[m] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    int code = __float_as_int(me.w);
    bool qactive = valid && code >= 0 && (!X_WEDGE || code % 3 == 0);
    int p = valid ? sidx[m] : 0;
    float q = qactive ? pts[p].w : 0.0f;
    float q2 = q*q;
    float cx, cy, cz;
    wave_centre(me.x, me.y, me.z, qactive, cx, cy, cz);
    float d[N_K];
    int c[N_K];
    if constexpr (TILE > 0)
        topk_wave_tiled<N_K, TILE>(rec, nodes, dim, max_level, me.x, me.y, me.z, q2, qactive, cx, cy, cz, d, c,
                                   stacks[wave], &tiles[wave][0][0], overflow);
    else
        topk_wave<N_K, DEDUP>(rec, nodes, dim, max_level, me.x, me.y, me.z, q2, qactive, cx, cy, cz, d, c,
                              stacks[wave], overflow);
    unsigned long long mask = 0ull;
    int g = qactive ? int(group[p]) : -1;
    if (qactive)
    {
        int s = pidx_slot[p];
        float max_d2 = 0.0f;
        int row[N_K];
#pragma unroll
        for (int k = 0; k < N_K; ++k)
        {
            if (c[k] != INT_MAX)
            {
                max_d2 = fmaxf(max_d2, d[k]);
                int j = c[k] / 3;
                row[k] = pidx_slot[j];
                mask |= 1ull << group[j];
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
                const int k = bitrev_slot(t);
                nbr_image[(long long)s*N_K + t] = static_cast<unsigned char>(c[k] != INT_MAX ? c[k] % 3 : 0);
            }
        }
        double radius = sqrt(static_cast<double>(max_d2));
        measure[s] = ball_measure(ld(u.y, p), ld(u.z, p), radius);
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

// runtime-K exact top-K for arbitrary query points (test access to the search contract)
template <int K>
__global__ __launch_bounds__(TOPK_BLOCK) void k_knn(const float4 *rec, const gdx_node *nodes, int dim, int max_level,
                                                   const float4 *queries, int nq, int *codes, float *dist,
                                                   unsigned int *overflow)
{
    __shared__ int stacks[TOPK_BLOCK / WAVE][TOPK_STACK];
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    long long m = ((long long)blockIdx.x*(TOPK_BLOCK / WAVE) + wave)*WAVE + lane;
    bool valid = m < nq;
    float4 qp = valid ? queries[m] : make_float4(0.f, 0.f, 0.f, 0.f);
    float q2 = qp.w*qp.w;
    float cx, cy, cz;
    wave_centre(qp.x, qp.y, qp.z, valid, cx, cy, cz);
    float d[K];
    int c[K];
    topk_wave<K>(rec, nodes, dim, max_level, qp.x, qp.y, qp.z, q2, valid, cx, cy, cz, d, c, stacks[wave], overflow);
    if (!valid) return;
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        codes[m*K + k] = c[k] == INT_MAX ? -1 : c[k];
        dist[m*K + k] = d[k];
    }
}

// ============================================================================================
// slot order and per-slot state
// ============================================================================================
__global__ __launch_bounds__(256) void k_slot_keys(const int *sidx, const unsigned char *active,
                                                  const unsigned char *group, int nr, int n, uint64_t *keys, int *vals)
{
    long long m = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (m >= nr) return;
    int p = sidx[m];
    keys[m] = p >= n ? 256ull : active[p] ? uint64_t(group[p]) : 255ull;
    vals[m] = p;
}

__global__ __launch_bounds__(256) void k_slot_maps(const int *order, const unsigned char *group,
                                                  const unsigned char *active, int n, int *slot_pidx, int *pidx_slot,
                                                  int *group_active, int *group_all)
{
    // per-block histograms: one global atomic per (block, group) instead of one per particle
    __shared__ int hist[2][LOCAL_GROUPS];
    for (int i = threadIdx.x; i < 2*LOCAL_GROUPS; i += blockDim.x) (&hist[0][0])[i] = 0;
    __syncthreads();
    long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (s < n)
    {
        int p = order[s];
        int g = group[p];
        if (active[p])
        {
            slot_pidx[s] = p;
            pidx_slot[p] = int(s);
            atomicAdd(&hist[0][g], 1);
        }
        else
        {
            pidx_slot[p] = -1;
        }
        atomicAdd(&hist[1][g], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < LOCAL_GROUPS; i += blockDim.x)
    {
        if (hist[0][i]) atomicAdd(group_active + i, hist[0][i]);
        if (hist[1][i]) atomicAdd(group_all + i, hist[1][i]);
    }
}

// the state in slot order, published as it is (the operator's publish writes only slots changed since)
__global__ __launch_bounds__(256) void k_gather_state(bound_state u, const int *slot_pidx, int n_active,
                                                     const uint64_t *rng_p, double2 *st, double2 *res, float2 *res32,
                                                     uint64_t *rng, gdx_env *env, double *hgs)
{
    long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (s >= n_active) return;
    int p = slot_pidx[s];
    const double2 v = make_double2(ld(u.size, p), ld(u.numr, p));
    st[s] = v;
    res[s] = v;
    res32[s] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
    rng[s] = rng_p[p];
    if constexpr (COAG_KERNEL == 3)
    {
        double y = ld(u.y, p), z = ld(u.z, p);
        env[s] = make_env(y, z);
        if constexpr (PARTNER_HG)
        {
            double R = cyl_R(y, z);
            hgs[s] = R*hg(R);
        }
    }
}

__global__ __launch_bounds__(256) void k_scatter_state(bound_state u, const int *slot_pidx, int n_active,
                                                      const double2 *st, const uint64_t *rng, uint64_t *rng_p)
{
    long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (s >= n_active) return;
    int p = slot_pidx[s];
    double2 v = st[s];
    *reinterpret_cast<double *>(const_cast<char *>(u.size.base) + (long long)p*u.size.stride) = v.x;
    *reinterpret_cast<double *>(const_cast<char *>(u.numr.base) + (long long)p*u.numr.stride) = v.y;
    rng_p[p] = rng[s];
}

// colstate_flag (src/swarm/colstate_flag.cu:13-29): the finiteness screen of a reused epoch
__global__ __launch_bounds__(256) void k_state_flag(bound_state u, int n, int *bad)
{
    long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool finite = isfinite(ld(u.x, i)) && isfinite(ld(u.y, i)) && isfinite(ld(u.z, i)) && isfinite(ld(u.size, i))
        && isfinite(ld(u.numr, i));
    if (u.vx.base) finite = finite && isfinite(ld(u.vx, i)) && isfinite(ld(u.vy, i)) && isfinite(ld(u.vz, i));
    if (!finite) atomicMin(bad, int(i) + 1);
}

inline int blocks_for(long long n, int per) { return int((n + per - 1) / per); }

}  // namespace

// first non-finite particle + 1, or 0
int state_flag(context &c)
{
    c.bad.alloc(1);
    int big = INT_MAX;
    hip_check(hipMemcpyAsync(c.bad.p, &big, sizeof(int), hipMemcpyHostToDevice, c.stream), "reset bad");
    k_state_flag<<<blocks_for(c.n, 256), 256, 0, c.stream>>>(c.user, c.n, c.bad.p);
    hip_check(hipGetLastError(), "launch state flag");
    int bad = 0;
    hip_check(hipMemcpyAsync(&bad, c.bad.p, sizeof(int), hipMemcpyDeviceToHost, c.stream), "read bad");
    hip_check(hipStreamSynchronize(c.stream), "state flag");
    return bad == INT_MAX ? 0 : bad;
}

// ================================================================================================
// host side
// ================================================================================================
void scan_exclusive_u32(const unsigned int *in, unsigned int *out, int n, sort_scratch &s, hipStream_t st)
{
    if (n <= 0) return;
    int nb = blocks_for(n, SCAN_TILE);
    s.partial.alloc(size_t(nb));
    k_scan_reduce<<<nb, SCAN_BLOCK, 0, st>>>(in, n, s.partial.p);
    k_scan_partials<<<1, SCAN_BLOCK, 0, st>>>(s.partial.p, nb);
    k_scan_down<<<nb, SCAN_BLOCK, 0, st>>>(in, out, n, s.partial.p);
    hip_check(hipGetLastError(), "launch scan");
}

void radix_sort_pairs(uint64_t *keys, int *vals, int n, int begin_bit, int end_bit, sort_scratch &s, hipStream_t st)
{
    if (n <= 1) return;
    int nb = blocks_for(n, SORT_TILE);
    s.keys_alt.alloc(size_t(n));
    s.vals_alt.alloc(size_t(n));
    s.hist.alloc(size_t(RADIX)*nb);
    uint64_t *kin = keys, *kout = s.keys_alt.p;
    int *vin = vals, *vout = s.vals_alt.p;
    for (int shift = begin_bit; shift < end_bit; shift += 8)
    {
        int bits = std::min(8, end_bit - shift);
        unsigned int mask = (1u << bits) - 1u;
        k_radix_hist<<<nb, SORT_BLOCK, 0, st>>>(kin, n, shift, mask, s.hist.p, nb);
        scan_exclusive_u32(s.hist.p, s.hist.p, RADIX*nb, s, st);
        k_radix_scatter<<<nb, SORT_BLOCK, 0, st>>>(kin, vin, n, shift, mask, s.hist.p, nb, kout, vout);
        hip_check(hipGetLastError(), "launch radix pass");
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys)
    {
        hip_check(hipMemcpyAsync(keys, kin, sizeof(uint64_t)*n, hipMemcpyDeviceToDevice, st), "sort copy-back");
        hip_check(hipMemcpyAsync(vals, vin, sizeof(int)*n, hipMemcpyDeviceToDevice, st), "sort copy-back");
    }
}

// morton_index::build (inc/swarm/morton/morton_index.cuh:168-229)
void morton_build(morton_index &ix, const float4 *points, const unsigned char *active, int n, float ox, float oy,
                  float oz, float width, int dim, int leaf_target, int max_level, sort_scratch &s, hipStream_t st,
                  const int *codes)
{
    if (n <= 0 || !(width > 0.0f)) throw status_error(12, "invalid adaptive Morton size");
    if (dim != 2 && dim != 3) throw status_error(12, "adaptive Morton dimension must be 2 or 3");
    if (leaf_target <= 0) throw status_error(12, "adaptive Morton leaf target must be positive");
    if (max_level <= 0 || max_level > 20) throw status_error(12, "adaptive Morton max level must be 1 through 20");
    ix.n = n; ix.dim = dim; ix.max_level = max_level; ix.leaf_target = leaf_target;
    ix.ox = ox; ix.oy = oy; ix.oz = oz; ix.width = width;
    ix.keys.alloc(size_t(n));
    ix.idx.alloc(size_t(n));
    ix.rec.alloc(size_t(n));
    k_morton_keys<<<blocks_for(n, 256), 256, 0, st>>>(points, n, ox, oy, oz, width, max_level, dim, ix.keys.p, ix.idx.p);
    hip_check(hipGetLastError(), "launch morton keys");
    radix_sort_pairs(ix.keys.p, ix.idx.p, n, 0, 3*max_level, s, st);
    k_records<<<blocks_for(n, 256), 256, 0, st>>>(ix.idx.p, points, active, codes, n, ix.rec.p);
    hip_check(hipGetLastError(), "launch morton records");

    size_t cap = std::max<size_t>(1024, size_t(n) / size_t(std::max(1, leaf_target))*4 + 64);
    if (ix.nodes.n < cap) ix.nodes.alloc(cap);
    k_root<<<1, 1, 0, st>>>(ix.nodes.p, n, ox, oy, oz, width);
    int node_count = 1, first = 0, count = 1;
    dbuf<unsigned int> children, offsets;
    for (int level = 0; level <= max_level && count; ++level)
    {
        children.alloc(size_t(count) + 1);
        offsets.alloc(size_t(count) + 1);
        k_node_counts<<<blocks_for(count, 256), 256, 0, st>>>(ix.nodes.p, ix.keys.p, first, count, level, max_level,
                                                             dim, leaf_target, children.p);
        scan_exclusive_u32(children.p, offsets.p, count + 1, s, st);
        unsigned int next = 0;
        hip_check(hipMemcpyAsync(&next, offsets.p + count, sizeof(next), hipMemcpyDeviceToHost, st), "frontier count");
        hip_check(hipStreamSynchronize(st), "frontier sync");
        if (!next) break;
        if (uint64_t(next) > uint64_t(INT_MAX - node_count)) throw status_error(11, "Morton node index overflow");
        size_t need = size_t(node_count) + next;
        if (need > ix.nodes.n) ix.nodes.grow(std::max(need, 2*ix.nodes.n), size_t(node_count), st);
        k_node_expand<<<blocks_for(count, 256), 256, 0, st>>>(ix.nodes.p, ix.keys.p, first, count, node_count,
                                                             offsets.p, level, max_level, dim);
        hip_check(hipGetLastError(), "launch morton expand");
        first = node_count;
        node_count += int(next);
        count = int(next);
    }
    ix.node_count = node_count;
    hip_check(hipStreamSynchronize(st), "complete Morton hierarchy");
}

void knn_query(const morton_index &ix, const float4 *queries, int nq, int k, int *codes, float *d2,
               unsigned int *overflow, hipStream_t st)
{
    if (nq <= 0) return;
    int nb = blocks_for(nq, TOPK_BLOCK);
    switch (k)
    {
    case 1: k_knn<1><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 2: k_knn<2><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 4: k_knn<4><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 8: k_knn<8><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 16: k_knn<16><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 32: k_knn<32><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    case 64: k_knn<64><<<nb, TOPK_BLOCK, 0, st>>>(ix.rec.p, ix.nodes.p, ix.dim, ix.max_level, queries, nq, codes, d2, overflow); break;
    default: throw status_error(15, "knn query K must be 1, 2, 4, 8, 16, 32 or 64");
    }
    hip_check(hipGetLastError(), "launch knn");
}

// the largest search cap and the unique-identifier test of swarm_runtime.cu:331-340, 413-415
static void wedge_cap(context &c)
{
    const int n = c.n;
    hipStream_t st = c.stream;
    const float x_min = static_cast<float>(X_MIN), x_max = static_cast<float>(X_MAX);
    dbuf<unsigned int> cap;
    cap.alloc(1);
    hip_check(hipMemsetAsync(cap.p, 0, sizeof(unsigned int), st), "zero cap maximum");
    k_cap_max<<<blocks_for(n, 256), 256, 0, st>>>(c.pts.p, n, cap.p);
    unsigned int bits = 0;
    hip_check(hipMemcpyAsync(&bits, cap.p, sizeof bits, hipMemcpyDeviceToHost, st), "read cap maximum");
    hip_check(hipStreamSynchronize(st), "cap maximum");
    std::memcpy(&c.q_max, &bits, sizeof bits);
    // _get_image_dist_min (inc/swarm/param_grid.cuh:86-89): the closest any two images of one particle get
    // (each sine rounded once from double to float, as in the ghost rotation below)
    auto sinf1 = [](float a) { return static_cast<float>(std::sin(static_cast<double>(a))); };
    const float image_dist_min = 2.0f*static_cast<float>(Y_MIN)
        *std::fmin(sinf1(static_cast<float>(Z_MIN)), sinf1(static_cast<float>(Z_MAX)))
        *std::fabs(sinf1(0.5f*(x_max - x_min)));
    c.unique_ids = image_dist_min < 0.0f || image_dist_min > 2.0f*c.q_max;
}

// morton_ghost_index::build with ghosts (morton_ghost.cuh:96-168): count, scan, write physical records
// then ghosts
static void wedge_records(context &c)
{
    const int n = c.n;
    hipStream_t st = c.stream;
    const float x_min = static_cast<float>(X_MIN), x_max = static_cast<float>(X_MAX);
    c.ghost_count.alloc(size_t(n));
    c.ghost_off.alloc(size_t(n));
    k_ghost_count<<<blocks_for(n, 256), 256, 0, st>>>(c.user, c.pts.p, n, c.q_max, x_min, x_max, c.ghost_count.p);
    hip_check(hipGetLastError(), "launch ghost count");
    scan_exclusive_u32(c.ghost_count.p, c.ghost_off.p, n, c.scratch, st);
    unsigned int last[2] = {0u, 0u};
    hip_check(hipMemcpyAsync(&last[0], c.ghost_count.p + (n - 1), sizeof(unsigned int), hipMemcpyDeviceToHost, st),
              "read last ghost count");
    hip_check(hipMemcpyAsync(&last[1], c.ghost_off.p + (n - 1), sizeof(unsigned int), hipMemcpyDeviceToHost, st),
              "read last ghost offset");
    hip_check(hipStreamSynchronize(st), "ghost offsets");
    const long long total = (long long)n + last[0] + last[1];
    if (total > INT_MAX / 3) throw status_error(12, "wedge record count exceeds the index limit");
    c.n_records = int(total);
    c.wedge_rec.alloc(size_t(total));
    c.wedge_code.alloc(size_t(total));
    const float width = x_max - x_min;
    const float sw = static_cast<float>(std::sin(static_cast<double>(width)));
    const float cw = static_cast<float>(std::cos(static_cast<double>(width)));
    k_ghost_write<<<blocks_for(n, 256), 256, 0, st>>>(c.user, c.pts.p, n, c.ghost_off.p, c.q_max, x_min, x_max, cw,
                                                     sw, c.wedge_rec.p, c.wedge_code.p);
    hip_check(hipGetLastError(), "launch ghost write");
}

// the evolve_collisions geometry branch (src/swarm/swarm_runtime.cu:384-448) for bound particle state
void geometry_epoch(context &c)
{
    const int n = c.n;
    hipStream_t st = c.stream;
    c.geometry_valid = false;
    c.state_loaded = false;
    c.pts.alloc(n);
    c.active.alloc(n);
    c.group.alloc(n);
    c.bad.alloc(1);
    c.overflow.alloc(1);
    c.graph.alloc(LOCAL_GROUPS);
    c.pidx_slot.alloc(n);
    c.slot_pidx.alloc(n);
    int big = INT_MAX;
    hip_check(hipMemcpyAsync(c.bad.p, &big, sizeof(int), hipMemcpyHostToDevice, st), "reset bad");
    k_site_init<<<blocks_for(n, 256), 256, 0, st>>>(c.user, n, c.pts.p, c.active.p, c.group.p, c.bad.p);
    hip_check(hipGetLastError(), "launch site init");
    int bad = 0;
    hip_check(hipMemcpyAsync(&bad, c.bad.p, sizeof(int), hipMemcpyDeviceToHost, st), "read bad");
    hip_check(hipStreamSynchronize(st), "site init");
    if (bad != INT_MAX)
        throw status_error(14, "non-finite particle state before collision search at particle " + std::to_string(bad - 1));

    // root cube of morton_ghost_index::build_index: width 2.0002*Y_MAX centred on the origin
    float root_width = 2.0002f*static_cast<float>(Y_MAX);
    float o = -0.5f*root_width;
    c.n_records = n;
    c.unique_ids = true;
    // the Morton search of a wedge indexes seam ghosts; the k-d search adds its own images (kd_build)
    if (X_WEDGE) wedge_cap(c);
    if (X_WEDGE && c.search == GDX_COL_SEARCH_MORTON)
    {
        wedge_records(c);
        morton_build(c.index, c.wedge_rec.p, c.active.p, c.n_records, o, o, o, root_width, DIM, c.leaf_target,
                     MORTON_MAX_LEVEL, c.scratch, st, c.wedge_code.p);
    }
    else
    {
        morton_build(c.index, c.pts.p, c.active.p, n, o, o, o, root_width, DIM, c.leaf_target, MORTON_MAX_LEVEL,
                     c.scratch, st);
    }
    const int nr = c.n_records;

    // slot order: active owners grouped by controller group, Morton order inside each group; ghost
    // records sort last (key 256) and are cut off
    dbuf<uint64_t> gkeys;
    dbuf<int> gvals;
    dbuf<int> gcount;
    gkeys.alloc(nr);
    gvals.alloc(nr);
    gcount.alloc(2*LOCAL_GROUPS);
    hip_check(hipMemsetAsync(gcount.p, 0, sizeof(int)*2*LOCAL_GROUPS, st), "zero group counts");
    k_slot_keys<<<blocks_for(nr, 256), 256, 0, st>>>(c.index.idx.p, c.active.p, c.group.p, nr, n, gkeys.p, gvals.p);
    radix_sort_pairs(gkeys.p, gvals.p, nr, 0, X_WEDGE ? 16 : 8, c.scratch, st);
    k_slot_maps<<<blocks_for(n, 256), 256, 0, st>>>(gvals.p, c.group.p, c.active.p, n, c.slot_pidx.p, c.pidx_slot.p,
                                                   gcount.p, gcount.p + LOCAL_GROUPS);
    hip_check(hipGetLastError(), "launch slot maps");
    std::vector<int> counts(2*LOCAL_GROUPS);
    hip_check(hipMemcpyAsync(counts.data(), gcount.p, sizeof(int)*2*LOCAL_GROUPS, hipMemcpyDeviceToHost, st), "group counts");
    hip_check(hipStreamSynchronize(st), "slot order");
    c.group_beg.assign(LOCAL_GROUPS, 0);
    c.group_cnt.assign(counts.begin(), counts.begin() + LOCAL_GROUPS);
    c.group_all.assign(counts.begin() + LOCAL_GROUPS, counts.end());
    int acc = 0;
    for (int g = 0; g < LOCAL_GROUPS; ++g) { c.group_beg[g] = acc; acc += c.group_cnt[g]; }
    c.n_active = acc;

    // neighbour lists, measures and the group dependency graph
    const size_t na = size_t(std::max(1, c.n_active));
    c.nbr.alloc(na*N_K);
    c.measure.alloc(na);
    hip_check(hipMemsetAsync(c.graph.p, 0, sizeof(unsigned long long)*LOCAL_GROUPS, st), "zero graph");
    hip_check(hipMemsetAsync(c.overflow.p, 0, sizeof(unsigned int), st), "zero overflow");
    if (c.search == GDX_COL_SEARCH_KDTREE)
    {
        kd_build(c);
        kd_query(c);
    }
    else
    {
        int waves = blocks_for(nr, WAVE);
        if (X_WEDGE) c.nbr_image.alloc(na*N_K);
#define GDX_TOPK_SHAPED(DEDUP, TILE, WPB)                                                                    \
    k_col_topk<DEDUP, TILE, WPB><<<xcd_grid(blocks_for(waves, WPB)), WPB*WAVE, 0, st>>>(                     \
        c.index.rec.p, c.index.nodes.p, nr, DIM, MORTON_MAX_LEVEL, c.index.idx.p, c.pts.p, c.pidx_slot.p,     \
        c.group.p, c.user, c.nbr.p, c.measure.p, c.graph.p, c.overflow.p, c.nbr_image.p)
#define GDX_TOPK_LAUNCH(DEDUP, TILE)                                                                         \
    do                                                                                                        \
    {                                                                                                         \
        if (c.topk_waves == 1) GDX_TOPK_SHAPED(DEDUP, TILE, 1);                                               \
        else if (c.topk_waves == 2) GDX_TOPK_SHAPED(DEDUP, TILE, 2);                                          \
        else GDX_TOPK_SHAPED(DEDUP, TILE, 4);                                                                 \
    } while (0)
        if (!c.unique_ids && X_WEDGE)
            GDX_TOPK_LAUNCH(bool(X_WEDGE), 0);
        else if (c.topk_tile == 8)
            GDX_TOPK_LAUNCH(false, 8);
        else if (c.topk_tile == 16)
            GDX_TOPK_LAUNCH(false, 16);
        else if (c.topk_tile == 32)
            GDX_TOPK_LAUNCH(false, 32);
        else
            GDX_TOPK_LAUNCH(false, 0);
#undef GDX_TOPK_LAUNCH
#undef GDX_TOPK_SHAPED
        hip_check(hipGetLastError(), "launch top-K");
    }
    unsigned int ovf = 0;
    std::vector<unsigned long long> graph(LOCAL_GROUPS);
    hip_check(hipMemcpyAsync(&ovf, c.overflow.p, sizeof(ovf), hipMemcpyDeviceToHost, st), "read overflow");
    hip_check(hipMemcpyAsync(graph.data(), c.graph.p, sizeof(unsigned long long)*LOCAL_GROUPS, hipMemcpyDeviceToHost, st), "read graph");
    hip_check(hipStreamSynchronize(st), "top-K");
    if (ovf) throw status_error(11, "traversal stack overflow in col_cache_get");
    c.edges.assign(size_t(LOCAL_GROUPS)*LOCAL_WORDS, 0u);
    for (int g = 0; g < LOCAL_GROUPS; ++g)
        for (int d = 0; d < LOCAL_GROUPS; ++d)
            if (graph[g] >> d & 1ull) c.edges[size_t(g)*LOCAL_WORDS + d / 32] |= 1u << (d % 32);
    c.geometry_valid = true;
    c.rev_valid = false;
    load_state(c);
}

void load_state(context &c)
{
    if (!c.geometry_valid) throw status_error(16, "collision geometry has not been built");
    hipStream_t st = c.stream;
    const size_t na = size_t(std::max(1, c.n_active));
    c.st.alloc(na);
    c.res.alloc(na);
    c.res32.alloc(na);
    c.cached.alloc(na);
    c.rng.alloc(na);
    c.queue.alloc(na);
    c.qstart.alloc(na);
    c.cqueue.alloc(na);
    c.cstart.alloc(na);
    c.lam.alloc(na);
    c.lam_add.alloc(na);
    c.dirty.alloc(na);
    c.stale.alloc(na);
    c.qsum.alloc(na);
    if (c.diag) c.evw.alloc(na);
    c.qevents.alloc(na);
    if (COAG_KERNEL == 3) c.env.alloc(na);
    if (PARTNER_HG) c.hgs.alloc(na);
    if (c.n_active > 0)
    {
        k_gather_state<<<blocks_for(c.n_active, 256), 256, 0, st>>>(c.user, c.slot_pidx.p, c.n_active, c.rng_p.p, c.st.p,
                                                                    c.res.p, c.res32.p, c.rng.p, c.env.p, c.hgs.p);
        hip_check(hipGetLastError(), "launch gather state");
    }
    // a new state: no kept bound is valid, no publish is pending
    hip_check(hipMemsetAsync(c.dirty.p, 0, na, st), "reset dirty flags");
    hip_check(hipMemsetAsync(c.lam_add.p, 0, sizeof(double)*na, st), "reset bound increments");
    invalidate_bounds(c);
    c.state_loaded = true;
}

void invalidate_bounds(context &c)
{
    // on the context's stream, completed before any run (batches launch on the first handle's stream)
    if (c.stale.p)
    {
        hip_check(hipMemsetAsync(c.stale.p, 1, c.stale.n, c.stream), "mark bounds stale");
        hip_check(hipStreamSynchronize(c.stream), "mark bounds stale");
    }
    c.model_version = majorant_version(c.model);
}

// reverse neighbour lists of the geometry epoch: for every slot j the owners whose rows hold j
__global__ __launch_bounds__(256) void k_rev_count(const int *nbr, long long entries, unsigned *count)
{
    const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= entries) return;
    const int j = nbr[i];
    if (j >= 0) atomicAdd(&count[j], 1u);
}

__global__ __launch_bounds__(256) void k_rev_fill(const int *nbr, long long entries, unsigned *cursor, int *rev)
{
    const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= entries) return;
    const int j = nbr[i];
    if (j >= 0) rev[atomicAdd(&cursor[j], 1u)] = int(i / N_K);
}

void build_reverse(context &c)
{
    const int n = c.n_active;
    if ((long long)n > MAJORANT_MAX_OWNERS)
        throw status_error(GDX_COL_ERR_ARG, "a swarm with a majorant holds at most " + std::to_string(MAJORANT_MAX_OWNERS)
                                            + " active owners (32-bit reverse neighbour offsets)");
    const size_t na = size_t(std::max(1, n));
    hipStream_t st = c.stream;
    c.rev_beg.alloc(na);
    c.rev_end.alloc(na);
    c.rev.alloc(na*N_K);
    if (n > 0)
    {
        const long long entries = (long long)n*N_K;
        const unsigned blocks = unsigned((entries + 255) / 256);
        hip_check(hipMemsetAsync(c.rev_end.p, 0, sizeof(unsigned)*na, st), "zero reverse counts");
        k_rev_count<<<blocks, 256, 0, st>>>(c.nbr.p, entries, c.rev_end.p);
        hip_check(hipGetLastError(), "launch reverse count");
        scan_exclusive_u32(c.rev_end.p, c.rev_beg.p, n, c.scratch, st);
        hip_check(hipMemcpyAsync(c.rev_end.p, c.rev_beg.p, sizeof(unsigned)*na, hipMemcpyDeviceToDevice, st),
                  "reverse cursors");
        // after the fill each cursor is the end of its slot's list
        k_rev_fill<<<blocks, 256, 0, st>>>(c.nbr.p, entries, c.rev_end.p, c.rev.p);
        hip_check(hipGetLastError(), "launch reverse fill");
    }
    c.rev_valid = true;
    invalidate_bounds(c);
}

void store_state(context &c)
{
    if (!c.state_loaded) throw status_error(16, "collision state has not been loaded");
    if (c.n_active > 0)
    {
        k_scatter_state<<<blocks_for(c.n_active, 256), 256, 0, c.stream>>>(c.user, c.slot_pidx.p, c.n_active, c.st.p,
                                                                          c.rng.p, c.rng_p.p);
        hip_check(hipGetLastError(), "launch scatter state");
    }
}

}  // namespace gdx
