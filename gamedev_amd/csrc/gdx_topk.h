// Register sorting networks of the tiled top-K selection (geometry.hip), host- and device-callable so
// that a host program can check them (topk_check.cpp).
//
// A candidate is the 64-bit key (bits(d^2) << 32) | code: for d^2 >= 0 and codes >= 0 the unsigned key
// order is the (d^2, code) order of the reference's neighbour comparison (morton_types.cuh:68-72), and
// the empty entry (inf, INT_MAX) is the largest key.
#ifndef GDX_TOPK_H
#define GDX_TOPK_H

#include <hip/hip_runtime.h>

#include <cstdint>

#define GDX_TOPK_FN __host__ __device__ __forceinline__

namespace gdx {

constexpr uint64_t TOPK_EMPTY = (uint64_t(0x7f800000u) << 32) | uint64_t(0x7fffffffu);

GDX_TOPK_FN void cx_up(uint64_t &a, uint64_t &b)
{
    const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// ascending bitonic sort of N keys (N a power of two)
template <int N>
GDX_TOPK_FN void bitonic_sort(uint64_t (&a)[N])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
    {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
        {
#pragma unroll
            for (int i = 0; i < N; ++i)
            {
                const int l = i ^ j;
                if (l > i)
                {
                    if ((i & k) == 0) cx_up(a[i], a[l]);
                    else cx_up(a[l], a[i]);
                }
            }
        }
    }
}

// ascending sort of a bitonic sequence of N keys
template <int N>
GDX_TOPK_FN void bitonic_merge(uint64_t (&a)[N])
{
#pragma unroll
    for (int j = N >> 1; j > 0; j >>= 1)
    {
#pragma unroll
        for (int i = 0; i < N; ++i)
        {
            const int l = i ^ j;
            if (l > i) cx_up(a[i], a[l]);
        }
    }
}

// L (K keys, ascending) becomes the K smallest keys of L and S (B keys, ascending): min(L[i], S[K-1-i])
// is bitonic and holds the K smallest of both (Batcher's merge), and one bitonic merge sorts it
template <int K, int B>
GDX_TOPK_FN void topk_merge_sorted(uint64_t (&L)[K], const uint64_t (&S)[B])
{
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        const int j = K - 1 - i;
        if (j < B) L[i] = L[i] < S[j] ? L[i] : S[j];
    }
    bitonic_merge<K>(L);
}

}  // namespace gdx

#endif  // GDX_TOPK_H
