// Host check of the tiled top-K selection's merge logic (gdx_topk.h): one query lane's candidate stream
// run through the lane's filter, LDS buffer and B-key merges, against the K smallest eligible keys by a
// full sort. Merges also happen early at random, as in the kernel, where any lane of the wavefront that
// fills its buffer merges every lane's buffer.
//
//   topk_check <trials> <seed>     exit 0 when every trial of every (K, B) agrees
//
// Built for the host only (hipcc -x hip --cuda-host-only) from the header the kernel uses.
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gdx_topk.h"

using namespace gdx;

namespace {

struct lcg
{
    uint64_t s;
    uint64_t next()
    {
        s = s*6364136223846793005ull + 1442695040888963407ull;
        return s >> 11;
    }
    double uniform() { return (double(next()) + 0.5)*0x1p-53; }
};

uint32_t fbits(float v)
{
    uint32_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

template <int K, int B>
int run(long trials, lcg &g, long &merges)
{
    int bad = 0;
    for (long t = 0; t < trials; ++t)
    {
        const int n = int(g.next() % 600);
        // distances from a small set of values (ties) or continuous, codes distinct, some inactive
        const bool ties = g.next() % 2 == 0;
        const float q2 = float(g.uniform()*1.2);
        std::vector<std::pair<float, int>> recs(n);
        std::vector<int> codes(n);
        for (int i = 0; i < n; ++i) codes[i] = int(g.next() % 1000000);
        std::sort(codes.begin(), codes.end());
        codes.erase(std::unique(codes.begin(), codes.end()), codes.end());
        recs.resize(codes.size());
        for (size_t i = 0; i < recs.size(); ++i)
        {
            float d2 = ties ? float(int(g.next() % 8))*0.125f : float(g.uniform());
            int code = g.next() % 10 == 0 ? -1 : codes[i];
            recs[i] = {d2, code};
        }
        for (size_t i = recs.size(); i > 1; --i) std::swap(recs[i - 1], recs[g.next() % i]);

        uint64_t L[K];
        for (int k = 0; k < K; ++k) L[k] = TOPK_EMPTY;
        uint64_t buf[B];
        int cnt = 0;
        auto flush = [&] {
            uint64_t S[B];
            for (int r = 0; r < B; ++r) S[r] = r < cnt ? buf[r] : TOPK_EMPTY;
            bitonic_sort<B>(S);
            topk_merge_sorted<K, B>(L, S);
            cnt = 0;
            ++merges;
        };
        std::vector<uint64_t> all;
        for (auto &r : recs)
        {
            const uint64_t key = (uint64_t(fbits(r.first)) << 32) | uint32_t(r.second);
            const bool eligible = r.second >= 0 && r.first <= q2;
            if (eligible) all.push_back(key);
            if (eligible && key < L[K - 1]) buf[cnt++] = key;
            if (cnt == B || g.next() % 20 == 0) flush();
        }
        if (cnt > 0) flush();
        std::sort(all.begin(), all.end());
        for (int k = 0; k < K; ++k)
        {
            const uint64_t want = k < int(all.size()) ? all[k] : TOPK_EMPTY;
            if (L[k] != want) { ++bad; break; }
        }
    }
    return bad;
}

}  // namespace

int main(int argc, char **argv)
{
    const long trials = argc > 1 ? std::atol(argv[1]) : 10000;
    lcg g{argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1ull};
    long merges = 0;
    int bad[6] = {run<16, 8>(trials, g, merges), run<16, 16>(trials, g, merges), run<16, 32>(trials, g, merges),
                  run<32, 8>(trials, g, merges), run<32, 16>(trials, g, merges), run<32, 32>(trials, g, merges)};
    std::printf("mismatches %d %d %d %d %d %d\n", bad[0], bad[1], bad[2], bad[3], bad[4], bad[5]);
    std::printf("merges %ld\n", merges);
    for (int b : bad)
        if (b) return 1;
    return 0;
}
