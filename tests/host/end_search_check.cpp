// CPU check of the octree fill's node-end search (gravhopper_amd/csrc/gh_search.h),
// built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Random sorted, distinct octant keys of 1 or 2 words (21 levels of 3 bits per word,
// level 0 in bits 62..60), with few distinct leading digits so that long shared
// prefixes occur.  dp[i] = shared leading levels of keys i and i+1, plus 1 (dp[n-1] =
// 0), and every search level is written from its definition: entry e of level k is the
// minimum of dp over positions [e 16^k, (e+1) 16^k), positions from n on counting as 0.
// The buffer is allocated at exactly search_buffer_bytes(n) with the unused bytes set
// to noise, so a read outside it fails under the sanitizer and a read of an unwritten
// byte that matters gives a wrong answer.  Each answer is compared with the end found
// by comparing octant digits one by one.
#include "../../gravhopper_amd/csrc/gh_search.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace gh;
typedef unsigned long long u64;

typedef std::array<u64, 2> Key;

static int digit(const Key& key, int level)
{
    return (int)((key[level / 21] >> (60 - 3 * (level % 21))) & 7);
}

int main()
{
    std::mt19937_64 rng(20261202);
    long checks = 0, bad = 0, large = 0;
    int max_levels = 0;
    for (int trial = 0; trial < 240; trial++) {
        const int W = 1 + trial % 2;
        int64_t n = 1 + (int64_t)(rng() % (trial < 120 ? 3000 : trial < 230 ? 70000 : 1100000));
        if (trial < 20) n = 1 + trial;
        std::vector<Key> k((size_t)n, Key{0, 0});
        const int narrow = 2 + trial % 9;  // leading levels drawn from 2 octants only
        for (auto& key : k)
            for (int w = 0; w < W; w++) {
                u64 x = 0;
                for (int l = 0; l < 21; l++) x = (x << 3) | (rng() % (w * 21 + l < narrow ? 2 : 8));
                key[w] = x;
            }
        std::sort(k.begin(), k.end());
        k.erase(std::unique(k.begin(), k.end()), k.end());
        n = (int64_t)k.size();

        const int64_t bytes = search_buffer_bytes(n);
        const int levels = search_levels(n);
        max_levels = std::max(max_levels, levels);
        unsigned char* buf = (unsigned char*)malloc((size_t)bytes);
        for (int64_t b = 0; b < bytes; b++) buf[b] = (unsigned char)rng();
        std::vector<int> dp((size_t)n, 0);
        for (int64_t i = 0; i + 1 < n; i++) {
            int d = 0;
            while (d < 21 * W && digit(k[i], d) == digit(k[i + 1], d)) d++;
            dp[i] = d + 1;
            buf[i] = (unsigned char)dp[i];
        }
        buf[n - 1] = 0;
        for (int lv = 1; lv < levels; lv++) {
            unsigned char* out = buf + search_level_offset(n, lv);
            int64_t span = 1;
            for (int j = 0; j < lv; j++) span *= 16;
            for (int64_t e = 0; e < search_level_size(n, lv); e++) {
                int v = 255;
                for (int64_t j = e * span; j < (e + 1) * span; j++) v = std::min(v, j < n ? dp[j] : 0);
                out[e] = (unsigned char)v;
            }
        }
        const EndSearch es = end_search(buf, n);
        for (int q = 0; q < 400; q++) {
            const int64_t s = (q % 16 == 0) ? 0 : (int64_t)(rng() % n);
            const int L = (q % 16 == 0) ? q % 5 : (int)(rng() % (21 * W + 1));
            int64_t want = s;  // last position whose key shares the first L digits of s's
            while (want + 1 < n) {
                bool same = true;
                for (int l = 0; l < L && same; l++) same = digit(k[want + 1], l) == digit(k[s], l);
                if (!same) break;
                want++;
            }
            const int64_t got = node_end_dp(es, n, s, L);
            checks++;
            if (want - s >= 4096) large++;
            if (got != want) {
                if (bad < 5)
                    printf("mismatch: n=%lld W=%d s=%lld L=%d got=%lld want=%lld\n", (long long)n, W,
                           (long long)s, L, (long long)got, (long long)want);
                bad++;
            }
        }
        free(buf);
    }
    printf("%ld searches (%ld over nodes of 4096+ particles), up to %d levels, %ld wrong\n",
           checks, large, max_levels, bad);
    return bad != 0 || large == 0 || max_levels < 6;
}
