// Node-end search of the octree fill, shared by the device fill kernel and the CPU test
// that checks it (plain C++ when compiled without hipcc).
//
// The particles are sorted by octant key; dp[i] = delta_i + 1, where delta_i is the
// number of leading octree levels that sorted particles i and i+1 share, and
// dp[n-1] = 0.  The node of level L whose particle range starts at s ends at the first
// j >= s with dp[j] <= L: particle j+1 then shares fewer than L levels with j, while
// every earlier pair in [s, j] shares at least L.  dp[n-1] = 0 ends every node.
//
// Search levels: level 0 is dp itself; level k >= 1 holds the minimum of each aligned
// group of 16 entries of level k-1 (entry e covers positions [e 16^k, (e+1) 16^k)), and
// entries whose positions all lie at n or beyond count as 0.  Levels are added until
// one fits in a single group of 16 (at least levels 0..2).  The search checks the rest
// of s's group on level 0, then the rest of the enclosing group on each level above,
// and from the first group with an entry <= L goes down again through the first such
// entry of each level.  The groups of levels 0..2 depend only on s and are loaded
// together; nodes of up to about 4096 particles (nearly all) need at most two more
// loads on the way down.  The search cannot run past the top level, whose one group
// holds the entry of position n-1 (value 0).
//
// Every load is a 16-byte-aligned group inside its level's region of one buffer
// (search_level_offset): each region starts on a 256-byte boundary and extends at
// least 64 bytes past its level's entries.  Bytes past the last valid entry of a level
// are never used, because the entry of position n-1 is found first.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define GH_HD __host__ __device__ __forceinline__
#else
#define GH_HD inline
#endif

namespace gh {

GH_HD int64_t search_level_size(int64_t n, int k)
{
    int64_t sz = n;
    for (int j = 0; j < k; j++) sz = (sz + 15) >> 4;
    return sz;
}

GH_HD int64_t search_level_offset(int64_t n, int k)
{
    int64_t off = 0;
    for (int j = 0; j < k; j++) off += (search_level_size(n, j) + 64 + 255) & ~(int64_t)255;
    return off;
}

GH_HD int search_levels(int64_t n)
{
    int k = 2;
    while (search_level_size(n, k) > 16) k++;
    return k + 1;
}

// bytes of the buffer holding every level
GH_HD int64_t search_buffer_bytes(int64_t n) { return search_level_offset(n, search_levels(n)); }

struct alignas(16) Bytes16 {
    uint32_t w[4];
};

GH_HD Bytes16 load16(const unsigned char* __restrict__ a, int64_t g)
{
    return *(const Bytes16*)(a + (g << 4));
}

// index of the first byte b >= from of the group with value <= L, 16 if none
GH_HD int first_le16(const Bytes16& g, int from, int L)
{
    int r = 16;
#pragma unroll
    for (int b = 15; b >= 0; b--) {
        const int v = (int)((g.w[b >> 2] >> (8 * (b & 3))) & 0xffu);
        if (b >= from && v <= L) r = b;
    }
    return r;
}

struct EndSearch {
    const unsigned char* base;  // level 0 (dp); level k at base + search_level_offset(n, k)
    const unsigned char* m1;    // level 1
    const unsigned char* m2;    // level 2
    int levels;                 // search_levels(n)
};

GH_HD EndSearch end_search(const unsigned char* base, int64_t n)
{
    EndSearch es;
    es.base = base;
    es.m1 = base + search_level_offset(n, 1);
    es.m2 = base + search_level_offset(n, 2);
    es.levels = search_levels(n);
    return es;
}

// The first groups a search from s reads, on levels 0..2: they depend on s only, so a
// caller ending several nodes that start at s loads them once.
struct EndStart {
    int64_t s, g0, p1, g1, p2, g2;
    Bytes16 w0, w1, w2;
};

GH_HD EndStart end_start(const EndSearch& es, int64_t s)
{
    EndStart st;
    st.s = s;
    st.g0 = s >> 4;
    st.p1 = st.g0 + 1;
    st.g1 = st.p1 >> 4;
    st.p2 = st.g1 + 1;
    st.g2 = st.p2 >> 4;
    st.w0 = load16(es.base, st.g0);
    st.w1 = load16(es.m1, st.g1);
    st.w2 = load16(es.m2, st.g2);
    return st;
}

// Last sorted position of the node of level L that starts at st.s.
GH_HD int64_t node_end_from(const EndSearch& es, int64_t n, const EndStart& st, int L)
{
    int r = first_le16(st.w0, (int)(st.s & 15), L);
    if (r < 16) return (st.g0 << 4) + r;
    r = first_le16(st.w1, (int)(st.p1 & 15), L);
    if (r < 16) {
        const int64_t q1 = (st.g1 << 4) + r;
        return (q1 << 4) + first_le16(load16(es.base, q1), 0, L);
    }
    r = first_le16(st.w2, (int)(st.p2 & 15), L);
    if (r < 16) {
        const int64_t q2 = (st.g2 << 4) + r;
        const int64_t q1 = (q2 << 4) + first_le16(load16(es.m1, q2), 0, L);
        return (q1 << 4) + first_le16(load16(es.base, q1), 0, L);
    }
    // large nodes: up one level at a time, then down through every level
    int64_t p = st.g2 + 1;
    for (int k = 3; k < es.levels; k++) {
        const int64_t g = p >> 4;
        r = first_le16(load16(es.base + search_level_offset(n, k), g), (int)(p & 15), L);
        if (r < 16) {
            int64_t q = (g << 4) + r;
            for (int j = k - 1; j >= 0; j--)
                q = (q << 4) + first_le16(load16(es.base + search_level_offset(n, j), q), 0, L);
            return q;
        }
        p = g + 1;
    }
    return n - 1;  // not reached: the top level's one group holds position n-1's entry
}

// Last sorted position of the node of level L that starts at s.
GH_HD int64_t node_end_dp(const EndSearch& es, int64_t n, int64_t s, int L)
{
    return node_end_from(es, n, end_start(es, s), L);
}

}  // namespace gh
