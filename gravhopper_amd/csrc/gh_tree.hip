// Barnes-Hut octree on the device, with the reference's exact topology.
//
// The reference (gravhopper/_jbgrav.c:360-462, 737-786) inserts particles one by one
// into a pointer octree: one particle per leaf, octant bit k = (x_k > c_k) strictly,
// branch = bit0 + 2 bit1 + 4 bit2, child centre c + (s * 0.5) * halfsize (FP64),
// child size = parent halfsize, root box from the particle extent with the size rule
// of :764-769.  The resulting tree does not depend on insertion order: it is the
// minimal octree in which every particle sits alone in a leaf, each particle's octant
// path being fixed by its own position and the FP64 centre recurrence.
//
// Device construction of that same tree:
//   1. k_bbox_*       extent -> root centre and size (the reference's arithmetic)
//   2. k_keys         per particle, octant digits of W*21 levels of the recurrence,
//                     21 levels (63 bits) per 64-bit word, level 0 most significant
//   3. radix sort     stable LSD sort of (key words, index), 8-bit digits
//                     (k_rs_hist / k_rs_scan / k_rs_scatter, wave-ballot ranking);
//                     one-word keys are sorted by their bits 24..63 (levels 0-12) and
//                     the short runs left equal there by the whole key (k_sort_runs)
//   4. k_delta        delta_i = number of leading levels shared by sorted keys i, i+1;
//                     keys equal over all W words -> more words (or coincident error);
//                     the same pass writes the node counts, the end-search levels and
//                     the particles in sorted order
//   5. nodes          in depth-first pre-order (children in branch order) the nodes
//                     whose particle range starts at sorted position i are the internal
//                     nodes of levels (delta_{i-1}, delta_i] followed by the leaf of i at
//                     level max(delta_{i-1}, delta_i) + 1, so one scan of
//                     max(0, delta_i - delta_{i-1}) + 1 numbers the whole tree; a node's
//                     range ends at the last j whose key shares the node's level count
//                     of leading digits (search over delta and its group minima,
//                     gh_search.h) and its skip pointer is the first node of position
//                     end+1
//   6. moments        node mass and first moment as differences of double-double prefix
//                     sums over the sorted particles (error far below one FP64 ulp)
//   7. walks          FP64: stackless pre-order walk with skip pointers.  A wavefront
//                     walks up to 64 targets in lockstep: the wave visits nodes in
//                     increasing order (next = n+1 if any waiting lane opened n, else
//                     skip[n]); a lane waits for every node from the end of the last
//                     subtree it accepted on, so each lane accepts exactly the nodes the
//                     reference's recursive walk (:487-541) accepts, while the node index
//                     stays wave-uniform and node records arrive through scalar loads.
//                     FP32: a group walk with one interaction list per 16 targets whose
//                     lists refine the reference's accepted nodes (k_walk_group).
#include <algorithm>
#include <cmath>

#include "gh_box.h"
#include "gh_internal.h"
#include "gh_search.h"

namespace gh {

typedef unsigned long long u64;
typedef unsigned int u32;

constexpr int kLevelsPerWord = 21;
constexpr int kMaxWords = 8;  // 168 levels

// FP64 node record of the walk and the dump
// FP64 record; a leaf carries its size with the sign bit set, so that the FP64 walk
// reads leaf-ness from the record itself
struct NodeD {
    double cx, cy, cz, size;
    double mx, my, mz, mass;  // centre of mass, mass
};
typedef double f64x8 __attribute__((ext_vector_type(8)));  // a NodeD as one load
static_assert(sizeof(NodeD) == sizeof(f64x8), "NodeD is 64 bytes");

// FP32 group walk's node record, 48 bytes read as three 16-byte loads: the FP64 centre
// (a leaf's is its particle's position, as a leaf is accepted without a test), the FP32
// size with a leaf's sign bit set, the FP32 mass, and the FP32 offset of the centre of
// mass from the centre (zero for a leaf, so that its entry is its exact position).
// Rounding the offset moves a monopole by at most 2^-24 of its node's size.
struct alignas(16) NodeW {
    double cx, cy, cz;
    float size, mass;
    float ox, oy, oz, pad;
};
static_assert(sizeof(NodeW) == 48, "NodeW is 48 bytes");

struct DD4 {
    double h[4], l[4];
};

struct TreeBufs {
    DevBuf keys[2];  // W words per element, word-major: keys[w * n + i]
    DevBuf idx[2];
    DevBuf hist, bbox_part, box, flags, delta, off, prefix, block_sums, block_counts, packed;
    DevBuf spos;  // (x, y, z, m) of the particles in sorted order
    DevBuf nodes64, nodesw, kids, skip, level, branch, leaf;
    DevBuf wprefix;  // FP32 walk prefixes per block of groups (k_walk_prefix)
    DevBuf tkeys[2], tidx[2];  // Morton order of position targets (tree_force_position)
    long long* readback = nullptr;  // pinned host words: flags[0..1], node count, flags[2]
    long long* readback_dev = nullptr;  // the same words' device address
    int64_t n_nodes = 0;
    int words = 0;
    int key_sel = 0;  // which of keys[0]/keys[1] holds the sorted keys
    int full_sorts = 0;  // builds left that sort one-word keys over all bits
};

// ------------------------------------------------------------------------------
// 1. bounding box
// ------------------------------------------------------------------------------
// Extent of each workgroup's share of the points: part[6 b ..] = (min, max).  The
// workgroups also clear zero[0 .. nzero) (the small sort's digit counts; may be null).
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const double* __restrict__ pos,
                                                         int64_t n, double* __restrict__ part,
                                                         u32* __restrict__ zero, int64_t nzero)
{
    __shared__ double smin[3][kBlock], smax[3][kBlock];
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nzero;
         j += (int64_t)gridDim.x * kBlock)
        zero[j] = 0u;
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) mn[k] = mx[k] = pos[k];  // particle 0 seeds, as :747-750
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        for (int k = 0; k < 3; k++) {
            const double q = pos[3 * i + k];
            if (q < mn[k]) mn[k] = q;
            if (q > mx[k]) mx[k] = q;
        }
    }
    for (int k = 0; k < 3; k++) {
        smin[k][threadIdx.x] = mn[k];
        smax[k][threadIdx.x] = mx[k];
    }
    block_extent(smin, smax);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            part[6 * blockIdx.x + k] = smin[k][0];
            part[6 * blockIdx.x + 3 + k] = smax[k][0];
        }
}

// box[0..2] = centre, box[3] = size (gravhopper/_jbgrav.c:764-772)
__global__ __launch_bounds__(kBlock) void k_bbox_final(const double* __restrict__ part,
                                                       int nparts, double eps,
                                                       double* __restrict__ box)
{
    __shared__ double smin[3][kBlock], smax[3][kBlock];
    double b[4];
    box_from_parts(part, nparts, eps, smin, smax, b);
    if (threadIdx.x < 4) box[threadIdx.x] = b[threadIdx.x];
}

// ------------------------------------------------------------------------------
// 2. keys from the FP64 centre recurrence
// ------------------------------------------------------------------------------
// One level: octant digit of x against centre c, then the child centre
// c + (s * 0.5) * h with s = +-1 and h the current node's halfsize (:406, :426, :456).
__device__ __forceinline__ int descend(double x, double y, double z, double& cx, double& cy,
                                       double& cz, double h)
{
    const bool gx = x > cx, gy = y > cy, gz = z > cz;
    cx = cx + (gx ? 0.5 : -0.5) * h;
    cy = cy + (gy ? 0.5 : -0.5) * h;
    cz = cz + (gz ? 0.5 : -0.5) * h;
    return (int)gx | ((int)gy << 1) | ((int)gz << 2);
}

// One workgroup per radix-sort chunk (kRsChunk particles, kKeysItems per thread).
// packed (nullable) receives (x, y, z, m) of each particle in input order, so that the
// gather into sorted order after the sort fetches one 32-byte record per particle
// rather than two scattered pieces (position and mass).  hist (nullable) receives the
// first radix pass's digit counts, those of bits hshift..hshift+7 of key word hword, per
// chunk: hist[d * gridDim.x + chunk], as k_rs_hist would count them.
constexpr int kKeysBlock = 1024;
constexpr int kKeysItems = kRsChunk / kKeysBlock;
static_assert(kKeysBlock >= 256 && kKeysItems * kKeysBlock == kRsChunk, "k_keys shape");
__global__ __launch_bounds__(kKeysBlock) void k_keys(
    const double* __restrict__ pos, int64_t n, const double* __restrict__ box, int words,
    u64* __restrict__ keys, u32* __restrict__ idx, int* __restrict__ flags,
    const double* __restrict__ mass, double4* __restrict__ packed, u32* __restrict__ hist,
    int hword, int hshift)
{
    __shared__ u32 h[256];
    if (hist) {
        if (threadIdx.x < 256) h[threadIdx.x] = 0u;
        __syncthreads();
    }
    if (flags && blockIdx.x == 0 && threadIdx.x < 4) flags[threadIdx.x] = 0;
    const double bx = box[0], by = box[1], bz = box[2], bs = box[3];
    for (int it = 0; it < kKeysItems; it++) {
        const int64_t i = (int64_t)blockIdx.x * kRsChunk + (int64_t)it * kKeysBlock + threadIdx.x;
        if (i >= n) break;
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if (packed) packed[i] = make_double4(x, y, z, mass[i]);
        double cx = bx, cy = by, cz = bz;
        double size = bs;
        for (int w = 0; w < words; w++) {
            u64 k = 0;
            for (int l = 0; l < kLevelsPerWord; l++) {
                const double hh = 0.5 * size;
                const int d = descend(x, y, z, cx, cy, cz, hh);
                k = (k << 3) | (u64)d;
                size = hh;
            }
            keys[(int64_t)w * n + i] = k;
            if (hist && w == hword) atomicAdd(&h[(u32)(k >> hshift) & 255u], 1u);
        }
        idx[i] = (u32)i;
    }
    if (hist) {
        __syncthreads();
        if (threadIdx.x < 256) hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
    }
}

// ------------------------------------------------------------------------------
// 3. LSD radix sort, 8-bit digits, stable
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_rs_hist(const u64* __restrict__ key, int64_t n,
                                                    int shift, u32* __restrict__ hist,
                                                    int nblocks)
{
    __shared__ u32 h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    u32 d[kRsItems];
#pragma unroll
    for (int it = 0; it < kRsItems; it++) {
        const int64_t i = base + (int64_t)it * kBlock + threadIdx.x;
        d[it] = i < n ? (u32)((key[i] >> shift) & 255u) : 256u;
    }
#pragma unroll
    for (int it = 0; it < kRsItems; it++)
        if (d[it] < 256u) atomicAdd(&h[d[it]], 1u);
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Per digit d (one workgroup each): exclusive prefix of hist[d][0..nblocks) in place,
// and the digit's total count in totals[d].  kRsScanItems consecutive entries per
// thread, a shuffle scan of the thread sums in each wavefront and one barrier for the
// wavefronts' totals per round of kBlock * kRsScanItems entries.
constexpr int kRsScanItems = 4;
__global__ __launch_bounds__(kBlock) void k_rs_scan(u32* __restrict__ hist, int nblocks,
                                                    u32* __restrict__ totals)
{
    __shared__ u32 wsum[kBlock / 64];
    u32* row = hist + (int64_t)blockIdx.x * nblocks;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += kBlock * kRsScanItems) {
        const int b = b0 + threadIdx.x * kRsScanItems;
        u32 v[kRsScanItems];
        u32 t = 0;
#pragma unroll
        for (int k = 0; k < kRsScanItems; k++) {
            v[k] = b + k < nblocks ? row[b + k] : 0u;
            t += v[k];
        }
        u32 incl = t;  // inclusive scan of t over the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const u32 o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        u32 before = carry, round = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wave) before += wsum[w];
            round += wsum[w];
        }
        before += incl - t;
#pragma unroll
        for (int k = 0; k < kRsScanItems; k++) {
            if (b + k < nblocks) row[b + k] = before;
            before += v[k];
        }
        carry += round;
        __syncthreads();  // wsum is rewritten by the next round
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one 8-bit digit.  Every element's output position is its digit's
// base (all elements of smaller digits) + the elements of the same digit in earlier
// blocks (hist, scanned) + its rank inside the block.  Each wavefront of the block owns
// a contiguous range of 64 * kRsItems elements and ranks them by itself (ballots per
// 64 elements, running per-digit counts in a row of LDS that only this wavefront
// touches), so the block synchronises only to add the counts of the earlier
// wavefronts.  STAGE (one key word only): the block's elements are first placed in LDS
// in output order and then written out by consecutive threads, so each digit's run of
// the block leaves as contiguous stores; otherwise every element is stored directly.
// DIRECT (STAGE only, at most kRsDirectChunks blocks): hist holds the unscanned digit
// counts of every block, and each block forms its bases from them itself (totals is
// unused); when hnext is not null the block also counts, per output chunk, the digits
// of the next pass (bits shift + 8 ..) into hnext, which must be zero beforehand.
template <bool STAGE, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_rs_scatter(
    const u64* __restrict__ kin, u64* __restrict__ kout, const u32* __restrict__ iin,
    u32* __restrict__ iout, int64_t n, int words, int wsel, int shift,
    const u32* __restrict__ hist, int nblocks, const u32* __restrict__ totals,
    u32* __restrict__ hnext)
{
    static_assert(!DIRECT || STAGE, "the direct form stages its output");
    constexpr int kWaves = kBlock / 64;
    __shared__ u32 gbase0[256];          // this block's first output position per digit
    __shared__ u32 lstart[256];          // this block's first LDS slot per digit (STAGE)
    __shared__ u32 wcount[kWaves][256];  // per-wave running counts, then wave offsets
    __shared__ u64 stage_k[STAGE ? kRsChunk : 1];
    __shared__ u32 stage_i[STAGE ? kRsChunk : 1];
    __shared__ u32 next_cnt[DIRECT ? kRsDirectChunks * 256 : 1];  // [output chunk][digit]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int dd = threadIdx.x;  // one digit per thread in the per-digit steps
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    const int64_t wbase = base + (int64_t)wave * 64 * kRsItems;  // this wave's range
    // every item's digit word and index is loaded up front, so that the loads are in
    // flight together while the digit bases are formed
    u64 kw[kRsItems];
    u32 id[kRsItems];
#pragma unroll
    for (int it = 0; it < kRsItems; it++) {
        const int64_t i = wbase + it * 64 + lane;
        kw[it] = i < n ? kin[(int64_t)wsel * n + i] : 0ull;
        id[it] = i < n ? iin[i] : 0u;
    }
    if (DIRECT) {
        // this digit's count in every block: total, earlier blocks, this block
        u32 tot = 0u, before = 0u, mine = 0u;
        for (int b = 0; b < nblocks; b++) {
            const u32 h = hist[dd * nblocks + b];
            tot += h;
            if (b < (int)blockIdx.x) before += h;
            if (b == (int)blockIdx.x) mine = h;
        }
        gbase0[dd] = tot;
        lstart[dd] = mine;
        for (int j = threadIdx.x; j < kRsDirectChunks * 256; j += kBlock) next_cnt[j] = 0u;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const u32 t = dd >= off ? gbase0[dd - off] : 0u;
            const u32 m = dd >= off ? lstart[dd - off] : 0u;
            __syncthreads();
            gbase0[dd] += t;
            lstart[dd] += m;
            __syncthreads();
        }
        gbase0[dd] += before - tot;  // exclusive base plus earlier blocks
        lstart[dd] -= mine;
    } else {
    // digit bases: exclusive scan of the 256 digit totals, plus earlier blocks
    {
        const u32 tot = totals[dd];
        gbase0[dd] = tot;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const u32 t = dd >= off ? gbase0[dd - off] : 0u;
            __syncthreads();
            gbase0[dd] += t;
            __syncthreads();
        }
        const u32 first = gbase0[dd] - tot + hist[(int64_t)dd * nblocks + blockIdx.x];
        __syncthreads();
        gbase0[dd] = first;
    }
    if (STAGE) {
        // this block's count of each digit, then their exclusive scan
        const u32 h0 = hist[(int64_t)dd * nblocks + blockIdx.x];
        const u32 h1 = (int)blockIdx.x + 1 < nblocks
                           ? hist[(int64_t)dd * nblocks + blockIdx.x + 1]
                           : totals[dd];
        const u32 cnt = h1 - h0;
        lstart[dd] = cnt;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const u32 t = dd >= off ? lstart[dd - off] : 0u;
            __syncthreads();
            lstart[dd] += t;
            __syncthreads();
        }
        lstart[dd] -= cnt;
    }
    }
    for (int w = 0; w < kWaves; w++) wcount[w][dd] = 0u;
    __syncthreads();
    // rank inside the wave: earlier items of this wave, then earlier lanes of this item
    const u64 lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    u32 dig[kRsItems], rank[kRsItems];
#pragma unroll
    for (int it = 0; it < kRsItems; it++) {
        const int64_t i = wbase + it * 64 + lane;
        const bool valid = i < n;
        const u32 d = valid ? (u32)((kw[it] >> shift) & 255u) : 0u;
        u64 peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const u64 m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        // every lane reads its digit's count before the leader of the digit adds this
        // item's peers (LDS operations of one wavefront take effect in program order)
        const u32 before = wcount[wave][d];
        if (valid && (peers & lt) == 0ull) wcount[wave][d] = before + (u32)__popcll(peers);
        dig[it] = d;
        rank[it] = before + (u32)__popcll(peers & lt);
    }
    __syncthreads();
    {
        u32 run = 0u;
        for (int w = 0; w < kWaves; w++) {
            const u32 t = wcount[w][dd];
            wcount[w][dd] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kRsItems; it++) {
        const int64_t i = wbase + it * 64 + lane;
        if (i >= n) continue;
        const u32 d = dig[it];
        const u32 inblk = wcount[wave][d] + rank[it];  // rank among the block's digit d
        if (STAGE) {
            const u32 slot = lstart[d] + inblk;
            stage_k[slot] = kw[it];
            stage_i[slot] = id[it];
        } else {
            const u32 dst = gbase0[d] + inblk;
            for (int w = 0; w < words; w++)
                kout[(int64_t)w * n + dst] = (w == wsel) ? kw[it] : kin[(int64_t)w * n + i];
            iout[dst] = id[it];
        }
    }
    if (STAGE) {
        __syncthreads();
        // LDS slot e holds the block's e-th element in output order
        const int64_t m = (n - base) < kRsChunk ? (n - base) : kRsChunk;
        for (int64_t e = threadIdx.x; e < m; e += kBlock) {
            const u64 k = stage_k[e];
            const u32 d = (u32)((k >> shift) & 255u);
            const u32 dst = gbase0[d] + (u32)(e - lstart[d]);
            kout[dst] = k;
            iout[dst] = stage_i[e];
            if (DIRECT && hnext)
                atomicAdd(&next_cnt[(dst / kRsChunk) * 256 + (u32)((k >> (shift + 8)) & 255u)],
                          1u);
        }
        if (DIRECT && hnext) {
            __syncthreads();
            for (int j = threadIdx.x; j < nblocks * 256; j += kBlock) {
                const u32 v = next_cnt[j];
                if (v) atomicAdd(&hnext[(j & 255) * nblocks + (j >> 8)], v);
            }
        }
    }
}

// One-word tree keys are radix-sorted by their bits kSortLowBits..63 only (the octant
// digits of levels 0-12), which leaves in input order the runs of positions whose keys
// agree there.  Such runs are few and short in practice (a galaxy of 2^20 particles has
// about 10^4, none longer than 4), so the thread at the start of each run sorts it by
// the whole key in place, stably (insertion sort), which gives the order of a sort over
// all bits.  A run longer than kRunMax is left alone and counted in *long_runs; the
// build then sorts again over all bits.  Permuting a run never changes bits 24..63 of
// any of its entries, so no thread's view of where runs start and end depends on when
// it reads its neighbours.
constexpr int kSortLowBits = 24;
constexpr int kRunMax = 64;
__global__ __launch_bounds__(kBlock) void k_sort_runs(u64* __restrict__ keys,
                                                      u32* __restrict__ idx, int64_t n,
                                                      int* __restrict__ long_runs)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i + 1 >= n) return;
    const u64 t = keys[i] >> kSortLowBits;
    if ((keys[i + 1] >> kSortLowBits) != t) return;
    if (i > 0 && (keys[i - 1] >> kSortLowBits) == t) return;
    int64_t e = i + 2;  // one past the run's end, or i + kRunMax + 1 for a long run
    while (e < n && e - i <= kRunMax && (keys[e] >> kSortLowBits) == t) e++;
    if (e - i > kRunMax) {
        atomicAdd(long_runs, 1);
        return;
    }
    for (int64_t j = i + 1; j < e; j++) {
        const u64 k = keys[j];
        const u32 p = idx[j];
        int64_t q = j;
        for (; q > i && keys[q - 1] > k; q--) {
            keys[q] = keys[q - 1];
            idx[q] = idx[q - 1];
        }
        keys[q] = k;
        idx[q] = p;
    }
}

// ------------------------------------------------------------------------------
// 4. shared levels of adjacent sorted keys
// ------------------------------------------------------------------------------
// delta_i + 1 for sorted positions i and i+1 (delta_i = their shared leading levels,
// words * 21 when the keys are equal over all words), 0 for i = n-1.
__device__ __forceinline__ int shared_levels_plus1(const u64* __restrict__ keys, int64_t n,
                                                   int words, int64_t i)
{
    if (i >= n - 1) return 0;
    int delta = words * kLevelsPerWord;
    for (int w = 0; w < words; w++) {
        const u64 a = keys[(int64_t)w * n + i], b = keys[(int64_t)w * n + i + 1];
        if (a != b) {
            const int lz = __clzll(a ^ b);  // >= 1: bit 63 is never used
            delta = w * kLevelsPerWord + (lz - 1) / 3;
            break;
        }
    }
    return delta + 1;
}

struct DeltaArgs {
    const u64* keys;    // sorted keys, word-major
    const u32* perm;         // sorted position -> particle
    const double4* packed;   // particles (x, y, z, m), input order
    unsigned char* search;  // end-search buffer: dp = level 0, then levels 1, 2, ...
    long long* cnt;         // nodes starting at each sorted position
    double4* spos;          // sorted (x, y, z, m)
    int* flags;             // [0] pairs equal over all words, [1] coincident pairs
                            // ([2] is k_sort_runs' count of long runs)
};

// Per sorted position i, after the sort:
//  - dp[i] = delta_i + 1 (dp[n-1] = 0); flags[0] counts pairs equal over all words,
//    flags[1] those at identical positions (both rare, so their atomics cost nothing);
//  - cnt[i] = internal nodes starting at i (levels delta_{i-1} + 1 .. delta_i) plus its
//    leaf, the input of the node numbering scan;
//  - spos[i] = (x, y, z, m) of the particle, read contiguously by the moment scan and the
//    node fill;
//  - the end-search levels 1 (minimum of dp over aligned groups of 16, a 16-lane
//    reduction inside a wavefront) and 2 (groups of 256, one per workgroup), with
//    positions from n on counting as 0.
static_assert(kBlock == 256, "k_delta writes one level-2 entry per workgroup of 256");
__global__ __launch_bounds__(kBlock) void k_delta(DeltaArgs a, int64_t n, int words)
{
    __shared__ int wave_min[kBlock / 64];
    __shared__ int lane_m[kBlock];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int m = i < n ? shared_levels_plus1(a.keys, n, words, i) : 0;
    unsigned char* dp = a.search;
    if (i < n) {
        dp[i] = (unsigned char)m;
        const u32 p = a.perm[i];
        const double4 r = a.packed[p];
        a.spos[i] = r;
        if (m == words * kLevelsPerWord + 1) {
            atomicAdd(&a.flags[0], 1);
            const double4 o = a.packed[a.perm[i + 1]];
            if (r.x == o.x && r.y == o.y && r.z == o.z) atomicAdd(&a.flags[1], 1);
        }
    }
    lane_m[threadIdx.x] = m;
    // level 1 and 2 minima
    int g = m;
    for (int o = 1; o < 16; o <<= 1) g = min(g, __shfl_xor(g, o));
    if ((threadIdx.x & 15) == 0) (dp + search_level_offset(n, 1))[i >> 4] = (unsigned char)g;
    for (int o = 16; o < 64; o <<= 1) g = min(g, __shfl_xor(g, o));
    if ((threadIdx.x & 63) == 0) wave_min[threadIdx.x >> 6] = g;
    __syncthreads();
    // node counts: levels (delta_{i-1}, delta_i] plus the leaf
    if (i < n) {
        const int mprev = threadIdx.x ? lane_m[threadIdx.x - 1]
                                      : (i ? shared_levels_plus1(a.keys, n, words, i - 1) : 0);
        const int lprev = mprev - 1, lcur = m - 1;
        a.cnt[i] = (long long)(lcur > lprev ? lcur - lprev : 0) + 1;
    }
    if (threadIdx.x == 0) {
        int b = wave_min[0];
        for (int k = 1; k < kBlock / 64; k++) b = min(b, wave_min[k]);
        (dp + search_level_offset(n, 2))[blockIdx.x] = (unsigned char)b;
    }
}

// End-search levels 3 and up, each from the level below, by one workgroup (the levels
// above 2 hold at most n / 4096 entries).
__device__ __forceinline__ void search_levels_block(unsigned char* __restrict__ dp, int64_t n,
                                                    int levels)
{
    for (int k = 3; k < levels; k++) {
        const unsigned char* lo = dp + search_level_offset(n, k - 1);
        unsigned char* hi = dp + search_level_offset(n, k);
        const int64_t lo_size = search_level_size(n, k - 1), size = search_level_size(n, k);
        for (int64_t e = threadIdx.x; e < size; e += kBlock) {
            int v = 255;
            for (int j = 0; j < 16; j++) {
                const int64_t q = 16 * e + j;
                v = min(v, q < lo_size ? (int)lo[q] : 0);
            }
            hi[e] = (unsigned char)v;
        }
        __syncthreads();  // the next level reads this one
    }
}

// ------------------------------------------------------------------------------
// generic two-phase scans (int64 sums and double-double 4-vectors)
// ------------------------------------------------------------------------------
__device__ __forceinline__ void dd_add(double& h, double& l, double bh, double bl)
{
    const double s = h + bh;
    const double bp = s - h;
    double e = (h - (s - bp)) + (bh - bp);
    e += l + bl;
    h = s + e;
    l = e - (h - s);
}

// A scan operation: the element type T, its identity and sum, and the input type In
// with the function that turns an input element into a T.
struct OpI64 {
    typedef long long T;
    typedef long long In;
    __device__ static T zero() { return 0; }
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static T load(const In* in, int64_t i) { return in[i]; }
};
// Moments of sorted particles (x, y, z, m): (m, m x, m y, m z) as double-double.
struct OpDD4 {
    typedef DD4 T;
    typedef double4 In;
    __device__ static T zero()
    {
        DD4 z;
        for (int k = 0; k < 4; k++) z.h[k] = z.l[k] = 0.0;
        return z;
    }
    __device__ static T add(T a, const T& b)
    {
        for (int k = 0; k < 4; k++) dd_add(a.h[k], a.l[k], b.h[k], b.l[k]);
        return a;
    }
    __device__ static T load(const In* in, int64_t i) { return of(in[i]); }
    __device__ static T of(const double4 q)
    {
        DD4 r;
        r.h[0] = q.w;
        r.h[1] = q.w * q.x;
        r.h[2] = q.w * q.y;
        r.h[3] = q.w * q.z;
        for (int k = 0; k < 4; k++) r.l[k] = 0.0;
        return r;
    }
};

constexpr int kScanItems = 4;
constexpr int kScanChunk = kBlock * kScanItems;

// Local phase of an exclusive scan of in[0..n), for this workgroup's chunk: out[i] =
// sum of the chunk's elements before i, and sums[b] = the chunk's total.  sh is LDS for
// kBlock elements.
template <class Op>
__device__ __forceinline__ void scan_local_block(const typename Op::In* __restrict__ in,
                                                 int64_t n, typename Op::T* __restrict__ out,
                                                 typename Op::T* __restrict__ sums,
                                                 typename Op::T* sh)
{
    typedef typename Op::T T;
    const int64_t b = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T run = Op::zero();
    for (int k = 0; k < kScanItems; k++) {
        v[k] = run;
        if (b + k < n) run = Op::add(run, Op::load(in, b + k));
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        T t = Op::zero();
        const bool take = threadIdx.x >= off;
        if (take) t = sh[threadIdx.x - off];
        __syncthreads();
        if (take) sh[threadIdx.x] = Op::add(t, sh[threadIdx.x]);
        __syncthreads();
    }
    const T before = threadIdx.x ? sh[threadIdx.x - 1] : Op::zero();
    for (int k = 0; k < kScanItems; k++)
        if (b + k < n) out[b + k] = Op::add(before, v[k]);
    if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = sh[kBlock - 1];
}

// Exclusive scan in place of the nb chunk totals by one workgroup (each thread a
// contiguous run); the total goes to *total and is returned to every thread.
template <class Op>
__device__ __forceinline__ typename Op::T scan_sums_block(typename Op::T* __restrict__ sums,
                                                          int64_t nb,
                                                          typename Op::T* __restrict__ total,
                                                          typename Op::T* sh)
{
    typedef typename Op::T T;
    const int64_t per = (nb + kBlock - 1) / kBlock;
    const int64_t b = (int64_t)threadIdx.x * per;
    const int64_t e = b + per < nb ? b + per : nb;
    T run = Op::zero();
    for (int64_t k = b; k < e; k++) run = Op::add(run, sums[k]);
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        T t = Op::zero();
        const bool take = threadIdx.x >= off;
        if (take) t = sh[threadIdx.x - off];
        __syncthreads();
        if (take) sh[threadIdx.x] = Op::add(t, sh[threadIdx.x]);
        __syncthreads();
    }
    T acc = threadIdx.x ? sh[threadIdx.x - 1] : Op::zero();
    for (int64_t k = b; k < e; k++) {
        const T t = sums[k];
        sums[k] = acc;
        acc = Op::add(acc, t);
    }
    const T tot = sh[kBlock - 1];
    if (threadIdx.x == kBlock - 1) *total = tot;
    return tot;
}

// The fill's two scans over the sorted particles, local phase: node counts cnt -> first
// node ids off, and moments (m, m x, m y, m z) of spos -> double-double prefixes pref.
// Both scans share their loads and one ladder over the threads' sums.
struct CountMoments {
    long long c;
    DD4 d;
};
__global__ __launch_bounds__(kBlock) void k_scan_local_fill(
    const long long* __restrict__ cnt, const double4* __restrict__ spos, int64_t n,
    long long* __restrict__ off, long long* __restrict__ off_sums, DD4* __restrict__ pref,
    DD4* __restrict__ pref_sums)
{
    __shared__ CountMoments sh[kBlock];
    const int64_t b = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    long long ci[kScanItems];
    double4 qi[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const bool in = b + k < n;
        ci[k] = in ? cnt[b + k] : 0;
        qi[k] = in ? spos[b + k] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    CountMoments run;
    run.c = 0;
    run.d = OpDD4::zero();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        run.c += ci[k];
        run.d = OpDD4::add(run.d, OpDD4::of(qi[k]));
    }
    sh[threadIdx.x] = run;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        CountMoments t;
        t.c = 0;
        t.d = OpDD4::zero();
        const bool take = threadIdx.x >= o;
        if (take) t = sh[threadIdx.x - o];
        __syncthreads();
        if (take) {
            sh[threadIdx.x].c += t.c;
            sh[threadIdx.x].d = OpDD4::add(t.d, sh[threadIdx.x].d);
        }
        __syncthreads();
    }
    CountMoments before;
    if (threadIdx.x) {
        before = sh[threadIdx.x - 1];
    } else {
        before.c = 0;
        before.d = OpDD4::zero();
    }
    // element k: the earlier threads' sum plus this thread's sum before k
    long long cv = 0;
    DD4 dv = OpDD4::zero();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (b + k < n) {
            off[b + k] = before.c + cv;
            pref[b + k] = OpDD4::add(before.d, dv);
        }
        cv += ci[k];
        dv = OpDD4::add(dv, OpDD4::of(qi[k]));
    }
    if (threadIdx.x == kBlock - 1) {
        off_sums[blockIdx.x] = sh[kBlock - 1].c;
        pref_sums[blockIdx.x] = sh[kBlock - 1].d;
    }
}

// Two workgroups after k_scan_local_fill, side by side: workgroup 1 scans the moments'
// chunk totals (total to pref[n]); workgroup 0 scans the node counts' (total to off[n]),
// forms the end-search levels 3 and up, and writes the words the host reads back to its
// pinned memory: the two flag counts of k_delta, the node count and the long-run count
// of k_sort_runs.
__global__ __launch_bounds__(kBlock) void k_scan_sums_fill(
    long long* __restrict__ off_sums, DD4* __restrict__ pref_sums, int64_t nb,
    long long* __restrict__ off_total, DD4* __restrict__ pref_total,
    unsigned char* __restrict__ dp, int64_t n, int levels, const int* __restrict__ flags,
    long long* __restrict__ host)
{
    __shared__ long long shi[kBlock];
    __shared__ DD4 shd[kBlock];
    if (blockIdx.x == 1) {
        scan_sums_block<OpDD4>(pref_sums, nb, pref_total, shd);
        return;
    }
    const long long nodes = scan_sums_block<OpI64>(off_sums, nb, off_total, shi);
    search_levels_block(dp, n, levels);
    if (threadIdx.x == 0) {
        int* hf = (int*)host;
        hf[0] = flags[0];
        hf[1] = flags[1];
        host[2] = flags[2];
        host[1] = nodes;
    }
}

// A scan kept in two parts, block-local exclusive prefixes and the exclusive prefix of
// the block totals, and added where it is read (the one reader visits each element
// about twice, so a separate pass adding them would cost more than it saves).
template <class Op>
struct ScanView {
    const typename Op::T* local;   // n+1 entries, local[n] = total
    const typename Op::T* blocks;  // exclusive prefix of the block totals
    int64_t n;
    // element j (0 <= j <= n) of the exclusive scan
    __device__ typename Op::T at(int64_t j) const
    {
        return j == n ? local[n] : Op::add(blocks[j / kScanChunk], local[j]);
    }
};

// ------------------------------------------------------------------------------
// 5./6. node ends, moments, node records
// ------------------------------------------------------------------------------
__device__ __forceinline__ double dd_diff(const DD4& a, const DD4& b, int k)
{
    return (a.h[k] - b.h[k]) + (a.l[k] - b.l[k]);
}

struct FillArgs {
    const u32* perm;
    const unsigned char* dp;
    EndSearch es;
    ScanView<OpI64> off;  // first node id of each sorted position, n+1 entries
    const double4* spos;  // sorted (x, y, z, m)
    const double* box;
    ScanView<OpDD4> pref;  // moment prefix sums, n+1 entries
    NodeD* nodes64;
    NodeW* nodesw;
    int* skip;
    short* level;
    signed char* branch;
    int* leaf;  // original particle index, -1 for internal nodes
    int what;   // kFill* bits: which records this build needs
};

constexpr int kFillF64 = 1;   // NodeD + leaf flags (FP64 walk, dump)
constexpr int kFillKids = 2;  // children lists (FP32 group walk)
constexpr int kFillDump = 4;  // levels and branches (canonical dump)
constexpr int kFillW = 8;     // NodeW records (FP32 group walk)

// Node id of level lv and branch br with centre c and size, ending at sorted position
// e, centre of mass mc and mass m; p = the particle of a leaf.
__device__ __forceinline__ void emit_node(const FillArgs& a, int64_t id, int lv, int br,
                                          double cx, double cy, double cz, double size,
                                          int64_t e, bool is_leaf, u32 p, double mcx,
                                          double mcy, double mcz, double m)
{
    NodeD d;
    d.cx = cx;
    d.cy = cy;
    d.cz = cz;
    d.size = is_leaf ? -size : size;
    d.mx = mcx;
    d.my = mcy;
    d.mz = mcz;
    d.mass = m;
    a.skip[id] = (int)a.off.at(e + 1);
    if (a.what & (kFillF64 | kFillDump)) {
        a.nodes64[id] = d;
        a.leaf[id] = is_leaf ? (int)p : -1;
    }
    if (a.what & kFillDump) {
        a.level[id] = (short)lv;
        a.branch[id] = (signed char)br;
    }
    if (a.what & kFillW) {
        NodeW w;
        if (is_leaf) {
            w.cx = mcx;
            w.cy = mcy;
            w.cz = mcz;
            w.ox = w.oy = w.oz = 0.f;
            w.size = -(float)size;
        } else {
            w.cx = cx;
            w.cy = cy;
            w.cz = cz;
            w.ox = (float)(mcx - cx);
            w.oy = (float)(mcy - cy);
            w.oz = (float)(mcz - cz);
            w.size = (float)size;
        }
        w.mass = (float)m;
        w.pad = 0.f;
        a.nodesw[id] = w;
    }
}

// Exclusive prefix sum of v over the wavefront's lanes, and its total (wave-uniform).
__device__ __forceinline__ int wave_scan_excl(int v, int& total)
{
    const int lane = threadIdx.x & 63;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    total = __shfl(incl, 63);
    return incl - v;
}

// An internal node waiting in the fill's queue: its centre and size, sorted start
// position, node id, level and branch.
struct FillEntry {
    double cx, cy, cz, size;
    u32 pos, id;
    int lv, br;
};
constexpr int kFillQueue = 128;  // entries per wavefront

// One thread per sorted particle i: the internal nodes starting at i (levels lprev+1 ..
// lcur) and the leaf of i (level lcur+1, or lprev+1 when no internal node starts
// here).  Each lane runs the FP64 centre recurrence down its own path and queues its
// internal nodes in LDS; the wavefront then finishes the queued nodes together (end
// search, moment prefixes, records), a lane per node, so that a lane with many internal
// nodes does not hold up the others (a particle starts 0.55 internal nodes on average
// at N=2^20, some start tens); a queue that would overflow is taken in rounds.  Last,
// every lane writes its leaf.  One wavefront per workgroup.
constexpr int kFillBlock = 64;
__global__ __launch_bounds__(kFillBlock) void k_fill_nodes(FillArgs a, int64_t n)
{
    __shared__ FillEntry queue[kFillQueue];
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kFillBlock + lane;
    const bool live = i < n;  // every lane takes part in the wave-wide steps
    u32 p = 0;
    double x = 0.0, y = 0.0, z = 0.0, m_leaf = 0.0;
    int n_int = 0, first = 0;
    int64_t base = 0;
    double cx = a.box[0], cy = a.box[1], cz = a.box[2], size = a.box[3];
    int br = -1;
    if (live) {
        p = a.perm[i];
        const double4 q = a.spos[i];
        x = q.x;
        y = q.y;
        z = q.z;
        m_leaf = q.w;
        const int lprev = (i == 0) ? -1 : (int)a.dp[i - 1] - 1;
        const int lcur = (int)a.dp[i] - 1;
        n_int = lcur > lprev ? lcur - lprev : 0;
        first = lprev + 1;  // level of the first node starting at i
        base = a.off.at(i);
        for (int lv = 0; lv < first; lv++) {
            const double h = 0.5 * size;
            br = descend(x, y, z, cx, cy, cz, h);
            size = h;
        }
    }
    int k = 0;  // this lane's internal nodes queued so far
    for (;;) {
        int total;
        const int pre = wave_scan_excl(n_int - k, total);
        if (total == 0) break;
        const int take = min(n_int - k, max(0, kFillQueue - pre));
        for (int t = 0; t < take; t++, k++) {
            FillEntry e;
            e.cx = cx;
            e.cy = cy;
            e.cz = cz;
            e.size = size;
            e.pos = (u32)i;
            e.id = (u32)(base + k);
            e.lv = first + k;
            e.br = br;
            queue[pre + t] = e;
            const double h = 0.5 * size;
            br = descend(x, y, z, cx, cy, cz, h);
            size = h;
        }
        __syncthreads();
        const int cnt = total < kFillQueue ? total : kFillQueue;
        for (int j = lane; j < cnt; j += kFillBlock) {
            const FillEntry e = queue[j];
            const int64_t s0 = e.pos;
            const int64_t end = node_end_from(a.es, n, end_start(a.es, s0), e.lv);
            const DD4 lo = a.pref.at(s0), hi = a.pref.at(end + 1);
            const double m = dd_diff(hi, lo, 0);
            emit_node(a, e.id, e.lv, e.br, e.cx, e.cy, e.cz, e.size, end, false, 0u,
                      dd_diff(hi, lo, 1) / m, dd_diff(hi, lo, 2) / m, dd_diff(hi, lo, 3) / m, m);
        }
        __syncthreads();  // the queue is read before the next round refills it
    }
    if (live)
        emit_node(a, base + n_int, first + n_int, br, cx, cy, cz, size, i, true, p, x, y, z,
                  m_leaf);
}

// ------------------------------------------------------------------------------
// 7. walks
// ------------------------------------------------------------------------------
// Scheduling (both walks).  Workgroups are dealt to the 8 XCDs round-robin (workgroup
// b runs on XCD b mod 8); xcd_block() renumbers them so that XCD x takes the x-th of 8
// contiguous, Morton-ordered ranges of walk units, and the L2 of XCD x holds the nodes
// of one region of the tree.  The FP64 walk runs one wavefront per workgroup, so that the
// hardware refills a SIMD slot as soon as its walk ends rather than when the slowest of
// several does; the FP32 group walk runs two independent ones per workgroup (below).
constexpr int kXcds = 8;
constexpr int64_t kWalkUnits = 4096;  // walks in flight for small target counts
constexpr int64_t kSortTargetsMin = 4096;  // position targets sorted from this many on
constexpr int kWalkBlock = 64;
__device__ __forceinline__ int64_t xcd_block(int64_t per_xcd)
{
    return (int64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
}

// FP32 walk: group opening test, shared interaction lists.
//
// A wavefront walks G consecutive targets in walk (Morton) order, G a power of two up to
// 16.  The group's frame is the FP64 position o of one of its targets; every coordinate
// below is relative to o, so FP32 rounding scales with the group's extent and the
// distances involved, not with where the system sits.  With the box B of the group's
// targets, an internal node is accepted for the whole group when
//     d_min(B, centre)^2 > (size / theta)^2,
// d_min the distance from the node's geometric centre to B, or else when every target x
// of the group has |centre - x|^2 > (size / theta)^2.  Either way every target passes the
// reference's test size/|centre - x| < theta (gravhopper/_jbgrav.c:496-502) there too:
// each target's interactions refine the ones the reference accepts for it (a node the
// reference accepts is either accepted here or opened and replaced by its children), and
// leaves are accepted as in the reference.
//
// Traversal: an LDS stack of node ids; each step the 64 lanes take up to 64 ids from
// its top and test them in parallel (the FP64 centre of the node's NodeW record,
// re-based to the frame, and its FP32 size).  An accepted node's centre of mass (centre
// plus offset, re-based in FP64) goes to the group's interaction list; an opened node's
// children (kids[8 n ..]) are pushed.  Each
// popped id pushes at most 8, so a step pops at most (kStackCap - top) / 7 ids; with
// less room left than one id needs, the wave walks the top id's subtree one node at a
// time with the skip pointers instead, pushing nothing.
//
// Evaluation: whenever the list holds 64 entries they are evaluated by the whole wave,
// lane l taking target l mod G and the entries of slice l / G (64 / G slices), in FP32
// (fused multiply-adds), each flush's FP32 sums added into FP64 accumulators; the
// slices are added in FP64 at the end.  F64 evaluates the same lists (the traversal and
// its tests are identical) in FP64, to check the FP32 arithmetic.  counts (nullable)
// receives each target's list length.
constexpr int kStackCap = 768;
constexpr int kListCap = 128;
// Walk prefixes (k_walk_prefix): groups per block, accepted nodes and steps at most, the
// stack ids a group's walk may start with, and the ints of one block's state (stack
// height, list length, stack, list).
constexpr int kPrefixGroups = 64;
constexpr int kPrefixCap = 1024;
constexpr int kPrefixSteps = 256;
constexpr int kPrefixStart = kStackCap - 7 * 64;
constexpr int kPrefixInts = 2 + kStackCap + kPrefixCap;
// Two wavefronts per workgroup, each walking its own group: a CU holds at most 16
// workgroups, so one-wavefront workgroups would cap it at 16 walks.
constexpr int kGroupWaves = 2;
constexpr int kGroupBlock = 64 * kGroupWaves;

template <bool F64>
struct GroupEntry {
    typedef float4 T;
};
template <>
struct GroupEntry<true> {
    typedef double4 T;
};

// Lanes' exclusive prefix sum and the wave total of a count 0..15, from four ballots
// (no cross-lane data movement); the total is wave-uniform.
__device__ __forceinline__ int wave_excl_count(int v, int& total)
{
    int pre = 0;
    total = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const u64 m = __ballot((v >> b) & 1);
        pre += (int)__builtin_amdgcn_mbcnt_hi((u32)(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo((u32)m, 0u)) << b;
        total += __popcll(m) << b;
    }
    return pre;
}

// Orders this wavefront's LDS accesses around it.  A wavefront's LDS instructions
// execute in program order, so within one wavefront this fence (no s_barrier) is all
// that a write by one lane followed by a read by another needs; the group walk keeps
// each wavefront's LDS to itself.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lanes_below(u64 m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// A NodeW in registers.
struct WalkRec {
    double cx, cy, cz;
    float size, mass, ox, oy, oz;
};

__device__ __forceinline__ WalkRec load_rec(const NodeW* __restrict__ nodes, int n)
{
    const char* p = (const char*)(nodes + n);
    const double2 a = *(const double2*)p;        // cx, cy
    const double2 b = *(const double2*)(p + 16); // cz, (size, mass)
    const float4 c = *(const float4*)(p + 32);   // offset of the centre of mass
    WalkRec r;
    r.cx = a.x;
    r.cy = a.y;
    r.cz = b.x;
    r.size = __int_as_float(__double2loint(b.y));
    r.mass = __int_as_float(__double2hiint(b.y));
    r.ox = c.x;
    r.oy = c.y;
    r.oz = c.z;
    return r;
}

// One walk with the opening box B (centre bc, half-sizes bh) in the frame o: traversal
// and evaluation as described above.  This lane's target is (dx0, dy0, dz0) = x - o
// (FP64; t its FP32 rounding), its slice of each flush `slice` of `nslices`, and the
// group has cnt targets.  list holds kListCap entries; fails is LDS scratch of 64
// entries.  prefix (nullable): the state k_walk_prefix left for the group's block, to
// start from instead of the root.  Adds the lane's partial field to (ax, ay, az);
// listed = the list length.
template <bool F64, bool GUARD>
__device__ __forceinline__ void walk_lists(
    const NodeW* __restrict__ nodes, const int* __restrict__ kids, const int* __restrict__ skip,
    int n_nodes, double ox, double oy, double oz, float bcx, float bcy, float bcz, float bhx,
    float bhy, float bhz, double dx0, double dy0, double dz0, int slice, int nslices,
    float inv_theta, double eps2, int* stack, typename GroupEntry<F64>::T* list,
    u64* fails, int cnt, double& ax, double& ay, double& az, int& listed,
    const int* __restrict__ prefix)
{
    typedef typename GroupEntry<F64>::T Entry;
    const int lane = threadIdx.x & 63;
    // the per-target tests' scratch: the list's upper half, free whenever they run (the
    // list then holds fewer than 64 entries)
    float4* const pend = (float4*)(list + 64);
    const int G = 64 / nslices;  // targets of the group; lane l holds target l mod G
    const float tx = (float)dx0, ty = (float)dy0, tz = (float)dz0;
    int held = 0;  // entries in the list (list[0 .. held))
    const float eps2f = in_vgpr((float)eps2);

    // evaluates `take` entries (<= 64) from the list head
    auto flush = [&](int take) {
        if (F64) {
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int j = slice; j < take; j += nslices) {
                const Entry e = list[j];
                const double ex = (double)e.x - dx0, ey = (double)e.y - dy0,
                             ez = (double)e.z - dz0;
                const double d = ex * ex + ey * ey + ez * ez + eps2;
                double inv = 0.0;
                if (d != 0.0) {
                    const double r = __builtin_amdgcn_rsq(d);
                    const double rn = r * __builtin_fma(-0.5 * d, r * r, 1.5);
                    inv = rn * rn * rn;
                }
                const double mi = (double)e.w * inv;
                fx += ex * mi; fy += ey * mi; fz += ez * mi;
            }
            ax += fx; ay += fy; az += fz;
        } else {
            float fx = 0.f, fy = 0.f, fz = 0.f;
            for (int j = slice; j < take; j += nslices) {
                const Entry e = list[j];
                const float ex = (float)e.x - tx, ey = (float)e.y - ty, ez = (float)e.z - tz;
                float r2 = __builtin_fmaf(ex, ex, eps2f);
                r2 = __builtin_fmaf(ey, ey, r2);
                r2 = __builtin_fmaf(ez, ez, r2);
                const float ri = __builtin_amdgcn_rsqf(r2);
                float sc = ((float)e.w * ri) * (ri * ri);
                if (GUARD) sc = clamp_finite(sc);
                fx = __builtin_fmaf(ex, sc, fx);
                fy = __builtin_fmaf(ey, sc, fy);
                fz = __builtin_fmaf(ez, sc, fz);
            }
            ax += (double)fx; ay += (double)fy; az += (double)fz;
        }
        // the entries after the evaluated ones move to the front (fewer than 64, so the
        // two ranges do not overlap)
        const int rest = held - take;
        Entry keep;
        if (lane < rest) keep = list[take + lane];
        wave_sync();
        if (lane < rest) list[lane] = keep;
        held = rest;
        wave_sync();
    };
    // the list entry of a node: centre of mass (centre plus offset) and mass, re-based to
    // the frame in FP64
    auto entry_of = [&](const WalkRec& r) -> Entry {
        Entry e;
        e.x = (r.cx - ox) + (double)r.ox;
        e.y = (r.cy - oy) + (double)r.oy;
        e.z = (r.cz - oz) + (double)r.oz;
        e.w = r.mass;
        return e;
    };
    // The group opening test (identical in both instances) of the node with record rec
    // (centre, size with the leaf's sign bit): accept = leaf, or the box B clears the
    // node's opening radius R; need = B does not clear it but its far corner does, so
    // that every target may still pass the reference's test on its own (then the
    // per-target tests below decide).  c: the centre in the frame; R2 = R^2.
    auto box_test = [&](const WalkRec& rec, float& cx, float& cy, float& cz, float& R2,
                        bool& need) -> bool {
        cx = (float)(rec.cx - ox);
        cy = (float)(rec.cy - oy);
        cz = (float)(rec.cz - oz);
        const float R = rec.size * inv_theta;  // +inf for theta = 0
        R2 = R * R;
        // branch-free, so that the whole record is loaded at once whatever the node is
        const bool leaf = __float_as_int(rec.size) < 0;
        const float ax_ = fabsf(cx - bcx), ay_ = fabsf(cy - bcy), az_ = fabsf(cz - bcz);
        const float ddx = fmaxf(0.f, ax_ - bhx), ddy = fmaxf(0.f, ay_ - bhy),
                    ddz = fmaxf(0.f, az_ - bhz);
        const bool clear = __builtin_fmaf(ddx, ddx, __builtin_fmaf(ddy, ddy, ddz * ddz)) > R2;
        const float fx = ax_ + bhx, fy = ay_ + bhy, fz = az_ + bhz;
        need = !leaf && !clear && cnt > 1 &&
               __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz)) > R2;
        return leaf || clear;
    };
    // Per-target tests of the nodes whose lanes have `need` set, spread over the wave as
    // (node, target) pairs: 64 / G nodes per pass, lane l testing node l / G of the pass
    // against target l mod G (its own).  A node is accepted when none of its G lanes
    // fails.  Returns the acceptance of this lane's node (false without `need`).
    auto target_tests = [&](bool need, float cx, float cy, float cz, float R2) -> bool {
        const u64 needm = __ballot(need);
        if (needm == 0) return false;
        const int rank = lanes_below(needm);
        if (need) pend[rank] = make_float4(cx, cy, cz, R2);
        wave_sync();
        const int m = (int)__popcll(needm);
        const int per = 64 / G;  // nodes per pass
        int pass = 0;
        for (int base = 0; base < m; base += per, pass++) {
            const int j = base + lane / G;
            bool fail = false;
            if (j < m) {
                const float4 q = pend[j];
                const float ex = q.x - tx, ey = q.y - ty, ez = q.z - tz;
                fail = !(__builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez)) > q.w);
            }
            const u64 fm = __ballot(fail);
            if (lane == 0) fails[pass] = fm;
        }
        wave_sync();
        bool ok = false;
        if (need) {
            const int ps = rank / per, sl = rank - ps * per;
            const u64 field =
                G == 64 ? fails[ps] : (fails[ps] >> (sl * G)) & ((1ull << G) - 1ull);
            ok = field == 0;
        }
        return ok;
    };

    wave_sync();  // the previous walk of this wavefront is done with stack and list
    int top = 0;
    if (prefix) {
        // the block's deferred nodes onto the stack, and its accepted nodes' entries into
        // the list, evaluated by 64 as they come in
        top = prefix[0];
        const int npre = prefix[1];
        const int* pstack = prefix + 2;
        const int* plist = pstack + top;
        for (int j = lane; j < top; j += 64) stack[j] = pstack[j];
        for (int b = 0; b < npre; b += 64) {
            const int take = npre - b < 64 ? npre - b : 64;
            const WalkRec rec = load_rec(nodes, lane < take ? plist[b + lane] : 0);
            if (lane < take) list[held + lane] = entry_of(rec);
            held += take;
            listed += take;
            wave_sync();
            while (held >= 64) flush(64);
        }
    } else {
        if (lane == 0) stack[0] = 0;
        top = n_nodes > 0 ? 1 : 0;
    }
    wave_sync();
    while (top > 0) {
        int k = top < 64 ? top : 64;
        const int room = (kStackCap - top) / 7;
        if (room < 1) {
            // no room for a parallel step: walk the top id's subtree sequentially
            const int n0 = stack[top - 1];
            top -= 1;
            int p = n0;
            const int end = skip[n0];
            while (p < end) {
                const WalkRec rec = load_rec(nodes, p);
                float cx, cy, cz, R2;
                bool need;
                bool acc = box_test(rec, cx, cy, cz, R2, need);
                if (!acc && need) {
                    // every lane's target against the node (lanes past the group's
                    // count repeat its last target)
                    const float ex = cx - tx, ey = cy - ty, ez = cz - tz;
                    acc = __ballot(!(__builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez)) >
                                     R2)) == 0;
                }
                if (acc) {
                    if (lane == 0) list[held] = entry_of(rec);
                    held += 1;
                    listed += 1;
                    p = skip[p];
                } else {
                    p = p + 1;
                }
                wave_sync();
                if (held >= 64) {
                    flush(64);
                    wave_sync();
                }
            }
            continue;
        }
        if (k > room) k = room;
        const bool active = lane < k;
        // the popped node's record (inactive lanes load the root's, and ignore it); its
        // children only where it is opened, so that the lanes of accepted nodes and
        // leaves (most of them) send no request for them
        const int n = active ? stack[top - 1 - lane] : 0;
        const WalkRec rec = load_rec(nodes, n);
        float cx, cy, cz, R2;
        bool need;
        bool acc_n = box_test(rec, cx, cy, cz, R2, need) && active;
        need = need && active;
        if (target_tests(need, cx, cy, cz, R2)) acc_n = true;
        int4 ka = make_int4(-1, -1, -1, -1), kb = ka;
        if (active && !acc_n) {
            ka = *(const int4*)(kids + 8 * (int64_t)n);
            kb = *(const int4*)(kids + 8 * (int64_t)n + 4);
        }
        // children come first in kids[8 n ..], then -1s: 8 plus the sign bits
        int nk = 8 + (ka.x >> 31) + (ka.y >> 31) + (ka.z >> 31) + (ka.w >> 31) + (kb.x >> 31) +
                 (kb.y >> 31) + (kb.z >> 31) + (kb.w >> 31);
        if (!active || acc_n) nk = 0;
        // appends to the list, in lane order
        const u64 am = __ballot(acc_n);
        const int before = lanes_below(am);
        const int added = (int)__popcll(am);
        // pushes: the popped ids' slots are reused, children in lane order
        int total;
        const int pre = wave_excl_count(nk, total);
        wave_sync();  // every lane has read its popped id
        if (acc_n) list[held + before] = entry_of(rec);
        if (nk > 0) {
            const int w = top - k + pre;
            const int kk[8] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y, kb.z, kb.w};
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < nk) stack[w + q] = kk[q];
        }
        held += added;
        listed += added;
        top = top - k + total;
        wave_sync();
        while (held >= 64) flush(64);
        wave_sync();
    }
    if (held > 0) flush(held);
}

constexpr int kLooseLog2 = 6;

__device__ __forceinline__ int exponent_f(float x)
{
    return (int)((__float_as_int(x) >> 23) & 0xff) - 127;  // 0 gives -127
}

// Walk prefixes.  Much of a group's walk is the same for its neighbours: far from all
// of them, nodes are accepted; near all of them, nodes are opened.  One wavefront per
// block of kPrefixGroups consecutive groups walks the tree once for the block's target
// box, deciding every node that is decided alike for all of the block's targets and
// deferring the others, and stores prefix[0] = h, prefix[1] = L, then the h deferred
// nodes and the L accepted ones.  Each group of the block starts its walk with the
// accepted nodes in its list and the deferred ones on its stack, so it accepts exactly
// the nodes its walk from the root accepts (whether a node is visited and accepted does
// not depend on the order of the walk), only in another order.  A node is decided alike
// when it is a leaf (always accepted), or when, with R = size / theta and the block's
// target box, its centre is farther than R + mu from all of the box (every target passes
// the reference's test there, so every group accepts it) or nearer than R - mu to all
// of it (no target passes, so every group opens it).  mu = 1e-4 (R + the box's largest
// half-size) covers the FP32 rounding of the groups' tests (about 1e-7 of the distances)
// and the widening of their boxes (1e-6 of their half-sizes).  The walk stops where its
// stack is nearly full, after kPrefixSteps steps, or before its list would overflow or
// its deferred nodes and stack together would exceed kPrefixStart ids (so that a group's
// walk starting with them still pops 64 per step); what is left on its stack is deferred
// too.  A block with a target that is not finite starts its walks at the root.
__global__ __launch_bounds__(64) void k_walk_prefix(
    const NodeW* __restrict__ nodes, const int* __restrict__ kids, int n_nodes,
    const double* __restrict__ tgt, const double4* __restrict__ stgt,
    const u32* __restrict__ order, int64_t nt, int G, double inv_theta,
    int* __restrict__ prefixes)
{
    __shared__ int stack[kStackCap];
    __shared__ int list[kPrefixCap];
    __shared__ int defer[kStackCap];
    const int lane = threadIdx.x;
    int* const out = prefixes + (int64_t)blockIdx.x * kPrefixInts;
    const int64_t b0 = (int64_t)blockIdx.x * kPrefixGroups * G;
    const int64_t b1 = b0 + (int64_t)kPrefixGroups * G < nt ? b0 + (int64_t)kPrefixGroups * G : nt;
    // the block's target box, in FP64
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    constexpr int kBatch = 8;  // loads in flight per lane
    bool bad = false;          // a coordinate that is not finite
    for (int64_t k0 = b0 + lane; k0 < b1; k0 += 64 * kBatch) {
        double q[kBatch][3];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int64_t k = k0 + 64 * u < b1 ? k0 + 64 * u : k0;  // k0 < b1 repeats
            if (stgt) {
                const double4 p = stgt[k];
                q[u][0] = p.x, q[u][1] = p.y, q[u][2] = p.z;
            } else {
                const int64_t t = order ? (int64_t)order[k] : k;
                q[u][0] = tgt[3 * t], q[u][1] = tgt[3 * t + 1], q[u][2] = tgt[3 * t + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            for (int d = 0; d < 3; d++) {
                bad = bad || !isfinite(q[u][d]);
                lo[d] = fmin(lo[d], q[u][d]);
                hi[d] = fmax(hi[d], q[u][d]);
            }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
        }
    double bc[3], bh[3], bhmax = 0.0;
    for (int d = 0; d < 3; d++) {
        bc[d] = 0.5 * (lo[d] + hi[d]);
        bh[d] = 0.5 * (hi[d] - lo[d]);
        bhmax = fmax(bhmax, bh[d]);
    }
    // a block with a target that is not finite decides nothing: its walks start at the root
    const bool finite = __ballot(bad) == 0 && isfinite(bhmax) && isfinite(bc[0]) &&
                        isfinite(bc[1]) && isfinite(bc[2]);
    int top = n_nodes > 0 ? 1 : 0, held = 0, nd = 0;
    if (lane == 0) stack[0] = 0;
    wave_sync();
    for (int step = 0; finite && step < kPrefixSteps && top > 0; step++) {
        int k = top < 64 ? top : 64;
        const int room = (kStackCap - top) / 7;
        if (room < 1) break;
        if (k > room) k = room;
        const bool active = lane < k;
        const int n = active ? stack[top - 1 - lane] : 0;
        const WalkRec rec = load_rec(nodes, n);
        bool acc = false, open = false;
        if (active) {
            if (__float_as_int(rec.size) < 0) {
                acc = true;
            } else {
                const double R = (double)rec.size * inv_theta;
                const double c[3] = {rec.cx, rec.cy, rec.cz};
                double dmin2 = 0.0, dmax2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double a = fabs(c[d] - bc[d]);
                    const double e = fmax(0.0, a - bh[d]);
                    dmin2 += e * e;
                    dmax2 += (a + bh[d]) * (a + bh[d]);
                }
                if (isinf(R)) {
                    open = true;
                } else {
                    const double mu = 1e-4 * (R + bhmax);
                    if (dmin2 > (R + mu) * (R + mu)) acc = true;
                    else if (R > mu && dmax2 < (R - mu) * (R - mu)) open = true;
                }
            }
        }
        const bool undecided = active && !acc && !open;
        const u64 am = __ballot(acc), um = __ballot(undecided);
        const int added = (int)__popcll(am), ndef = (int)__popcll(um);
        int4 ka = make_int4(-1, -1, -1, -1), kb = ka;
        if (open) {
            ka = *(const int4*)(kids + 8 * (int64_t)n);
            kb = *(const int4*)(kids + 8 * (int64_t)n + 4);
        }
        int nk = 8 + (ka.x >> 31) + (ka.y >> 31) + (ka.z >> 31) + (ka.w >> 31) + (kb.x >> 31) +
                 (kb.y >> 31) + (kb.z >> 31) + (kb.w >> 31);
        if (!open) nk = 0;
        int total;
        const int pre = wave_excl_count(nk, total);
        // a step whose results would not fit is not taken: its popped ids stay on the
        // stack and are deferred with the rest (so a walk starts with at most
        // kPrefixStart ids on its stack, and its first steps can still pop 64)
        if (held + added > kPrefixCap || nd + ndef + (top - k + total) > kPrefixStart) break;
        wave_sync();  // every lane has read its popped id
        if (acc) list[held + lanes_below(am)] = n;
        if (undecided) defer[nd + lanes_below(um)] = n;
        if (nk > 0) {
            const int w = top - k + pre;
            const int kk[8] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y, kb.z, kb.w};
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < nk) stack[w + q] = kk[q];
        }
        held += added;
        nd += ndef;
        top = top - k + total;
        wave_sync();
    }
    // the deferred nodes, then what is left on the stack, go right after the two counts,
    // the list right after them
    int* const pstack = out + 2;
    if (nd + top > kStackCap) {
        // too many for a walk's stack: the block's walks start at the root
        if (lane == 0) {
            pstack[0] = 0;
            out[0] = n_nodes > 0 ? 1 : 0;
            out[1] = 0;
        }
        return;
    }
    for (int j = lane; j < nd; j += 64) pstack[j] = defer[j];
    for (int j = lane; j < top; j += 64) pstack[nd + j] = stack[j];
    int* const plist = pstack + nd + top;
    for (int j = lane; j < held; j += 64) plist[j] = list[j];
    if (lane == 0) {
        out[0] = nd + top;
        out[1] = held;
    }
}

template <bool F64, bool GUARD>
__global__ __launch_bounds__(kGroupBlock) void k_walk_group(
    const NodeW* __restrict__ nodes, const int* __restrict__ kids, const int* __restrict__ skip,
    int n_nodes, const double* __restrict__ tgt, const double4* __restrict__ stgt,
    const u32* __restrict__ order, int64_t nt, int G, float inv_theta, double eps2,
    double* __restrict__ acc, int* __restrict__ counts, int64_t per_xcd,
    const int* __restrict__ prefixes)
{
    typedef typename GroupEntry<F64>::T Entry;
    __shared__ int stacks[kGroupWaves][kStackCap];
    __shared__ Entry lists[kGroupWaves][kListCap];
    __shared__ u64 failss[kGroupWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int* const stack = stacks[wave];
    Entry* const list = lists[wave];
    u64* const fails = failss[wave];
    // each wavefront walks its own group (no synchronisation between them)
    const int64_t first = (xcd_block(per_xcd) * kGroupWaves + wave) * G;
    if (first >= nt) return;
    const int cnt = (int)(nt - first < G ? nt - first : G);
    const int ti = lane & (G - 1);  // this lane's target within the group
    const int* const prefix =
        prefixes ? prefixes + (first / G / kPrefixGroups) * kPrefixInts : nullptr;
    auto target = [&](int64_t k) -> double4 {
        if (stgt) return stgt[k];
        const int64_t t = order ? (int64_t)order[k] : k;
        return make_double4(tgt[3 * t], tgt[3 * t + 1], tgt[3 * t + 2], 0.0);
    };
    auto store = [&](int64_t k, double fx, double fy, double fz, int listed) {
        const int64_t t = order ? (int64_t)order[k] : k;
        acc[3 * t] = fx;
        acc[3 * t + 1] = fy;
        acc[3 * t + 2] = fz;
        if (counts) counts[t] = listed;
    };
    // the group's frame: one of its targets, in FP64
    const double4 o = target(first + (G / 2 < cnt - 1 ? G / 2 : cnt - 1));
    // this lane's target (lanes past the group's count repeat its last target)
    const double4 q = target(first + (ti < cnt ? ti : cnt - 1));
    const double dx0 = q.x - o.x, dy0 = q.y - o.y, dz0 = q.z - o.z;
    const float tx = (float)dx0, ty = (float)dy0, tz = (float)dz0;
    // the targets' box, in FP32 in the frame (every lane holds one of the targets)
    float lox = tx, loy = ty, loz = tz, hix = tx, hiy = ty, hiz = tz;
    for (int s = 32; s > 0; s >>= 1) {
        lox = fminf(lox, __shfl_xor(lox, s)); hix = fmaxf(hix, __shfl_xor(hix, s));
        loy = fminf(loy, __shfl_xor(loy, s)); hiy = fmaxf(hiy, __shfl_xor(hiy, s));
        loz = fminf(loz, __shfl_xor(loz, s)); hiz = fmaxf(hiz, __shfl_xor(hiz, s));
    }
    // Looseness: the box extent against the geometric mean of the gaps between
    // consecutive targets (max norm), through binary exponents.  A group whose sorted
    // targets jump between distant systems has targets far from its frame, whose FP32
    // coordinates would then carry 2^-24 times that distance; such a group is walked
    // target by target instead.  (With 16 or more targets one long jump moves the
    // geometric mean of the gaps little; groups of 2 to 8 are not checked.)
    const float nx = __shfl(tx, (lane + 1) & 63), ny = __shfl(ty, (lane + 1) & 63),
                nz = __shfl(tz, (lane + 1) & 63);
    int eg = lane < cnt - 1 && lane + 1 < G
                 ? exponent_f(fmaxf(fabsf(nx - tx), fmaxf(fabsf(ny - ty), fabsf(nz - tz))))
                 : 0;
    for (int s = 32; s > 0; s >>= 1) eg += __shfl_xor(eg, s);
    const int ext = exponent_f(fmaxf(hix - lox, fmaxf(hiy - loy, hiz - loz)));
    const bool loose = G >= 16 && cnt > 1 && ext * (cnt - 1) - eg > kLooseLog2 * (cnt - 1);

    if (!loose) {
        // box centre and half-sizes, the half-sizes widened by a few ulp so that rounding
        // never shrinks the box
        const float bcx = 0.5f * (lox + hix), bcy = 0.5f * (loy + hiy),
                    bcz = 0.5f * (loz + hiz);
        const float bhx = fmaxf(hix - bcx, bcx - lox) * 1.000001f;
        const float bhy = fmaxf(hiy - bcy, bcy - loy) * 1.000001f;
        const float bhz = fmaxf(hiz - bcz, bcz - loz) * 1.000001f;
        double ax = 0.0, ay = 0.0, az = 0.0;
        int listed = 0;
        walk_lists<F64, GUARD>(nodes, kids, skip, n_nodes, o.x, o.y, o.z, bcx, bcy, bcz, bhx,
                               bhy, bhz, dx0, dy0, dz0, lane / G, 64 / G, inv_theta, eps2,
                               stack, list, fails, cnt, ax, ay, az, listed, prefix);
        // add the slices of each target (lanes ti, ti + G, ...), in FP64
        for (int s = G; s < 64; s <<= 1) {
            ax += __shfl_xor(ax, s);
            ay += __shfl_xor(ay, s);
            az += __shfl_xor(az, s);
        }
        if (lane < cnt) store(first + lane, ax, ay, az, listed);
        return;
    }
    // target by target, each in its own frame with a point as box (the reference's own
    // opening test), 64 slices
    for (int k = 0; k < cnt; k++) {
        const double4 p = target(first + k);
        double ax = 0.0, ay = 0.0, az = 0.0;
        int listed = 0;
        walk_lists<F64, GUARD>(nodes, kids, skip, n_nodes, p.x, p.y, p.z, 0.f, 0.f, 0.f, 0.f,
                               0.f, 0.f, 0.0, 0.0, 0.0, lane, 64, inv_theta, eps2, stack, list,
                               fails, 1, ax, ay, az, listed, prefix);
        for (int s = 1; s < 64; s <<= 1) {
            ax += __shfl_xor(ax, s);
            ay += __shfl_xor(ay, s);
            az += __shfl_xor(az, s);
        }
        if (lane == 0) store(first + k, ax, ay, az, listed);
    }
}

// Children of every node, in pre-order: kids[8 n + k], -1 after the last.  The children
// of n are n + 1 and then, while below skip[n], the skip pointer of the previous child; a
// leaf (skip[n] = n + 1) has none.  The walk reads only the lists of nodes it opens, but
// leaves' lists are written too: they share cache lines with internal nodes' lists, and
// a line left partly unwritten costs the walk a fetch of the rest from memory.
__global__ __launch_bounds__(kBlock) void k_kids(const int* __restrict__ skip, int n_nodes,
                                                 int* __restrict__ kids)
{
    const int n = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (n >= n_nodes) return;
    const int end = skip[n];
    int k[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int c = n + 1, q = 0;
    while (c < end && q < 8) {
        k[q++] = c;
        c = skip[c];
    }
    int4* o = (int4*)(kids + 8 * (int64_t)n);
    o[0] = make_int4(k[0], k[1], k[2], k[3]);
    o[1] = make_int4(k[4], k[5], k[6], k[7]);
}

// FP64 walk: the reference's opening test decided exactly, the monopole to a few ulp.
//  - Opening test (gravhopper/_jbgrav.c:496-502): the reference accepts an internal
//    node when fl(size / fl(sqrt(d2))) < theta, with d2 = ((dx^2 + dy^2) + dz^2) formed
//    in its operation order.  d2 is formed the same way here.  With R2 = (size/theta)^2
//    the real-number test is d2 > R2, and the two roundings of the reference move q by
//    less than 3 ulp, so d2 > R2 (1 + 1e-13) accepts and d2 < R2 (1 - 1e-13) opens
//    exactly as the reference does.  Only a lane between the two bounds (rare) takes the
//    reference's sqrt and division.  c_lo/c_hi = (1 -/+ 1e-13) / theta^2 (+inf for
//    theta <= 0: nothing is accepted, as size/dist < theta never holds).
//  - Leaves are accepted without the test (:502).
//  - Monopole (:506-524): inv = (1/d)/sqrt(d), d = |com - x|^2 + eps^2, 0 if d == 0, here
//    as r^3 with r = 1/sqrt(d) from v_rsq_f64 and one Newton step.  Measured on gfx950
//    over 2^20 values of d in [1e-30, 1e30] and [1, 4): v_rsq_f64 alone is within
//    2.5e8 ulp, one step within 19.4 ulp (4e-15), two within 1.7 ulp.  One step is far
//    inside the FP64 walk's tolerance against the reference (1e-10), which moment
//    rounding and summation order set anyway, and 4 % faster than two.
// COUNT: counts receives the accepted-node count per target.
__device__ __forceinline__ double rsqrt_f64(double d)
{
    const double r = __builtin_amdgcn_rsq(d);
    return r * __builtin_fma(-0.5 * d, r * r, 1.5);
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_walk_f64(const NodeD* __restrict__ nodes,
                                                     const int* __restrict__ skip,
                                                     int64_t n_nodes,
                                                     const double* __restrict__ tgt,
                                                     const double4* __restrict__ stgt,
                                                     const u32* __restrict__ order, int64_t nt,
                                                     double eps, double theta, double c_lo,
                                                     double c_hi, double* __restrict__ acc,
                                                     int* __restrict__ counts,
                                                     int64_t per_xcd, int lanes)
{
    // one walk of `lanes` consecutive targets per wavefront (lanes <= 64)
    const int lane = threadIdx.x & 63;
    const int64_t first = (xcd_block(per_xcd) * (kWalkBlock / 64) + (threadIdx.x >> 6)) * lanes;
    if (first >= nt) return;  // whole wavefront idle
    const int64_t k = first + lane;
    const bool live = lane < lanes && k < nt;
    int64_t t = 0;
    double x = 0.0, y = 0.0, z = 0.0;
    if (live) {
        t = order ? (int64_t)order[k] : k;
        if (stgt) {  // self walk: the particles in walk order, read contiguously
            const double4 q = stgt[k];
            x = q.x;
            y = q.y;
            z = q.z;
        } else {
            x = tgt[3 * t];
            y = tgt[3 * t + 1];
            z = tgt[3 * t + 2];
        }
    }
    const double eps2 = eps * eps;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int cnt = 0;
    int my_skip = live ? 0 : 0x7fffffff;  // waiting iff node >= my_skip
    const int n_nodes32 = (int)n_nodes;  // below 2^31 (checked by the build)
    int node = 0;
    while (node < n_nodes32) {
        // the whole record (one 64-byte scalar load) and the skip pointer, issued
        // together; a leaf's size has its sign bit set
        const f64x8 nd = *(const f64x8*)(nodes + node);
        const int sk = skip[node];
        const bool is_leaf = __double_as_longlong(nd[3]) < 0;
        const bool active = (u32)my_skip <= (u32)node;
        bool take = active;
        int next = node + 1;
        if (!is_leaf) {  // internal node (a uniform test): opening test
            const double ddx = nd[0] - x, ddy = nd[1] - y, ddz = nd[2] - z;
            double d2 = ddx * ddx;
            d2 = d2 + ddy * ddy;
            d2 = d2 + ddz * ddz;
            const double s2 = nd[3] * nd[3];
            bool far = d2 > s2 * c_hi;
            const bool unsure = active && !far && !(d2 < s2 * c_lo);
            if (__builtin_amdgcn_uicmp((u32)unsure, 0u, 33) != 0ull) {
                if (unsure) far = (nd[3] / sqrt(d2)) < theta;
            }
            const bool opens = active && !far;
            take = active && far;
            next = (__builtin_amdgcn_uicmp((u32)opens, 0u, 33) != 0ull) ? node + 1 : sk;
        }
        if (take) {
            const double dx = nd[4] - x, dy = nd[5] - y, dz = nd[6] - z;
            double d2 = dx * dx;
            d2 = d2 + dy * dy;
            d2 = d2 + dz * dz;
            const double d = d2 + eps2;
            double inv = 0.0;
            if (d != 0.0) {
                const double r = rsqrt_f64(d);
                inv = r * r * r;
            }
            const double mi = nd[7] * inv;
            ax = ax + dx * mi;
            ay = ay + dy * mi;
            az = az + dz * mi;
            if (COUNT) cnt++;
            my_skip = sk;
        }
        node = __builtin_amdgcn_readfirstlane(next);
    }
    if (live) {
        acc[3 * t] = ax;
        acc[3 * t + 1] = ay;
        acc[3 * t + 2] = az;
        if (COUNT) counts[t] = cnt;
    }
}


// ------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------
// Stable LSD sort of (keys[0], idx[0]) by the key bits [lo_shift, hi_shift) of every
// word (word-major, `words` words); returns which of the two buffer pairs holds the
// result.
int radix_sort(gh_ctx* c, DevBuf* keys, DevBuf* idx, DevBuf& hist, int64_t n, int words,
               int lo_shift, hipStream_t s, int* sel_out, int hi_shift, bool have_hist0)
{
    const int nblocks = (int)((n + kRsChunk - 1) / kRsChunk);
    GH_TRY(ensure(c, hist, sizeof(u32) * 256 * (size_t)(nblocks + 1)));
    int sel = 0;
    for (int w = words - 1; w >= 0; w--) {
        for (int shift = lo_shift; shift < hi_shift; shift += 8) {
            const u64* kin = keys[sel].as<u64>();
            if (!(have_hist0 && w == words - 1 && shift == lo_shift)) {
                hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(kBlock), 0, s,
                                   kin + (int64_t)w * n, n, shift, hist.as<u32>(), nblocks);
                GH_TRY(check_launch(c, "k_rs_hist"));
            }
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(kBlock), 0, s, hist.as<u32>(), nblocks,
                               hist.as<u32>() + (int64_t)256 * nblocks);
            GH_TRY(check_launch(c, "k_rs_scan"));
            if (words == 1)
                hipLaunchKernelGGL((k_rs_scatter<true, false>), dim3(nblocks), dim3(kBlock), 0,
                                   s, kin, keys[sel ^ 1].as<u64>(), idx[sel].as<u32>(),
                                   idx[sel ^ 1].as<u32>(), n, words, w, shift, hist.as<u32>(),
                                   nblocks, hist.as<u32>() + (int64_t)256 * nblocks, nullptr);
            else
                hipLaunchKernelGGL((k_rs_scatter<false, false>), dim3(nblocks), dim3(kBlock), 0,
                                   s, kin, keys[sel ^ 1].as<u64>(), idx[sel].as<u32>(),
                                   idx[sel ^ 1].as<u32>(), n, words, w, shift, hist.as<u32>(),
                                   nblocks, hist.as<u32>() + (int64_t)256 * nblocks, nullptr);
            GH_TRY(check_launch(c, "k_rs_scatter"));
            sel ^= 1;
        }
    }
    *sel_out = sel;
    return 0;
}

int radix_sort_direct(gh_ctx* c, DevBuf* keys, DevBuf* idx, DevBuf& hist, int64_t n,
                      int lo_shift, int hi_shift, hipStream_t s, int* sel_out)
{
    const int nblocks = (int)((n + kRsChunk - 1) / kRsChunk);
    if (nblocks > kRsDirectChunks) return fail(c, GH_EARG, "internal: direct sort of %d chunks", nblocks);
    int sel = 0;
    u32* h = hist.as<u32>();
    for (int shift = lo_shift; shift < hi_shift; shift += 8) {
        u32* hn = shift + 8 < hi_shift ? h + (int64_t)256 * nblocks : nullptr;
        hipLaunchKernelGGL((k_rs_scatter<true, true>), dim3(nblocks), dim3(kBlock), 0, s,
                           keys[sel].as<u64>(), keys[sel ^ 1].as<u64>(), idx[sel].as<u32>(),
                           idx[sel ^ 1].as<u32>(), n, 1, 0, shift, h, nblocks, nullptr, hn);
        GH_TRY(check_launch(c, "k_rs_scatter"));
        h += (int64_t)256 * nblocks;
        sel ^= 1;
    }
    *sel_out = sel;
    return 0;
}

// Extents of n points per workgroup into part (*nparts of them), clearing zero[0 ..
// nzero) on the way.
int bbox_partial_dev(gh_ctx* c, const double* d_pos, int64_t n, DevBuf& part, u32* zero,
                     int64_t nzero, hipStream_t s, int* nparts)
{
    const int np = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
    GH_TRY(ensure(c, part, sizeof(double) * 6 * (size_t)np));
    hipLaunchKernelGGL(k_bbox_partial, dim3(np), dim3(kBlock), 0, s, d_pos, n,
                       part.as<double>(), zero, nzero);
    *nparts = np;
    return check_launch(c, "k_bbox_partial");
}

// Root box of the reference (centre box[0..2], size box[3], gravhopper/_jbgrav.c:747-772)
// of n points, written to the device array box; part holds the partial extents.
int bbox_dev(gh_ctx* c, const double* d_pos, int64_t n, double eps, DevBuf& part, double* box,
             hipStream_t s)
{
    int nparts = 0;
    GH_TRY(bbox_partial_dev(c, d_pos, n, part, nullptr, 0, s, &nparts));
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, s, part.as<double>(), nparts,
                       eps, box);
    return check_launch(c, "k_bbox_final");
}

// Builds the tree of (d_pos, d_mass) into c->tree.  One host synchronisation reads the
// node count and the key checks.
static int build_tree(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t n,
                      double eps, double theta, int what, hipStream_t s)
{
    if (!c->tree) c->tree = new TreeBufs();
    TreeBufs* tb = c->tree;
    TimedScope ts(c, s, GH_T_TREE_BUILD);

    // 1. box
    GH_TRY(ensure(c, tb->box, sizeof(double) * 8));
    GH_TRY(bbox_dev(c, d_pos, n, eps, tb->bbox_part, tb->box.as<double>(), s));

    GH_TRY(ensure(c, tb->idx[0], sizeof(u32) * (size_t)n));
    GH_TRY(ensure(c, tb->idx[1], sizeof(u32) * (size_t)n));
    GH_TRY(ensure(c, tb->delta, (size_t)search_buffer_bytes(n)));
    unsigned char* const dp = tb->delta.as<unsigned char>();
    const int search_lv = search_levels(n);
    GH_TRY(ensure(c, tb->flags, sizeof(int) * 4));
    if (!tb->readback) {
        // coherent, so that the kernel's stores reach the host without any cache flush
        GH_HIP(c, hipHostMalloc((void**)&tb->readback, 4 * sizeof(long long),
                                hipHostMallocMapped | hipHostMallocCoherent));
        GH_HIP(c, hipHostGetDevicePointer((void**)&tb->readback_dev, tb->readback, 0));
    }
    const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
    GH_TRY(ensure(c, tb->off, sizeof(long long) * (size_t)(2 * n + 2)));
    GH_TRY(ensure(c, tb->block_sums, sizeof(DD4) * (size_t)(nb + 1)));
    GH_TRY(ensure(c, tb->block_counts, sizeof(long long) * (size_t)(nb + 1)));
    GH_TRY(ensure(c, tb->prefix, sizeof(DD4) * (size_t)(n + 1)));
    GH_TRY(ensure(c, tb->spos, sizeof(double4) * (size_t)n));
    GH_TRY(ensure(c, tb->packed, sizeof(double4) * (size_t)n));
    long long* cnt = tb->off.as<long long>() + (n + 1);
    long long* off = tb->off.as<long long>();
    DD4* pref = tb->prefix.as<DD4>();
    ScanView<OpI64> off_view;
    ScanView<OpDD4> pref_view;
    long long total = 0;
    // Whether one-word keys are sorted over all bits, or by bits kSortLowBits.. and then
    // run by run (k_sort_runs).  After a build with a run too long for k_sort_runs the
    // next kFullSortBuilds builds of the context sort over all bits.
    constexpr int kFullSortBuilds = 64;
    const int sort_mode = c->tune[GH_TUNE_TREE_SORT];
    bool all_bits = sort_mode < 0 || (sort_mode == 0 && tb->full_sorts > 0);
    if (sort_mode == 0 && tb->full_sorts > 0) tb->full_sorts--;
    // Keys with `words` words, sort, shared levels, numbering and moments, then one
    // readback of (flags, node count).  Keys still equal over all words (rare) mean
    // particles closer than words * 21 levels: the pass is redone with one more word.
    for (int words = 1;; words++) {
        if (words > kMaxWords)
            return fail(c, GH_EDEPTH,
                        "particles closer than %d octree levels below the root box",
                        kMaxWords * kLevelsPerWord);
        GH_TRY(ensure(c, tb->keys[0], sizeof(u64) * (size_t)n * words));
        GH_TRY(ensure(c, tb->keys[1], sizeof(u64) * (size_t)n * words));
        // 2. keys; the kernel also counts the first radix pass's digits, and its first
        // threads clear the flags read after k_delta and k_sort_runs
        const bool by_runs = words == 1 && !all_bits;
        const int lo_bits = by_runs ? kSortLowBits : 0;
        GH_TRY(ensure(c, tb->hist, sizeof(u32) * 256 * (size_t)(blocks_for(n, kRsChunk) + 1)));
        hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, kRsChunk)), dim3(kKeysBlock), 0, s, d_pos,
                           n, tb->box.as<double>(), words, tb->keys[0].as<u64>(),
                           tb->idx[0].as<u32>(), tb->flags.as<int>(), d_mass,
                           tb->packed.as<double4>(), tb->hist.as<u32>(), words - 1, lo_bits);
        GH_TRY(check_launch(c, "k_keys"));
        // 3. sort
        GH_TRY(radix_sort(c, tb->keys, tb->idx, tb->hist, n, words, lo_bits, s, &tb->key_sel,
                          64, true));
        if (by_runs) {
            hipLaunchKernelGGL(k_sort_runs, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s,
                               tb->keys[tb->key_sel].as<u64>(), tb->idx[tb->key_sel].as<u32>(),
                               n, tb->flags.as<int>() + 2);
            GH_TRY(check_launch(c, "k_sort_runs"));
        }
        // 4./5. shared levels, end-search levels, node counts, sorted particles
        DeltaArgs da;
        da.keys = tb->keys[tb->key_sel].as<u64>();
        da.perm = tb->idx[tb->key_sel].as<u32>();
        da.packed = tb->packed.as<double4>();
        da.search = dp;
        da.cnt = cnt;
        da.spos = tb->spos.as<double4>();
        da.flags = tb->flags.as<int>();
        hipLaunchKernelGGL(k_delta, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, da, n,
                           words);
        GH_TRY(check_launch(c, "k_delta"));
        // 5./6. first node ids (scan of the node counts) and moments (double-double
        // prefix sums of (m, m x, m y, m z)), each kept as chunk prefixes plus the
        // prefix of the chunk totals (ScanView); the end-search levels above 2; and the
        // readback of (flags, node count)
        if (nb > 0) {
            hipLaunchKernelGGL(k_scan_local_fill, dim3((unsigned)nb), dim3(kBlock), 0, s, cnt,
                               tb->spos.as<double4>(), n, off,
                               tb->block_counts.as<long long>(), pref,
                               tb->block_sums.as<DD4>());
            GH_TRY(check_launch(c, "k_scan_local_fill"));
        }
        tb->readback[0] = tb->readback[1] = -1;
        hipLaunchKernelGGL(k_scan_sums_fill, dim3(2), dim3(kBlock), 0, s,
                           tb->block_counts.as<long long>(), tb->block_sums.as<DD4>(), nb,
                           off + n, pref + n, dp, n, search_lv, tb->flags.as<int>(),
                           tb->readback_dev);
        GH_TRY(check_launch(c, "k_scan_sums_fill"));
        off_view.local = off;
        off_view.blocks = tb->block_counts.as<long long>();
        off_view.n = n;
        pref_view.local = pref;
        pref_view.blocks = tb->block_sums.as<DD4>();
        pref_view.n = n;
        GH_HIP(c, hipStreamSynchronize(s));
        if (tb->readback[1] < 0)
            return fail(c, GH_EHIP, "internal: the tree build's readback was not written");
        if (tb->readback[2] > 0) {
            // a run too long for k_sort_runs, so the order (and what k_delta found in it)
            // is not the keys' order: the same words again, sorted over all bits
            all_bits = true;
            if (sort_mode == 0) tb->full_sorts = kFullSortBuilds;
            words--;
            continue;
        }
        const int* got = (const int*)tb->readback;
        if (got[1] > 0)
            return fail(c, GH_ECOINCIDENT,
                        "%d pair(s) of particles share a position; the octree cannot "
                        "separate them", got[1]);
        if (got[0] == 0) {
            tb->words = words;
            total = tb->readback[1];
            break;
        }
    }
    const u32* perm = tb->idx[tb->key_sel].as<u32>();
    if (total <= 0 || total >= ((long long)1 << 31))
        return fail(c, GH_EDEPTH, "octree of %lld nodes is out of range", total);
    tb->n_nodes = total;
    const size_t T = (size_t)total;
    GH_TRY(ensure(c, tb->skip, sizeof(int) * T));
    if (what & (kFillF64 | kFillDump)) {
        GH_TRY(ensure(c, tb->nodes64, sizeof(NodeD) * T));
        GH_TRY(ensure(c, tb->leaf, sizeof(int) * T));
    }
    if (what & kFillKids) GH_TRY(ensure(c, tb->kids, 8 * sizeof(int) * T));
    if (what & kFillW) GH_TRY(ensure(c, tb->nodesw, sizeof(NodeW) * T));
    if (what & kFillDump) {
        GH_TRY(ensure(c, tb->level, sizeof(short) * T));
        GH_TRY(ensure(c, tb->branch, T));
    }

    FillArgs fa;
    fa.perm = perm;
    fa.dp = dp;
    fa.es = end_search(dp, n);
    fa.off = off_view;
    fa.spos = tb->spos.as<double4>();
    fa.box = tb->box.as<double>();
    fa.pref = pref_view;
    fa.nodes64 = tb->nodes64.as<NodeD>();
    fa.nodesw = tb->nodesw.as<NodeW>();
    fa.skip = tb->skip.as<int>();
    fa.level = tb->level.as<short>();
    fa.branch = tb->branch.as<signed char>();
    fa.leaf = tb->leaf.as<int>();
    fa.what = what;
    hipLaunchKernelGGL(k_fill_nodes, dim3(blocks_for(n, kFillBlock)), dim3(kFillBlock), 0, s, fa,
                       n);
    GH_TRY(check_launch(c, "k_fill_nodes"));
    if (what & kFillKids) {
        hipLaunchKernelGGL(k_kids, dim3(blocks_for(total, kBlock)), dim3(kBlock), 0, s,
                           tb->skip.as<int>(), (int)total, tb->kids.as<int>());
        GH_TRY(check_launch(c, "k_kids"));
    }
    c->last_tree_nodes = total;
    return 0;
}

// Walks nt targets through the tree built last: the particles themselves (self) or the
// positions tgt.  Targets are walked in octant-path (Morton) order, so that neighbouring
// lanes take neighbouring targets and a wavefront's paths through the tree nearly
// coincide: self targets in the particles' sorted order, kSortTargetsMin or more
// position targets in an order sorted here by the top 40 key bits (13 levels of the
// particles' root box; targets outside the box fall in its edge cells).  Results are
// written in the caller's order.
static int walk(gh_ctx* c, const double* tgt, bool self, int64_t nt, double eps, double theta,
                int precision, double* d_acc, int* counts, hipStream_t s)
{
    TreeBufs* tb = c->tree;
    if (nt <= 0) return 0;
    TimedScope ts(c, s, GH_T_TREE_WALK);
    const u32* order = nullptr;
    const double4* stgt = nullptr;  // self walks read the sorted particles
    if (self) {
        order = tb->idx[tb->key_sel].as<u32>();
        stgt = tb->spos.as<double4>();
    } else if (nt >= kSortTargetsMin) {
        for (int k = 0; k < 2; k++) {
            GH_TRY(ensure(c, tb->tkeys[k], sizeof(u64) * (size_t)nt));
            GH_TRY(ensure(c, tb->tidx[k], sizeof(u32) * (size_t)nt));
        }
        hipLaunchKernelGGL(k_keys, dim3(blocks_for(nt, kRsChunk)), dim3(kKeysBlock), 0, s, tgt,
                           nt, tb->box.as<double>(), 1, tb->tkeys[0].as<u64>(),
                           tb->tidx[0].as<u32>(), (int*)nullptr, (const double*)nullptr,
                           (double4*)nullptr, (u32*)nullptr, 0, 0);
        GH_TRY(check_launch(c, "k_keys"));
        int sel = 0;
        GH_TRY(radix_sort(c, tb->tkeys, tb->tidx, tb->hist, nt, 1, 24, s, &sel));
        order = tb->tidx[sel].as<u32>();
    }
    // Targets per walk: 64 (a full wavefront) when there are enough targets to fill the
    // GPU; fewer targets give each walk fewer lanes, so that about kWalkUnits walks run
    // side by side (a walk's length grows with the spread of its targets, and a few
    // long walks would leave most of the GPU idle).
    const int lanes = (int)std::min<int64_t>(64, std::max<int64_t>(1, (nt + kWalkUnits - 1) /
                                                                          kWalkUnits));
    const int waves_per_block = kWalkBlock / 64;
    if (precision == GH_FP64) {
        const int64_t n_walks = (nt + lanes - 1) / lanes;
        const int64_t per = (blocks_for(n_walks, waves_per_block) + kXcds - 1) / kXcds;
        const double inf = __builtin_huge_val();
        const double c_lo = theta > 0.0 ? (1.0 - 1e-13) / (theta * theta) : inf;
        const double c_hi = theta > 0.0 ? (1.0 + 1e-13) / (theta * theta) : inf;
        if (counts)
            hipLaunchKernelGGL(k_walk_f64<true>, dim3((unsigned)(per * kXcds)), dim3(kWalkBlock),
                               0, s, tb->nodes64.as<NodeD>(), tb->skip.as<int>(),
                               tb->n_nodes, tgt, stgt, order, nt, eps, theta,
                               c_lo, c_hi, d_acc, counts, per, lanes);
        else
            hipLaunchKernelGGL(k_walk_f64<false>, dim3((unsigned)(per * kXcds)), dim3(kWalkBlock),
                               0, s, tb->nodes64.as<NodeD>(), tb->skip.as<int>(),
                               tb->n_nodes, tgt, stgt, order, nt, eps, theta,
                               c_lo, c_hi, d_acc, counts, per, lanes);
        return check_launch(c, "k_walk_f64");
    }
    // FP32: the group walk, 16 targets per wavefront, or one target per wavefront when
    // there are too few targets for kWalkUnits groups of 16 (then every target walks in
    // its own frame with the reference's opening test)
    int G = c->tune[GH_TUNE_WALK_GROUP];
    if (!(G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64))
        G = (nt >= 16 * kWalkUnits) ? 16 : 1;
    const bool f64eval = c->tune[GH_TUNE_WALK_F64EVAL] > 0;
    const bool guard = fp32_guard(eps);
    const float inv_theta = theta > 0.0 ? (float)(1.0 / theta) : __builtin_huge_valf();
    const int64_t n_groups = (nt + G - 1) / G;
    const int64_t n_blocks = (n_groups + kGroupWaves - 1) / kGroupWaves;
    const int64_t per = (n_blocks + kXcds - 1) / kXcds;
    const dim3 grid((unsigned)(per * kXcds));
    const NodeW* nd = tb->nodesw.as<NodeW>();
    const int* kd = tb->kids.as<int>();
    const int* sk = tb->skip.as<int>();
    const int T = (int)tb->n_nodes;
    const double eps2 = eps * eps;
    // the walks' shared first steps, per block of kPrefixGroups groups
    int* pf = nullptr;
    if (c->tune[GH_TUNE_WALK_PREFIX] >= 0) {
        const int64_t nblk = (n_groups + kPrefixGroups - 1) / kPrefixGroups;
        GH_TRY(ensure(c, tb->wprefix, sizeof(int) * (size_t)kPrefixInts * (size_t)nblk));
        pf = tb->wprefix.as<int>();
        hipLaunchKernelGGL(k_walk_prefix, dim3((unsigned)nblk), dim3(64), 0, s, nd, kd, T, tgt,
                           stgt, order, nt, G, theta > 0.0 ? 1.0 / theta : HUGE_VAL, pf);
        GH_TRY(check_launch(c, "k_walk_prefix"));
    }
#define GH_GWALK(F, GU)                                                                    \
    hipLaunchKernelGGL((k_walk_group<F, GU>), grid, dim3(kGroupBlock), 0, s, nd, kd, sk, T, tgt, \
                       stgt, order, nt, G, inv_theta, eps2, d_acc, counts, per, pf)
    if (f64eval)
        GH_GWALK(true, false);
    else if (guard)
        GH_GWALK(false, true);
    else
        GH_GWALK(false, false);
#undef GH_GWALK
    return check_launch(c, "k_walk_group");
}

int tree_dev(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np,
             const double* d_fpos, int64_t nf, double eps, double theta, int precision,
             double* d_acc, hipStream_t s)
{
    const bool self = (d_fpos == nullptr);
    const int64_t nt = self ? np : nf;
    if (nt == 0) return 0;
    if (np == 0) return zero_dev(c, d_acc, 3 * nt, s);
    if (precision != GH_FP64 && precision != GH_FP32)
        return fail(c, GH_EARG, "unknown precision %d", precision);
    GH_TRY(build_tree(c, d_pos, d_mass, np, eps, theta,
                      precision == GH_FP64 ? kFillF64 : (kFillW | kFillKids), s));
    return walk(c, self ? d_pos : d_fpos, self, nt, eps, theta, precision, d_acc, nullptr, s);
}

void tree_free(gh_ctx* c)
{
    if (!c->tree) return;
    TreeBufs* tb = c->tree;
    DevBuf* bufs[] = {&tb->keys[0], &tb->keys[1], &tb->idx[0],    &tb->idx[1],  &tb->hist,
                      &tb->bbox_part, &tb->box,   &tb->flags,     &tb->delta,   &tb->off,
                      &tb->prefix,  &tb->spos,      &tb->block_sums, &tb->block_counts,
                      &tb->packed,
                      &tb->nodes64, &tb->nodesw, &tb->kids,  &tb->wprefix,
                      &tb->skip,    &tb->level,   &tb->branch,    &tb->leaf,    &tb->tkeys[0],
                      &tb->tkeys[1], &tb->tidx[0], &tb->tidx[1]};
    for (DevBuf* b : bufs) release(c, *b);
    if (tb->readback) (void)hipHostFree(tb->readback);
    delete tb;
    c->tree = nullptr;
}

}  // namespace gh

using namespace gh;

extern "C" int gh_tree_dump(gh_ctx* c, const double* pos, const double* mass, int64_t np,
                            double eps, gh_node_rec* out, int64_t cap, int64_t* n_nodes)
{
    if (!c) return GH_EARG;
    if (np < 1) return fail(c, GH_EARG, "the tree needs at least one particle");
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    GH_TRY(ensure(c, c->pos64, sizeof(double) * 3 * (size_t)np));
    GH_TRY(ensure(c, c->mass64, sizeof(double) * (size_t)np));
    GH_HIP(c, hipMemcpyAsync(c->pos64.p, pos, sizeof(double) * 3 * (size_t)np,
                             hipMemcpyHostToDevice, s));
    GH_HIP(c, hipMemcpyAsync(c->mass64.p, mass, sizeof(double) * (size_t)np,
                             hipMemcpyHostToDevice, s));
    GH_TRY(build_tree(c, c->pos64.as<double>(), c->mass64.as<double>(), np, eps, 0.7,
                      kFillF64 | kFillDump, s));
    TreeBufs* tb = c->tree;
    const int64_t T = tb->n_nodes;
    if (n_nodes) *n_nodes = T;
    const int64_t m = cap < T ? cap : T;
    if (m <= 0 || !out) return 0;
    std::vector<NodeD> nd(m);
    std::vector<short> lv(m);
    std::vector<signed char> br(m);
    std::vector<int> lf(m);
    GH_HIP(c, hipMemcpyAsync(nd.data(), tb->nodes64.p, sizeof(NodeD) * m, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(lv.data(), tb->level.p, sizeof(short) * m, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(br.data(), tb->branch.p, (size_t)m, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(lf.data(), tb->leaf.p, sizeof(int) * m, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipStreamSynchronize(s));
    for (int64_t k = 0; k < m; k++) {
        gh_node_rec& r = out[k];
        r.level = lv[k];
        r.branch = br[k];
        r.leaf = lf[k];
        r.center[0] = nd[k].cx;
        r.center[1] = nd[k].cy;
        r.center[2] = nd[k].cz;
        r.size = std::fabs(nd[k].size);
        r.mass = nd[k].mass;
        r.com[0] = nd[k].mx;
        r.com[1] = nd[k].my;
        r.com[2] = nd[k].mz;
    }
    return 0;
}

// Interaction counts per target: the FP64 walk's accepted nodes (comparable with the
// reference's walk) or the FP32 group walk's list lengths.
static int tree_counts(gh_ctx* c, const double* pos, const double* mass, int64_t np,
                       const double* force_pos, int64_t nf, double eps, double theta,
                       int precision, double* acc, int64_t* counts)
{
    if (!c) return GH_EARG;
    if (np < 1) return fail(c, GH_EARG, "the tree needs at least one particle");
    const bool self = force_pos == nullptr;
    const int64_t nt = self ? np : nf;
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    GH_TRY(ensure(c, c->pos64, sizeof(double) * 3 * (size_t)np));
    GH_TRY(ensure(c, c->mass64, sizeof(double) * (size_t)np));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1)));
    GH_TRY(ensure(c, c->scratch, sizeof(int) * (size_t)(nt > 0 ? nt : 1)));
    GH_HIP(c, hipMemcpyAsync(c->pos64.p, pos, sizeof(double) * 3 * (size_t)np,
                             hipMemcpyHostToDevice, s));
    GH_HIP(c, hipMemcpyAsync(c->mass64.p, mass, sizeof(double) * (size_t)np,
                             hipMemcpyHostToDevice, s));
    const double* tgt = c->pos64.as<double>();
    if (!self) {
        GH_TRY(ensure(c, c->fpos64, sizeof(double) * 3 * (size_t)(nt > 0 ? nt : 1)));
        GH_HIP(c, hipMemcpyAsync(c->fpos64.p, force_pos, sizeof(double) * 3 * (size_t)nt,
                                 hipMemcpyHostToDevice, s));
        tgt = c->fpos64.as<double>();
    }
    GH_TRY(build_tree(c, c->pos64.as<double>(), c->mass64.as<double>(), np, eps, theta,
                      precision == GH_FP64 ? kFillF64 : (kFillW | kFillKids), s));
    GH_TRY(walk(c, tgt, self, nt, eps, theta, precision, c->acc64.as<double>(),
                c->scratch.as<int>(), s));
    std::vector<int> cnt((size_t)(nt > 0 ? nt : 1));
    GH_HIP(c, hipMemcpyAsync(acc, c->acc64.p, sizeof(double) * 3 * (size_t)nt,
                             hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(cnt.data(), c->scratch.p, sizeof(int) * (size_t)nt,
                             hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipStreamSynchronize(s));
    for (int64_t k = 0; k < nt; k++) counts[k] = cnt[k];
    return 0;
}

extern "C" int gh_tree_interactions(gh_ctx* c, const double* pos, const double* mass,
                                    int64_t np, const double* force_pos, int64_t nf, double eps,
                                    double theta, double* acc, int64_t* counts)
{
    return tree_counts(c, pos, mass, np, force_pos, nf, eps, theta, GH_FP64, acc, counts);
}

extern "C" int gh_tree_interactions_fp32(gh_ctx* c, const double* pos, const double* mass,
                                         int64_t np, const double* force_pos, int64_t nf,
                                         double eps, double theta, double* acc, int64_t* counts)
{
    return tree_counts(c, pos, mass, np, force_pos, nf, eps, theta, GH_FP32, acc, counts);
}
