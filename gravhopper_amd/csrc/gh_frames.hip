// Spatial order and local frames of the FP32 pair kernels.
//
// An FP32 pair force needs the separation x_j - x_i.  Casting absolute FP64 positions to
// FP32 first and subtracting afterwards loses 2^-24 |x| per coordinate, which for a
// star cluster 8 kpc from the coordinate origin is a relative error of order 1e-4 on
// its internal forces (the reference forms every difference in FP64,
// gravhopper/_jbgrav.c:155-168).  The FP32 kernels therefore never cast an absolute
// position.  Points are sorted along a Morton curve (about 11 levels in their root box) and
// cut into groups of consecutive sorted points (the kernels' blocks); each group g has
// an FP64 origin o_g, one of its own points, and the kernels work in the frame of one
// group: a point of group h enters as fl(x - o_h) + fl(o_h - o_g), both formed from FP64
// values.  The rounding error of a separation is then bounded by 2^-24 times the extents
// of the groups involved plus the separation itself, independent of where the system
// sits.
//
// A sorted group is compact unless the curve jumps between two clusters inside it (two
// galaxies far apart, or the rare large jump of the curve itself).  k_frames measures
// this per group: R = max |x - o| against the geometric mean s of the gaps between
// consecutive sorted points of the group; R > 2^kFrameLog2 s flags the group, and pairs
// involving a flagged group are computed in double-single form (hi + lo FP32 pairs of
// FP64 differences), whose separations are exact to 2^-24 |x_j - x_i| whatever the
// group's shape.
#include <algorithm>

#include "gh_box.h"
#include "gh_internal.h"

namespace gh {

// 21-bit integer spread to every third bit (bit k of v goes to bit 3k).
__device__ __forceinline__ u64 spread3(u64 v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// The sort uses key bits [kKeyShift, 63) of the 63-bit Morton key.
constexpr int kKeyShift = 31;

__device__ __forceinline__ u32 quantize(double x, double lo, double scale)
{
    double q = (x - lo) * scale;
    q = q > 0.0 ? q : 0.0;  // also maps NaN to 0
    return q < 2097151.0 ? (u32)q : 2097151u;
}

// Morton key of 21 levels (bit 3k + a: level 20 - k, axis a) of point i in the box
// box[0..2] = centre, box[3] = size.
__device__ __forceinline__ u64 morton_key(const double* __restrict__ pos, int64_t i,
                                          const double* box)
{
    const double size = box[3];
    const double scale = size > 0.0 ? 2097152.0 / size : 0.0;
    u64 k = 0;
    for (int a = 0; a < 3; a++)
        k |= spread3(quantize(pos[3 * i + a], box[a] - 0.5 * size, scale)) << a;
    return k;
}

// Morton keys of the points; idx = identity.
__global__ __launch_bounds__(kBlock) void k_morton(const double* __restrict__ pos, int64_t n,
                                                   const double* __restrict__ box,
                                                   u64* __restrict__ keys,
                                                   u32* __restrict__ idx,
                                                   int* __restrict__ nflag)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) *nflag = 0;
    if (i >= n) return;
    double b[4];
    for (int k = 0; k < 4; k++) b[k] = box[k];
    keys[i] = morton_key(pos, i, b);
    idx[i] = (u32)i;
}

// The same for a sort of at most kRsDirectChunks chunks: every workgroup forms the root
// box from the nparts extents itself (workgroup 0 stores it), and the digit counts of
// the sort's first pass (key bits kKeyShift ..) are added into h0 (digit d, chunk b at
// d * nb + b; zero beforehand).
__global__ __launch_bounds__(kBlock) void k_morton_small(const double* __restrict__ pos,
                                                         int64_t n,
                                                         const double* __restrict__ part,
                                                         int nparts, double* __restrict__ box,
                                                         u64* __restrict__ keys,
                                                         u32* __restrict__ idx,
                                                         int* __restrict__ nflag,
                                                         u32* __restrict__ h0, int nb)
{
    __shared__ double smin[3][kBlock], smax[3][kBlock];
    __shared__ u32 h[256];
    double b[4];
    box_from_parts(part, nparts, 0.0, smin, smax, b);
    if (blockIdx.x == 0 && threadIdx.x < 4) box[threadIdx.x] = b[threadIdx.x];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) *nflag = 0;
    if (i < n) {
        const u64 k = morton_key(pos, i, b);
        keys[i] = k;
        idx[i] = (u32)i;
        atomicAdd(&h[(u32)(k >> kKeyShift) & 255u], 1u);
    }
    __syncthreads();
    const int chunk = (int)((int64_t)blockIdx.x * kBlock / kRsChunk);
    const u32 v = h[threadIdx.x];
    if (v) atomicAdd(&h0[threadIdx.x * nb + chunk], v);
}

int spatial_order(gh_ctx* c, OrderBufs& o, const double* d_pos, int64_t n, hipStream_t s)
{
    o.sorted = nullptr;
    if (n <= 0) return 0;
    for (int k = 0; k < 2; k++) {
        GH_TRY(ensure(c, o.keys[k], sizeof(u64) * (size_t)n));
        GH_TRY(ensure(c, o.idx[k], sizeof(u32) * (size_t)n));
    }
    GH_TRY(ensure(c, o.box, sizeof(double) * 8));
    GH_TRY(ensure(c, o.nflag, 256));
    const int nb = (int)((n + kRsChunk - 1) / kRsChunk);
    if (nb <= kRsDirectChunks && c->tune[GH_TUNE_SORT_DIRECT] >= 0) {
        // few chunks: six kernels instead of fifteen (the same order, bit for bit)
        const int passes = (63 - kKeyShift + 7) / 8;
        const int64_t nh = (int64_t)256 * nb * passes;
        GH_TRY(ensure(c, o.hist, sizeof(u32) * (size_t)nh));
        int nparts = 0;
        GH_TRY(bbox_partial_dev(c, d_pos, n, o.boxpart, o.hist.as<u32>(), nh, s, &nparts));
        hipLaunchKernelGGL(k_morton_small, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s,
                           d_pos, n, o.boxpart.as<double>(), nparts, o.box.as<double>(),
                           o.keys[0].as<u64>(), o.idx[0].as<u32>(), o.nflag.as<int>(),
                           o.hist.as<u32>(), nb);
        GH_TRY(check_launch(c, "k_morton_small"));
        GH_TRY(radix_sort_direct(c, o.keys, o.idx, o.hist, n, kKeyShift, 63, s, &o.sel));
        o.sorted = d_pos;
        o.sorted_n = n;
        return 0;
    }
    GH_TRY(bbox_dev(c, d_pos, n, 0.0, o.boxpart, o.box.as<double>(), s));
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, d_pos, n,
                       o.box.as<double>(), o.keys[0].as<u64>(), o.idx[0].as<u32>(),
                       o.nflag.as<int>());
    GH_TRY(check_launch(c, "k_morton"));
    // bits 31..62 (the top 10 2/3 levels): four 8-bit passes.  Deeper keys change the
    // blocks' compactness, and so the forces' rounding, by little (Plummer sphere,
    // 65536 particles, 1024-particle blocks: median relative error 3.0e-8 with 10 levels,
    // 3.4e-8 with 16) and cost two more passes.
    GH_TRY(radix_sort(c, o.keys, o.idx, o.hist, n, 1, kKeyShift, s, &o.sel, 63));
    o.sorted = d_pos;
    o.sorted_n = n;
    return 0;
}

// Root box (as k_bbox_final with eps = 0) of each segment's real points, one workgroup
// per segment: box[4 g ..] = centre, size.
__global__ __launch_bounds__(kBlock) void k_seg_box(const double* __restrict__ pos,
                                                    const int64_t* __restrict__ start,
                                                    const int64_t* __restrict__ count,
                                                    double* __restrict__ box)
{
    __shared__ double smin[3][kBlock], smax[3][kBlock];
    const int64_t a = start[blockIdx.x], n = count[blockIdx.x];
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) mn[k] = mx[k] = pos[3 * a + k];
    for (int64_t i = threadIdx.x; i < n; i += kBlock)
        for (int k = 0; k < 3; k++) {
            const double q = pos[3 * (a + i) + k];
            mn[k] = q < mn[k] ? q : mn[k];
            mx[k] = q > mx[k] ? q : mx[k];
        }
    for (int k = 0; k < 3; k++) {
        smin[k][threadIdx.x] = mn[k];
        smax[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                if (smin[k][threadIdx.x + w] < smin[k][threadIdx.x])
                    smin[k][threadIdx.x] = smin[k][threadIdx.x + w];
                if (smax[k][threadIdx.x + w] > smax[k][threadIdx.x])
                    smax[k][threadIdx.x] = smax[k][threadIdx.x + w];
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double bs = smax[0][0] - smin[0][0];
    for (int k = 1; k < 3; k++)
        if ((smax[k][0] - smin[k][0]) > bs) bs = smax[k][0] - smin[k][0];
    for (int k = 0; k < 3; k++) box[4 * blockIdx.x + k] = 0.5 * (smin[k][0] + smax[k][0]);
    box[4 * blockIdx.x + 3] = bs;
}

// Keys of the padded, segmented layout: position p of block b belongs to segment
// seg[b]; its key is (segment << 32) | the key bits spatial_order sorts by (in the
// segment's box), and padding positions take their segment's largest key, so that a
// stable sort keeps every segment in place, in the order a single sort of it gives,
// with its padding last.
__global__ __launch_bounds__(kBlock) void k_morton_seg(const double* __restrict__ pos,
                                                       int64_t n_pad, int blk,
                                                       const int* __restrict__ seg,
                                                       const int64_t* __restrict__ start,
                                                       const int64_t* __restrict__ count,
                                                       const double* __restrict__ box,
                                                       u64* __restrict__ keys,
                                                       u32* __restrict__ idx,
                                                       int* __restrict__ nflag)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) *nflag = 0;
    if (p >= n_pad) return;
    const int g = seg[p / blk];
    u64 k = ((u64)1 << (63 - kKeyShift)) - 1;
    if (p - start[g] < count[g]) {
        const double* bx = box + 4 * g;
        const double size = bx[3];
        const double scale = size > 0.0 ? 2097152.0 / size : 0.0;
        u64 m = 0;
        for (int a = 0; a < 3; a++)
            m |= spread3(quantize(pos[3 * p + a], bx[a] - 0.5 * size, scale)) << a;
        k = m >> kKeyShift;
    }
    keys[p] = ((u64)g << (63 - kKeyShift)) | k;
    idx[p] = (u32)p;
}

int spatial_order_seg(gh_ctx* c, OrderBufs& o, const double* d_pos, int64_t n_pad, int blk,
                      int nseg, const int* seg, const int64_t* start, const int64_t* count,
                      hipStream_t s)
{
    o.sorted = nullptr;
    if (n_pad <= 0) return 0;
    for (int k = 0; k < 2; k++) {
        GH_TRY(ensure(c, o.keys[k], sizeof(u64) * (size_t)n_pad));
        GH_TRY(ensure(c, o.idx[k], sizeof(u32) * (size_t)n_pad));
    }
    GH_TRY(ensure(c, o.box, sizeof(double) * 4 * (size_t)nseg));
    GH_TRY(ensure(c, o.nflag, 256));
    hipLaunchKernelGGL(k_seg_box, dim3((unsigned)nseg), dim3(kBlock), 0, s, d_pos, start, count,
                       o.box.as<double>());
    GH_TRY(check_launch(c, "k_seg_box"));
    hipLaunchKernelGGL(k_morton_seg, dim3(blocks_for(n_pad, kBlock)), dim3(kBlock), 0, s, d_pos,
                       n_pad, blk, seg, start, count, o.box.as<double>(), o.keys[0].as<u64>(),
                       o.idx[0].as<u32>(), o.nflag.as<int>());
    GH_TRY(check_launch(c, "k_morton_seg"));
    int bits = 63 - kKeyShift;
    while (bits < 64 && ((int64_t)1 << (bits - (63 - kKeyShift))) < nseg) bits++;
    return radix_sort(c, o.keys, o.idx, o.hist, n_pad, 1, 0, s, &o.sel, bits);
}

__device__ __forceinline__ int exponent_of(double x)
{
    // binary exponent of |x| (x = 0 and subnormals give -1023)
    return (int)((__double_as_longlong(x) >> 52) & 0x7ff) - 1023;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ long long wave_sum(long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One workgroup per group of G consecutive sorted points.  Origin: the group's point at
// offset min(G/2, count-1) (a point of the group's majority part, so that a group with a
// few strays keeps a useful frame).  Writes loc / s64 for the group's G slots (padding:
// massless at the origin), frame = (origin, R) and the compactness flag.
template <int G>
__global__ __launch_bounds__(kBlock) void k_frames(const double* __restrict__ pos,
                                                   const double* __restrict__ mass,
                                                   const u32* __restrict__ perm, int64_t n,
                                                   const int* __restrict__ valid,
                                                   float4* __restrict__ loc,
                                                   double4* __restrict__ s64,
                                                   double4* __restrict__ frame,
                                                   int* __restrict__ flag,
                                                   int* __restrict__ nflag)
{
    __shared__ double sR[kBlock / 64];
    __shared__ long long sE[kBlock / 64], sC[kBlock / 64];
    const int64_t p0 = (int64_t)blockIdx.x * G;
    int cnt = valid ? valid[blockIdx.x] : (int)(n - p0 < G ? n - p0 : G);
    if (cnt < 0) cnt = 0;
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (cnt > 0) {
        const int64_t pm = p0 + (G / 2 < cnt - 1 ? G / 2 : cnt - 1);
        const int64_t jm = perm ? (int64_t)perm[pm] : pm;
        ox = pos[3 * jm];
        oy = pos[3 * jm + 1];
        oz = pos[3 * jm + 2];
    }
    double R = 0.0;
    long long esum = 0, gcnt = 0;
    for (int k = threadIdx.x; k < G; k += kBlock) {
        const int64_t p = p0 + k;
        if (k < cnt) {
            const int64_t j = perm ? (int64_t)perm[p] : p;
            const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
            const double m = mass ? mass[j] : 0.0;
            const double dx = x - ox, dy = y - oy, dz = z - oz;
            R = fmax(R, fmax(fabs(dx), fmax(fabs(dy), fabs(dz))));
            loc[p] = make_float4((float)dx, (float)dy, (float)dz, (float)m);
            if (s64) s64[p] = make_double4(x, y, z, m);
            if (k + 1 < cnt) {
                const int64_t j2 = perm ? (int64_t)perm[p + 1] : p + 1;
                const double g = fmax(fabs(pos[3 * j2] - x),
                                      fmax(fabs(pos[3 * j2 + 1] - y), fabs(pos[3 * j2 + 2] - z)));
                esum += exponent_of(g);
                gcnt += 1;
            }
        } else {
            loc[p] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (s64) s64[p] = make_double4(ox, oy, oz, 0.0);
        }
    }
    R = wave_max(R);
    esum = wave_sum(esum);
    gcnt = wave_sum(gcnt);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sR[wave] = R;
        sE[wave] = esum;
        sC[wave] = gcnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) {
            R = fmax(R, sR[w]);
            esum += sE[w];
            gcnt += sC[w];
        }
        // log2 R - mean log2 gap > kFrameLog2, in integer exponents
        const long long eR = exponent_of(R);
        const bool loose = gcnt > 0 && eR * gcnt - esum > (long long)kFrameLog2 * gcnt;
        frame[blockIdx.x] = make_double4(ox, oy, oz, R);
        flag[blockIdx.x] = loose ? 1 : 0;
        if (loose) atomicAdd(nflag, 1);
    }
}

int gather_frames(gh_ctx* c, OrderBufs& o, const double* d_pos, const double* d_mass,
                  const u32* perm, int64_t n, int64_t n_pad, int group, const int* valid,
                  bool want_s64, hipStream_t s)
{
    if (n_pad <= 0) return 0;
    if (n_pad % group != 0) return fail(c, GH_EARG, "internal: group padding");
    const int64_t ng = n_pad / group;
    GH_TRY(ensure(c, o.loc, sizeof(float4) * (size_t)n_pad));
    if (want_s64) GH_TRY(ensure(c, o.s64, sizeof(double4) * (size_t)n_pad));
    GH_TRY(ensure(c, o.frame, sizeof(double4) * (size_t)ng));
    GH_TRY(ensure(c, o.flag, sizeof(int) * (size_t)ng));
    GH_TRY(ensure(c, o.nflag, 256));
    double4* s64 = want_s64 ? o.s64.as<double4>() : nullptr;
#define GH_FRAMES(GG)                                                                       \
    hipLaunchKernelGGL(k_frames<GG>, dim3((unsigned)ng), dim3(kBlock), 0, s, d_pos, d_mass,  \
                       perm, n, valid, o.loc.as<float4>(), s64, o.frame.as<double4>(),       \
                       o.flag.as<int>(), o.nflag.as<int>())
    switch (group) {
        case 256: GH_FRAMES(256); break;
        case 512: GH_FRAMES(512); break;
        case 1024: GH_FRAMES(1024); break;
        default: return fail(c, GH_EARG, "internal: group size %d", group);
    }
#undef GH_FRAMES
    return check_launch(c, "k_frames");
}

}  // namespace gh
