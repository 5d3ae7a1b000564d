// Direct-summation kernels for gfx950.
//
//  k_direct_f64   IEEE FP64, one target per lane, sources in ascending index order,
//                 the operation order of gravhopper/_jbgrav.c:140-193 (self) and
//                 :299-353 (targets), so the result is bit-identical to the reference.
//                 Built with -ffp-contract=off; double '/' and sqrt are the compiler's
//                 correctly rounded expansions.
//  k_direct_f64_chain, k_direct_f64_spread  the same sums for up to 65536 targets:
//                 a target's terms are computed by several lanes, then added in the
//                 same order, so the results are the same bits.
//  k_direct_f32   FP32 pair forces with v_rsq_f32, IPT targets per lane, 256-source
//                 tiles staged in LDS as float4 (x, y, z, m) and read as wave-wide
//                 broadcasts, FP32 sums per tile and an FP32 two-sum (hi, lo) across
//                 tiles; FP64 out.  Sources may be split over grid.y when there are too
//                 few targets to fill 256 CUs; the splits are added in split order.
//  k_direct_sym_f32  the self field with each unordered pair computed once.
//  Both FP32 kernels form coordinates only relative to FP64 origins of groups of
//  spatially sorted particles (gh_frames.hip), so their accuracy does not depend on
//  where the system sits; non-compact groups take double-single instances.
#include <cmath>

#include "gh_internal.h"

namespace gh {

int64_t pad_to(int64_t n, int64_t m) { return (n + m - 1) / m * m; }

// ------------------------------------------------------------------------------
// FP64, reference operation order
// ------------------------------------------------------------------------------
template <bool SELF>
__global__ __launch_bounds__(kBlock) void k_direct_f64(const double4* __restrict__ src,
                                                       int64_t np,
                                                       const double* __restrict__ tgt,
                                                       int64_t nt, double eps2,
                                                       double* __restrict__ acc)
{
    __shared__ double4 tile[kBlock];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < nt;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (live) {
        xi = tgt[3 * i];
        yi = tgt[3 * i + 1];
        zi = tgt[3 * i + 2];
    }
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int64_t base = 0; base < np; base += kBlock) {
        __syncthreads();
        if (base + threadIdx.x < np) tile[threadIdx.x] = src[base + threadIdx.x];
        __syncthreads();
        const int cnt = (int)((np - base) < kBlock ? (np - base) : kBlock);
        if (!live) continue;
        for (int jj = 0; jj < cnt; jj++) {
            const double4 s = tile[jj];
            // dpos[i][j] = x_j - x_i (the reference stores -(x_i - x_j) or (x_j - x_i),
            // which are the same IEEE value); d2 = ((0 + dx^2) + dy^2) + dz^2.
            const double dx = s.x - xi, dy = s.y - yi, dz = s.z - zi;
            double d2 = dx * dx;
            d2 = d2 + dy * dy;
            d2 = d2 + dz * dz;
            const double d = d2 + eps2;
            double inv;
            if (SELF) {
                inv = 1.0 / d / sqrt(d);
            } else {
                inv = (d == 0.0) ? 0.0 : 1.0 / d / sqrt(d);
            }
            const double c0 = s.w * dx * inv;
            const double c1 = s.w * dy * inv;
            const double c2 = s.w * dz * inv;
            if (!SELF || base + jj != i) {
                f0 = f0 + c0;
                f1 = f1 + c1;
                f2 = f2 + c2;
            }
        }
    }
    if (live) {
        acc[3 * i] = f0;
        acc[3 * i + 1] = f1;
        acc[3 * i + 2] = f2;
    }
}

// The same sums for few targets, when one lane per target would leave most of the GPU
// idle (each lane's serial loop over all sources, with an IEEE division and square root
// per pair, is then the whole run time).  TPW targets per wavefront, P = 64 / TPW lanes
// per target: per chunk of P sources, the P lanes of a target compute one term each
// (the same expressions as k_direct_f64), and the lanes then add the P terms in
// ascending source order.  Every target's sum is formed term by term in the reference's
// order, so the results are bit-identical to k_direct_f64's.
template <bool SELF, int TPW>
__global__ __launch_bounds__(kBlock) void k_direct_f64_spread(const double4* __restrict__ src,
                                                              int64_t np,
                                                              const double* __restrict__ tgt,
                                                              int64_t nt, double eps2,
                                                              double* __restrict__ acc)
{
    constexpr int P = 64 / TPW;
    __shared__ double terms[kBlock / 64][3][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / P, q = lane % P;  // target of this lane in the wave, source slot
    const int64_t i = ((int64_t)blockIdx.x * (kBlock / 64) + wave) * TPW + g;
    const bool live = i < nt;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (live) {
        xi = tgt[3 * i];
        yi = tgt[3 * i + 1];
        zi = tgt[3 * i + 2];
    }
    // one term of source j, computed without branches (the source index is clamped, and
    // the term is +0.0 for a source past np or for the target itself); adding +0.0 to a
    // sum that starts at +0.0 never changes it, so these terms leave the reference's
    // sums untouched
    auto term = [&](int64_t j, double& c0, double& c1, double& c2) {
        const double4 s = src[j < np ? j : np - 1];
        const double dx = s.x - xi, dy = s.y - yi, dz = s.z - zi;
        double d2 = dx * dx;
        d2 = d2 + dy * dy;
        d2 = d2 + dz * dz;
        const double d = d2 + eps2;
        double inv;
        if (SELF) {
            inv = 1.0 / d / sqrt(d);
        } else {
            inv = (d == 0.0) ? 0.0 : 1.0 / d / sqrt(d);
        }
        const bool use = j < np && !(SELF && j == i);
        c0 = use ? s.w * dx * inv : 0.0;
        c1 = use ? s.w * dy * inv : 0.0;
        c2 = use ? s.w * dz * inv : 0.0;
    };
    // Software pipelined: while a group adds the P terms of one chunk in source order,
    // its lanes already compute the terms of the next chunk; both are in one branch-free
    // block, so their dependent chains interleave.
    double c0, c1, c2;
    term(q, c0, c1, c2);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;  // the sums (every lane of a group holds them)
    const double* r0 = &terms[wave][0][g * P];
    const double* r1 = &terms[wave][1][g * P];
    const double* r2 = &terms[wave][2][g * P];
    for (int64_t base = 0; base < np; base += P) {
        terms[wave][0][lane] = c0;
        terms[wave][1][lane] = c1;
        terms[wave][2][lane] = c2;
        __syncthreads();
        term(base + P + q, c0, c1, c2);
#pragma unroll
        for (int k = 0; k < P; k++) {
            f0 = f0 + r0[k];
            f1 = f1 + r1[k];
            f2 = f2 + r2[k];
        }
        __syncthreads();
    }
    if (live && q == 0) {
        acc[3 * i] = f0;
        acc[3 * i + 1] = f1;
        acc[3 * i + 2] = f2;
    }
}

// The same sums with a lane of its own for each sum.  A one-wavefront workgroup takes
// TPW targets, whose 3 TPW component sums are held by lanes 0 .. 3 TPW - 1.  Per chunk
// of 64 sources, lane l computes the terms of source base + l for every target (the
// same expressions as k_direct_f64, +0.0 for the target itself and for sources past
// np) into LDS; lane L then adds the 64 terms of its sum in source order.  Each pass
// stores one chunk's terms, computes the next chunk's and adds the stored ones (two
// LDS buffers; a single wavefront reads and writes LDS in program order, so one
// barrier per chunk orders each buffer's writes after the reads two chunks earlier).
template <bool SELF, int TPW>
__global__ __launch_bounds__(64) void k_direct_f64_chain(const double* __restrict__ pos,
                                                         const double* __restrict__ mass,
                                                         int64_t np,
                                                         const double* __restrict__ tgt,
                                                         int64_t nt, double eps2,
                                                         double* __restrict__ acc)
{
    constexpr int NS = 3 * TPW;  // sums per wavefront
    // T[b][k][L]: terms 2k and 2k + 1 of sum L in chunk buffer b
    __shared__ double2 T[2][32][NS];
    const int lane = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * TPW;
    double xt[TPW], yt[TPW], zt[TPW];
#pragma unroll
    for (int t = 0; t < TPW; t++) {
        // a target past nt takes the last target's position; its sums are not stored
        const int64_t i = i0 + t < nt ? i0 + t : nt - 1;
        xt[t] = tgt[3 * i];
        yt[t] = tgt[3 * i + 1];
        zt[t] = tgt[3 * i + 2];
    }
    double c[TPW][3];
    // sources are read in place from the (n,3) positions and the masses: the chunks
    // are small, and a packing pass would cost a launch per force evaluation
    auto terms = [&](int64_t base) {
        const int64_t j = base + lane;
        const int64_t jc = j < np ? j : np - 1;
        const double sx = pos[3 * jc], sy = pos[3 * jc + 1], sz = pos[3 * jc + 2];
        const double sw = mass[jc];
#pragma unroll
        for (int t = 0; t < TPW; t++) {
            const double dx = sx - xt[t], dy = sy - yt[t], dz = sz - zt[t];
            double d2 = dx * dx;
            d2 = d2 + dy * dy;
            d2 = d2 + dz * dz;
            const double d = d2 + eps2;
            double inv;
            if (SELF) {
                inv = 1.0 / d / sqrt(d);
            } else {
                inv = (d == 0.0) ? 0.0 : 1.0 / d / sqrt(d);
            }
            const bool use = j < np && !(SELF && j == i0 + t);
            c[t][0] = use ? sw * dx * inv : 0.0;
            c[t][1] = use ? sw * dy * inv : 0.0;
            c[t][2] = use ? sw * dz * inv : 0.0;
        }
    };
    terms(0);
    const int L = lane < NS ? lane : NS - 1;
    double f = 0.0;
    int b = 0;
    for (int64_t base = 0; base < np; base += 64) {
        double* w = reinterpret_cast<double*>(&T[b][lane >> 1][0]) + (lane & 1);
#pragma unroll
        for (int t = 0; t < TPW; t++) {
            w[2 * (3 * t)] = c[t][0];
            w[2 * (3 * t + 1)] = c[t][1];
            w[2 * (3 * t + 2)] = c[t][2];
        }
        __syncthreads();
        terms(base + 64);
#pragma unroll
        for (int k = 0; k < 32; k++) {
            const double2 v = T[b][k][L];
            f = f + v.x;
            f = f + v.y;
        }
        b ^= 1;
    }
    if (lane < NS && i0 + lane / 3 < nt) acc[3 * i0 + lane] = f;
}

// Kernel shape for the FP64 sums, as targets per wavefront: negative -v selects
// k_direct_f64_chain with v targets, positive v k_direct_f64_spread with v targets
// (64: k_direct_f64, one target per lane).  Automatic: the chain kernel up to
// kF64ChainMax targets (1, 2 or 4 per wavefront by target count), then the spread
// kernel up to kF64SpreadMax (2 to 16 targets, so that at most about kF64Waves
// wavefronts run), then one target per lane.  Measured on Plummer spheres (kernel ms):
//   targets per wavefront    chain 1/2/4/8/16          spread 1/2/4/8/16/32, plain
//   N=2000                   0.034/0.032/0.040/0.062/0.14  0.071/0.064/0.086/0.13/0.25/0.43, 0.41
//   N=16384                  1.31/1.02/0.99/1.52/2.23      3.43/1.95/1.36/1.26/1.99/3.50, 3.42
//   N=65536                  18.6/14.4/13.8/17.0/26.9      48.3/27.8/19.0/15.7/13.6/16.5, 13.8
//   N=262144                 -/-/208/254/393               plain 141
constexpr int64_t kF64Waves = 2048;
constexpr int64_t kF64ChainMax = 32768;
constexpr int64_t kF64SpreadMax = 65536;
static int f64_targets_per_wave(const gh_ctx* c, int64_t nt)
{
    const int t = c->tune[GH_TUNE_DIRECT_F64_TPW];
    if (t == 1 || t == 2 || t == 4 || t == 8 || t == 16 || t == 32 || t == 64) return t;
    if (t == -1 || t == -2 || t == -4 || t == -8 || t == -16) return t;
    if (nt <= 1024) return -1;
    if (nt <= 4096) return -2;
    if (nt <= kF64ChainMax) return -4;
    if (nt > kF64SpreadMax) return 64;
    int tpw = 2;
    while (tpw < 16 && (int64_t)tpw * kF64Waves < nt) tpw *= 2;
    return tpw;
}

// ------------------------------------------------------------------------------
// FP32 pairs, compensated accumulation
// ------------------------------------------------------------------------------
// Every FP32 kernel works in local frames (gh_frames.hip): the particles are sorted
// along a Morton curve, each block of consecutive sorted particles has an FP64 origin,
// and a kernel forms its FP32 coordinates relative to one block's origin from FP64
// values.  Blocks whose flag is set (not compact) are handled by the double-single
// (DS) instances, which carry every coordinate as hi + lo FP32 words of an FP64
// difference and form separations as (hi_j - hi_i) + (lo_j - lo_i).
__device__ __forceinline__ void two_sum_into(float& hi, float& lo, float t)
{
    // Knuth's branch-free two-sum: hi + lo + t == s + (lo + e) exactly before the
    // final rounding of the low word.
    const float s = hi + t;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (t - bp);
    hi = s;
    lo = lo + e;
}

// hi + lo split of an FP64 value: hi = fl32(d), lo = fl32(d - hi) (d - hi is exact).
__device__ __forceinline__ void split_ds(double d, float& hi, float& lo)
{
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// One-sided kernel: IPT targets per lane, 256-source tiles staged in LDS and read as
// wave-wide broadcasts, FP32 sums per tile and an FP32 two-sum (hi, lo) across tiles;
// FP64 out.  Workgroup b holds the sorted targets of group b (256 IPT of them), whose
// frame is tframe[b]; every source tile is re-based into that frame from the FP64
// sources (src64, any order, massless padding).  A workgroup runs only in the
// instance matching its group's flag (DS = 1: double-single separations).  Sources may
// be split over grid.y when there are too few targets to fill 256 CUs; the splits are
// added in split order.  perm (sorted target -> caller's index) places the results;
// with splits the per-split fields go to out in sorted order instead.
template <int IPT, bool GUARD, bool DS>
__global__ __launch_bounds__(kBlock) void k_direct_f32(const double4* __restrict__ src64,
                                                       int64_t np_pad, int64_t j_per_split,
                                                       const float4* __restrict__ tloc,
                                                       const double4* __restrict__ t64,
                                                       const double4* __restrict__ tframe,
                                                       const int* __restrict__ tflag,
                                                       const u32* __restrict__ perm,
                                                       int64_t nt, float eps2,
                                                       double* __restrict__ out)
{
    __shared__ float4 tile[kBlock];
    __shared__ float4 tile_lo[DS ? kBlock : 1];
    if ((tflag[blockIdx.x] != 0) != DS) return;
    const double4 o = tframe[blockIdx.x];
    const int64_t i0 = (int64_t)blockIdx.x * (kBlock * IPT) + threadIdx.x;
    const int64_t j_begin = (int64_t)blockIdx.y * j_per_split;
    const int64_t j_end = min(np_pad, j_begin + j_per_split);

    eps2 = in_vgpr(eps2);
    float xi[IPT], yi[IPT], zi[IPT], xl[IPT], yl[IPT], zl[IPT];
    float hx[IPT], hy[IPT], hz[IPT], lx[IPT], ly[IPT], lz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int64_t i = i0 + (int64_t)k * kBlock;
        xi[k] = yi[k] = zi[k] = xl[k] = yl[k] = zl[k] = 0.f;
        if (i < nt) {
            if (DS) {
                const double4 q = t64[i];
                split_ds(q.x - o.x, xi[k], xl[k]);
                split_ds(q.y - o.y, yi[k], yl[k]);
                split_ds(q.z - o.z, zi[k], zl[k]);
            } else {
                const float4 t = tloc[i];
                xi[k] = t.x;
                yi[k] = t.y;
                zi[k] = t.z;
            }
        }
        hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.f;
    }

    for (int64_t base = j_begin; base < j_end; base += kBlock) {
        __syncthreads();
        {
            const double4 q = src64[base + threadIdx.x];
            if (DS) {
                float4 h, l;
                split_ds(q.x - o.x, h.x, l.x);
                split_ds(q.y - o.y, h.y, l.y);
                split_ds(q.z - o.z, h.z, l.z);
                h.w = (float)q.w;
                l.w = 0.f;
                tile[threadIdx.x] = h;
                tile_lo[threadIdx.x] = l;
            } else {
                tile[threadIdx.x] = make_float4((float)(q.x - o.x), (float)(q.y - o.y),
                                                (float)(q.z - o.z), (float)q.w);
            }
        }
        __syncthreads();
        float ax[IPT], ay[IPT], az[IPT];
#pragma unroll
        for (int k = 0; k < IPT; k++) ax[k] = ay[k] = az[k] = 0.f;
#pragma unroll 8
        for (int jj = 0; jj < kBlock; jj++) {
            const float4 s = tile[jj];
            float4 sl = make_float4(0.f, 0.f, 0.f, 0.f);
            if (DS) sl = tile_lo[jj];
#pragma unroll
            for (int k = 0; k < IPT; k++) {
                float dx = s.x - xi[k];
                float dy = s.y - yi[k];
                float dz = s.z - zi[k];
                if (DS) {
                    dx = dx + (sl.x - xl[k]);
                    dy = dy + (sl.y - yl[k]);
                    dz = dz + (sl.z - zl[k]);
                }
                float r2 = __builtin_fmaf(dx, dx, eps2);
                r2 = __builtin_fmaf(dy, dy, r2);
                r2 = __builtin_fmaf(dz, dz, r2);
                const float ri = __builtin_amdgcn_rsqf(r2);
                const float ri2 = ri * ri;
                float sc = (s.w * ri) * ri2;
                if (GUARD) sc = clamp_finite(sc);
                ax[k] = __builtin_fmaf(dx, sc, ax[k]);
                ay[k] = __builtin_fmaf(dy, sc, ay[k]);
                az[k] = __builtin_fmaf(dz, sc, az[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            two_sum_into(hx[k], lx[k], ax[k]);
            two_sum_into(hy[k], ly[k], ay[k]);
            two_sum_into(hz[k], lz[k], az[k]);
        }
    }

    const bool split = gridDim.y > 1;
    double* o_split = out + (int64_t)blockIdx.y * nt * 3;
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int64_t i = i0 + (int64_t)k * kBlock;
        if (i < nt) {
            double* d = split ? o_split + 3 * i : out + 3 * (int64_t)perm[i];
            d[0] = (double)hx[k] + (double)lx[k];
            d[1] = (double)hy[k] + (double)ly[k];
            d[2] = (double)hz[k] + (double)lz[k];
        }
    }
}

// ------------------------------------------------------------------------------
// FP32 pairs, each unordered pair computed once (self field only)
// ------------------------------------------------------------------------------
// Particles (in sorted order) are cut into blocks of BLK = 256*IPT.  A workgroup takes
// one block pair (I, J), I <= J.  Its 4 wavefronts hold the 256*IPT i-particles of
// block I in registers (IPT per lane); block J is staged in LDS.  For I < J, block J is
// walked in chunks of 64: at step s lane l pairs its i-particles with j = (l + s) mod 64
// of the chunk and accumulates both the force on its i-particles and the force on that
// j in a register that moves to lane l-1 after every step (DPP wavefront rotation), so
// after 64 steps lane l holds the whole wavefront's force on j = l.  The four
// wavefronts' j-partials are added in LDS.  A diagonal pair I = J is done one-sided.
// The workgroup works in block I's frame: i-particles are block I's local coordinates,
// block J's enter as fl(loc_J + fl(o_J - o_I)).  Pairs with a flagged block are left to
// the DS instance, which stages both sides as hi + lo words of the FP64 differences
// from o_I.
//
// Results are partial fields part[K][p] (FP32, 3 per particle): particle p of block B
// receives one partial from every block K (K = B for its diagonal); k_reduce_sym adds
// them in FP64, in an order fixed by their count.  The j-side partials can go to a
// second array (SymOut): a pass over the block pairs of a band of columns J keeps only
// the band's slots (see direct_sym).
constexpr int kRotFromNext = 0x134;  // DPP wave_rol:1 -- lane l reads lane l+1 (mod 64)

__device__ __forceinline__ float rot_from_next(float v)
{
    // every lane has a source lane, so no "old" value is needed
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kRotFromNext, 0xf, 0xf,
                                                   false));
}

// Where the partial fields go: the i-side of an item to part[slot_i][p] (stride np_pad
// particles), the j-side to jpart[slot_j][p - joff] (stride jstride particles).  One
// array for both (jpart = part, jstride = np_pad, joff = 0) outside banded passes.
struct SymOut {
    float* part;
    float* jpart;
    int64_t jstride;
    int64_t joff;
};

// One block pair.  Work item {I, J, slot_i, slot_j}: global block indices I <= J (of
// one simulation), the partial slot the i-side field goes to (J's index within its
// simulation) and the slot of the j-side field (I's index within its simulation).
template <int IPT, bool GUARD, bool DS>
__device__ __forceinline__ void sym_item(const float4* __restrict__ src,
                                         const double4* __restrict__ s64,
                                         const double4* __restrict__ frame, int64_t np_pad,
                                         int4 it, float eps2, const SymOut out,
                                         float4* tile, float4* tile_lo, float (*jp)[3][64])
{
    constexpr int BLK = kBlock * IPT;
    constexpr int NCH = BLK / 64;
    const int I = it.x, J = it.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ibase = (int64_t)I * BLK + wave * 64 * IPT + lane;
    const int64_t jbase = (int64_t)J * BLK;
    eps2 = in_vgpr(eps2);
    const double4 oI = frame[I];

    float xi[IPT], yi[IPT], zi[IPT], mi[IPT], ax[IPT], ay[IPT], az[IPT];
    float xl[IPT], yl[IPT], zl[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        if (DS) {
            const double4 q = s64[ibase + k * 64];
            split_ds(q.x - oI.x, xi[k], xl[k]);
            split_ds(q.y - oI.y, yi[k], yl[k]);
            split_ds(q.z - oI.z, zi[k], zl[k]);
            mi[k] = (float)q.w;
        } else {
            const float4 t = src[ibase + k * 64];
            xi[k] = t.x;
            yi[k] = t.y;
            zi[k] = t.z;
            mi[k] = t.w;
            xl[k] = yl[k] = zl[k] = 0.f;
        }
        ax[k] = ay[k] = az[k] = 0.f;
    }
    if (DS) {
        for (int q = threadIdx.x; q < NCH * 128; q += kBlock) {
            const double4 v = s64[jbase + (q >> 7) * 64 + (q & 63)];
            float4 h, l;
            split_ds(v.x - oI.x, h.x, l.x);
            split_ds(v.y - oI.y, h.y, l.y);
            split_ds(v.z - oI.z, h.z, l.z);
            h.w = (float)v.w;
            l.w = 0.f;
            tile[q] = h;
            tile_lo[q] = l;
        }
    } else {
        const double4 oJ = frame[J];
        const float offx = (float)(oJ.x - oI.x), offy = (float)(oJ.y - oI.y),
                    offz = (float)(oJ.z - oI.z);
        for (int q = threadIdx.x; q < NCH * 128; q += kBlock) {
            float4 v = src[jbase + (q >> 7) * 64 + (q & 63)];
            v.x += offx;
            v.y += offy;
            v.z += offz;
            tile[q] = v;
        }
    }
    __syncthreads();

    if (I == J) {
        for (int c = 0; c < NCH; c++) {
#pragma unroll 8
            for (int q = 0; q < 64; q++) {
                const float4 s = tile[c * 128 + q];
                float4 sl = make_float4(0.f, 0.f, 0.f, 0.f);
                if (DS) sl = tile_lo[c * 128 + q];
#pragma unroll
                for (int k = 0; k < IPT; k++) {
                    float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
                    if (DS) {
                        dx = dx + (sl.x - xl[k]);
                        dy = dy + (sl.y - yl[k]);
                        dz = dz + (sl.z - zl[k]);
                    }
                    float r2 = __builtin_fmaf(dx, dx, eps2);
                    r2 = __builtin_fmaf(dy, dy, r2);
                    r2 = __builtin_fmaf(dz, dz, r2);
                    const float ri = __builtin_amdgcn_rsqf(r2);
                    float sc = (s.w * ri) * (ri * ri);
                    if (GUARD) sc = clamp_finite(sc);
                    ax[k] = __builtin_fmaf(dx, sc, ax[k]);
                    ay[k] = __builtin_fmaf(dy, sc, ay[k]);
                    az[k] = __builtin_fmaf(dz, sc, az[k]);
                }
            }
        }
    } else {
        for (int c = 0; c < NCH; c++) {
            const float4* row = &tile[c * 128 + lane];
            const float4* row_lo = &tile_lo[DS ? c * 128 + lane : 0];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            for (int s0 = 0; s0 < 64; s0 += 8) {
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const float4 sj = row[s0 + s];
                    float4 sl = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (DS) sl = row_lo[s0 + s];
#pragma unroll
                    for (int k = 0; k < IPT; k++) {
                        float dx = sj.x - xi[k], dy = sj.y - yi[k], dz = sj.z - zi[k];
                        if (DS) {
                            dx = dx + (sl.x - xl[k]);
                            dy = dy + (sl.y - yl[k]);
                            dz = dz + (sl.z - zl[k]);
                        }
                        float r2 = __builtin_fmaf(dx, dx, eps2);
                        r2 = __builtin_fmaf(dy, dy, r2);
                        r2 = __builtin_fmaf(dz, dz, r2);
                        const float ri = __builtin_amdgcn_rsqf(r2);
                        const float ri3 = ri * ri * ri;
                        float si = sj.w * ri3;
                        if (GUARD) si = clamp_finite(si);
                        ax[k] = __builtin_fmaf(dx, si, ax[k]);
                        ay[k] = __builtin_fmaf(dy, si, ay[k]);
                        az[k] = __builtin_fmaf(dz, si, az[k]);
                        float sm = mi[k] * ri3;
                        if (GUARD) sm = clamp_finite(sm);
                        fx = __builtin_fmaf(dx, sm, fx);
                        fy = __builtin_fmaf(dy, sm, fy);
                        fz = __builtin_fmaf(dz, sm, fz);
                    }
                    fx = rot_from_next(fx);
                    fy = rot_from_next(fy);
                    fz = rot_from_next(fz);
                }
            }
            // lane l now holds this wavefront's pull of its i-particles by j = l; the
            // force on j is the opposite
            jp[wave][0][lane] = fx;
            jp[wave][1][lane] = fy;
            jp[wave][2][lane] = fz;
            __syncthreads();
            if (threadIdx.x < 3 * 64) {
                const int comp = threadIdx.x >> 6, t = threadIdx.x & 63;
                float sacc = jp[0][comp][t];
#pragma unroll
                for (int w = 1; w < kBlock / 64; w++) sacc += jp[w][comp][t];
                out.jpart[((int64_t)it.w * out.jstride + (jbase - out.joff) + c * 64 + t) * 3 +
                          comp] = -sacc;
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        float* o = out.part + ((int64_t)it.z * np_pad + ibase + k * 64) * 3;
        o[0] = ax[k];
        o[1] = ay[k];
        o[2] = az[k];
    }
}

// One launch over a list of block pairs; items with a flagged block are skipped (the DS
// kernel takes them).  queue == nullptr: workgroup b takes item b.  Otherwise the
// workgroups are persistent and take items from *queue (which must be 0 at launch)
// until the list is exhausted.
template <int IPT, bool GUARD>
__global__ __launch_bounds__(kBlock) void k_direct_sym_f32(const float4* __restrict__ src,
                                                           const double4* __restrict__ frame,
                                                           const int* __restrict__ flag,
                                                           int64_t np_pad,
                                                           const int4* __restrict__ work,
                                                           int64_t items,
                                                           const float* __restrict__ eps2_blk,
                                                           const SymOut out,
                                                           int* __restrict__ queue)
{
    constexpr int NCH = kBlock * IPT / 64;
    // each 64-particle chunk of block J is stored twice in a row, so that lane l's
    // rotating index l + s needs no wrap-around
    __shared__ float4 tile[NCH * 128];
    __shared__ float jp[kBlock / 64][3][64];
    __shared__ int next;
    if (!queue) {
        const int4 it = work[blockIdx.x];
        if (flag[it.x] | flag[it.y]) return;
        sym_item<IPT, GUARD, false>(src, nullptr, frame, np_pad, it, eps2_blk[it.x], out, tile,
                                    nullptr, jp);
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int64_t k = next;
        __syncthreads();
        if (k >= items) break;
        const int4 it = work[k];
        if (flag[it.x] | flag[it.y]) continue;
        sym_item<IPT, GUARD, false>(src, nullptr, frame, np_pad, it, eps2_blk[it.x], out, tile,
                                    nullptr, jp);
        __syncthreads();
    }
}

// The items the fast kernel skips, in double-single form: a persistent grid strides
// over the whole list and takes the items with a flagged block.
template <int IPT, bool GUARD>
__global__ __launch_bounds__(kBlock) void k_direct_sym_ds(const double4* __restrict__ s64,
                                                          const double4* __restrict__ frame,
                                                          const int* __restrict__ flag,
                                                          const int* __restrict__ nflag,
                                                          int64_t np_pad,
                                                          const int4* __restrict__ work,
                                                          int64_t items,
                                                          const float* __restrict__ eps2_blk,
                                                          const SymOut out)
{
    constexpr int NCH = kBlock * IPT / 64;
    __shared__ float4 tile[NCH * 128];
    __shared__ float4 tile_lo[NCH * 128];
    __shared__ float jp[kBlock / 64][3][64];
    if (*nflag == 0) return;  // no flagged block: nothing to do
    for (int64_t k = blockIdx.x; k < items; k += gridDim.x) {
        const int4 it = work[k];
        if (!(flag[it.x] | flag[it.y])) continue;
        sym_item<IPT, GUARD, true>(nullptr, s64, frame, np_pad, it, eps2_blk[it.x], out, tile,
                                   tile_lo, jp);
        __syncthreads();
    }
}

// acc[perm[p]] = sum over the partial slots of p's simulation in FP64 (perm == nullptr:
// acc[p]).  nb_blk[b] (nullable): number of slots of the simulation that block b belongs
// to; otherwise every particle has nb slots.  Four consecutive components per lane, read
// as one 16-byte load per partial slab (each slab, 3 np_pad floats with np_pad a multiple
// of the block size, starts on a 16-byte boundary, and a group of four never straddles
// two blocks, as 3 x 256 is a multiple of 4).
__global__ __launch_bounds__(kBlock) void k_reduce_sym(const float* __restrict__ part, int nb,
                                                       const int* __restrict__ nb_blk,
                                                       int blk, int64_t np_pad, int64_t n3,
                                                       const u32* __restrict__ perm,
                                                       double* __restrict__ acc)
{
    // Each workgroup takes 64 groups of 4 floats, one per lane; its wavefront w adds the
    // slots w, w + kSlotLanes, ... in order (so each load reads 64 consecutive float4),
    // and the wavefronts' sums are added in a fixed order.  The order depends on the slot
    // count only, so an ensemble's particles sum exactly as they would alone.
    static_assert(kBlock == 64 * kSlotLanes, "one slice per wavefront");
    __shared__ double sh[kSlotLanes][4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t k0 = ((int64_t)blockIdx.x * 64 + lane) * 4;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (k0 < n3) {
        const int slots = nb_blk ? nb_blk[(k0 / 3) / blk] : nb;
        const float4* p = (const float4*)(part + k0);
        const int64_t stride = np_pad * 3 / 4;  // float4 per slab
        for (int b = w; b < slots; b += kSlotLanes) {
            const float4 v = p[(int64_t)b * stride];
            s[0] += (double)v.x;
            s[1] += (double)v.y;
            s[2] += (double)v.z;
            s[3] += (double)v.w;
        }
    }
    for (int j = 0; j < 4; j++) sh[w][j][lane] = s[j];
    __syncthreads();
    // wavefront w stores component w of its lanes' groups: (s0 + s1) + (s2 + s3)
    const int64_t k = k0 + w;
    if (k0 < n3 && k < n3) {
        const double t = (sh[0][w][lane] + sh[1][w][lane]) + (sh[2][w][lane] + sh[3][w][lane]);
        const int64_t q = k / 3;
        acc[perm ? 3 * (int64_t)perm[q] + (k - 3 * q) : k] = t;
    }
}

// One pass of the banded pair-once self field (direct_sym): the block pairs (I, J) with
// J in the band [c0, c0 + w) of blocks.  Particle p of block P has this pass's partials
// from partners q (blocks) in partner order: the j-side of (q, P) for q < P when P is
// in the band (partj[q][p - c0 blk], jw blocks of particles per slot), then the i-side
// of (P, q) for q >= P in the band (parti[q - c0][p]).  Their FP64 sum is added to
// acc[perm[p]] (written on the first pass).  w: the band's width in blocks.
__global__ __launch_bounds__(kBlock) void k_reduce_band(const float* __restrict__ parti,
                                                        const float* __restrict__ partj,
                                                        int blk, int c0, int w, int jw,
                                                        int64_t np_pad,
                                                        int64_t n3, const u32* __restrict__ perm,
                                                        bool first, double* __restrict__ acc)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n3) return;
    const int64_t p = k / 3;
    const int comp = (int)(k - 3 * p);
    const int P = (int)(p / blk);
    double sum = 0.0;
    const bool in_band = P >= c0 && P < c0 + w;
    if (in_band) {
        const int64_t pj = p - (int64_t)c0 * blk;
        for (int q = 0; q < P; q++) sum += (double)partj[((int64_t)q * jw * blk + pj) * 3 + comp];
    }
    for (int q = P > c0 ? P : c0; q < c0 + w; q++)
        sum += (double)parti[((int64_t)(q - c0) * np_pad + p) * 3 + comp];
    const int64_t o = 3 * (int64_t)perm[p] + comp;
    acc[o] = first ? sum : acc[o] + sum;
}

// Sum of per-split partial fields (sorted target order), in split order, written to
// acc[perm[i]].
__global__ __launch_bounds__(kBlock) void k_reduce_splits(const double* __restrict__ part,
                                                          int splits, int64_t n3,
                                                          const u32* __restrict__ perm,
                                                          double* __restrict__ acc)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n3) return;
    double s = part[k];
    for (int p = 1; p < splits; p++) s += part[(int64_t)p * n3 + k];
    const int64_t q = k / 3;
    acc[3 * (int64_t)perm[q] + (k - 3 * q)] = s;
}

// (x, y, z, m) in input order, entries n..n_pad-1 massless at the origin
__global__ __launch_bounds__(kBlock) void k_pack_src64(const double* __restrict__ pos,
                                                       const double* __restrict__ mass,
                                                       int64_t n, int64_t n_pad,
                                                       double4* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        out[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    else if (i < n_pad)
        out[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// ------------------------------------------------------------------------------
// dispatch
// ------------------------------------------------------------------------------
namespace {

struct F32Shape {
    int ipt;
    int splits;
};

// Targets per lane and source splits: enough workgroups for every CU, no more
// source splitting than needed.
F32Shape pick_shape(gh_ctx* c, int64_t nt, int64_t np_pad)
{
    int dev_cus = 256;
    (void)c;
    F32Shape sh{1, 1};
    if (nt >= (int64_t)kBlock * 2 * 2 * dev_cus) sh.ipt = 2;
    const int64_t wgs = (nt + (int64_t)kBlock * sh.ipt - 1) / ((int64_t)kBlock * sh.ipt);
    const int64_t want = 4LL * dev_cus;
    if (wgs < want) {
        int64_t sp = (want + wgs - 1) / wgs;
        const int64_t max_sp = np_pad / kBlock;  // at least one tile per split
        if (sp > max_sp) sp = max_sp;
        if (sp < 1) sp = 1;
        sh.splits = (int)sp;
    }
    return sh;
}

template <int IPT, bool DS>
void launch_f32(bool guard, dim3 grid, hipStream_t s, const double4* src64, int64_t np_pad,
                int64_t jps, const OrderBufs& t, int64_t nt, float eps2, double* out)
{
    const float4* tloc = t.loc.as<float4>();
    const double4* t64 = t.s64.as<double4>();
    const double4* fr = t.frame.as<double4>();
    const int* fl = t.flag.as<int>();
    const u32* perm = t.perm();
    if (guard)
        hipLaunchKernelGGL((k_direct_f32<IPT, true, DS>), grid, dim3(kBlock), 0, s, src64,
                           np_pad, jps, tloc, t64, fr, fl, perm, nt, eps2, out);
    else
        hipLaunchKernelGGL((k_direct_f32<IPT, false, DS>), grid, dim3(kBlock), 0, s, src64,
                           np_pad, jps, tloc, t64, fr, fl, perm, nt, eps2, out);
}

}  // namespace

// FP32 pair arithmetic needs the zero-separation guard whenever a coincident pair (or
// the self pair of the one-sided kernels) could produce 0 * inf: m / eps^3 overflows FP32
// for eps^2 below about 1e-15 once masses reach 1e15 Msun, and the guard is also what
// the reference does for targets at zero distance (:327-330).
bool fp32_guard(double eps)
{
    const float eps2 = (float)(eps * eps);
    return !(eps2 >= 1e-15f);
}

// Symmetric (pair-once) kernel for the self field: particles per lane, or 0 for the
// one-sided kernel.  Block pairs must number a few thousand to fill 256 CUs.
// The pair-once kernel keeps one FP32 partial field per block: 12 B x np x (np / block)
// of scratch, which grows as np^2 (12.9 GB at np = 2^20).  Above kSymScratchMax the
// block pairs are taken in bands of columns, one pass each, whose partials take at most
// that much (direct_sym).
constexpr double kSymScratchMax = 13.0 * (1 << 30);

int sym_ipt(gh_ctx* c, int64_t np)
{
    const int forced = c->tune[GH_TUNE_DIRECT_SELF];
    if (forced < 0) return 0;
    int ipt = 0;
    if (forced == 1 || forced == 2 || forced == 4)
        ipt = forced;
    else if (np >= 65536)
        ipt = 4;
    else if (np >= 32768)
        ipt = 2;
    else if (np >= 16384)
        ipt = 1;
    return ipt;
}

template <int IPT>
static void launch_sym_t(bool guard, unsigned grid, hipStream_t s, const float4* src,
                         const double4* s64, const double4* frame, const int* flag,
                         const int* nflag, int64_t np_pad, const int4* work, int64_t items,
                         const float* eps2_blk, const SymOut out, int* queue, unsigned ds_grid)
{
    if (guard) {
        hipLaunchKernelGGL((k_direct_sym_f32<IPT, true>), dim3(grid), dim3(kBlock), 0, s, src,
                           frame, flag, np_pad, work, items, eps2_blk, out, queue);
        hipLaunchKernelGGL((k_direct_sym_ds<IPT, true>), dim3(ds_grid), dim3(kBlock), 0, s, s64,
                           frame, flag, nflag, np_pad, work, items, eps2_blk, out);
    } else {
        hipLaunchKernelGGL((k_direct_sym_f32<IPT, false>), dim3(grid), dim3(kBlock), 0, s, src,
                           frame, flag, np_pad, work, items, eps2_blk, out, queue);
        hipLaunchKernelGGL((k_direct_sym_ds<IPT, false>), dim3(ds_grid), dim3(kBlock), 0, s,
                           s64, frame, flag, nflag, np_pad, work, items, eps2_blk, out);
    }
}

int launch_sym(gh_ctx* c, int ipt, bool guard, const float4* src, const double4* s64,
               const double4* frame, const int* flag, const int* nflag, int64_t np_pad,
               const int4* work, int64_t items, const float* eps2_blk, float* part, int* queue,
               unsigned persistent_grid, hipStream_t s, float* jpart, int64_t jstride,
               int64_t joff)
{
    SymOut out;
    out.part = part;
    out.jpart = jpart ? jpart : part;
    out.jstride = jpart ? jstride : np_pad;
    out.joff = jpart ? joff : 0;
    if (items > 0x7fffffff) return fail(c, GH_EARG, "too many block pairs for one launch");
    const unsigned grid = queue ? persistent_grid : (unsigned)items;
    if (grid == 0) return 0;
    const unsigned ds_grid = (unsigned)std::min<int64_t>(items, 2 * (int64_t)c->n_cu);
    switch (ipt) {
        case 1: launch_sym_t<1>(guard, grid, s, src, s64, frame, flag, nflag, np_pad, work, items, eps2_blk, out, queue, ds_grid); break;
        case 2: launch_sym_t<2>(guard, grid, s, src, s64, frame, flag, nflag, np_pad, work, items, eps2_blk, out, queue, ds_grid); break;
        case 4: launch_sym_t<4>(guard, grid, s, src, s64, frame, flag, nflag, np_pad, work, items, eps2_blk, out, queue, ds_grid); break;
        default: return fail(c, GH_EARG, "unsupported particles-per-lane %d", ipt);
    }
    return check_launch(c, "k_direct_sym_f32");
}

int reduce_sym(gh_ctx* c, const float* part, int nb, const int* nb_blk, int blk,
               int64_t np_pad, int64_t n, const u32* perm, double* acc, hipStream_t s)
{
    if (n <= 0) return 0;
    if ((np_pad * 3) % 4 != 0 || (3 * blk) % 4 != 0)
        return fail(c, GH_EARG, "partial slabs are not 16-byte aligned");
    hipLaunchKernelGGL(k_reduce_sym, dim3(blocks_for((3 * n + 3) / 4, 64)), dim3(kBlock), 0, s,
                       part, nb, nb_blk, blk, np_pad, 3 * n, perm, acc);
    return check_launch(c, "k_reduce_sym");
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

int direct_sym(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np, int ipt,
               double eps, double* d_acc, hipStream_t s)
{
    const int64_t blk = (int64_t)kBlock * ipt;
    const int64_t np_pad = pad_to(np, blk);
    const int nb = (int)(np_pad / blk);
    // band width in blocks: all of them when one partial per block fits kSymScratchMax,
    // else as many as keep the band's partials (2 w np_pad of 12 bytes) within it
    int w = nb;
    if (c->tune[GH_TUNE_SYM_BAND] > 0)
        w = std::min(nb, c->tune[GH_TUNE_SYM_BAND]);
    else if (12.0 * (double)np_pad * nb > kSymScratchMax)
        w = (int)std::max<double>(1.0, std::floor(kSymScratchMax / (24.0 * (double)np_pad)));
    const int bands = (nb + w - 1) / w;
    const int64_t items = (int64_t)nb * (nb + 1) / 2;
    if (items > 0x7fffffff) return fail(c, GH_EARG, "too many particles for one launch");
    const size_t part_floats = bands == 1 ? 3 * (size_t)np_pad * nb : 6 * (size_t)np_pad * w;
    GH_TRY(ensure(c, c->scratch, sizeof(float) * part_floats));
    GH_TRY(ensure(c, c->eps_blk, sizeof(float) * (size_t)nb));
    const float eps2 = (float)(eps * eps);
    if (c->sym_nb != nb || c->sym_ipt != ipt || c->sym_w != w) {
        // block pairs (I, J), I <= J, band by band of columns J, then ordered by I;
        // slots: the partner block, within the band for the i-side
        std::vector<int4> wl((size_t)items);
        c->sym_band_start.assign(1, 0);
        size_t k = 0;
        for (int b = 0; b < bands; b++) {
            const int c0 = b * w, c1 = std::min(nb, c0 + w);
            for (int I = 0; I < c1; I++)
                for (int J = std::max(I, c0); J < c1; J++)
                    wl[k++] = make_int4(I, J, bands == 1 ? J : J - c0, I);
            c->sym_band_start.push_back((int64_t)k);
        }
        GH_TRY(ensure(c, c->work, sizeof(int4) * (size_t)items));
        GH_HIP(c, hipMemcpyAsync(c->work.p, wl.data(), sizeof(int4) * (size_t)items,
                                 hipMemcpyHostToDevice, s));
        GH_HIP(c, hipStreamSynchronize(s));
        c->sym_nb = nb;
        c->sym_ipt = ipt;
        c->sym_w = w;
        c->sym_eps2 = -1.f;
    }
    if (c->sym_eps2 != eps2) {  // the block list is rebuilt above whenever nb changes
        hipLaunchKernelGGL(k_fill_f32, dim3(blocks_for(nb, kBlock)), dim3(kBlock), 0, s,
                           c->eps_blk.as<float>(), (int64_t)nb, eps2);
        GH_TRY(check_launch(c, "k_fill_f32"));
        c->sym_eps2 = eps2;
    }
    OrderBufs& o = c->ord_p;
    GH_TRY(spatial_order(c, o, d_pos, np, s));
    GH_TRY(gather_frames(c, o, d_pos, d_mass, o.perm(), np, np_pad, (int)blk, nullptr, true, s));
    float* part = c->scratch.as<float>();
    const bool guard = fp32_guard(eps);
    if (bands == 1) {
        {
            TimedScope ts(c, s, GH_T_DIRECT_F32);
            GH_TRY(launch_sym(c, ipt, guard, o.loc.as<float4>(), o.s64.as<double4>(),
                              o.frame.as<double4>(), o.flag.as<int>(), o.nflag.as<int>(), np_pad,
                              c->work.as<int4>(), items, c->eps_blk.as<float>(), part, nullptr,
                              0, s));
        }
        return reduce_sym(c, part, nb, nullptr, (int)blk, np_pad, np, o.perm(), d_acc, s);
    }
    float* partj = part + 3 * (size_t)np_pad * w;
    for (int b = 0; b < bands; b++) {
        const int c0 = b * w, wb = std::min(nb, c0 + w) - c0;
        const int64_t i0 = c->sym_band_start[b], ni = c->sym_band_start[b + 1] - i0;
        {
            TimedScope ts(c, s, GH_T_DIRECT_F32);
            GH_TRY(launch_sym(c, ipt, guard, o.loc.as<float4>(), o.s64.as<double4>(),
                              o.frame.as<double4>(), o.flag.as<int>(), o.nflag.as<int>(), np_pad,
                              c->work.as<int4>() + i0, ni, c->eps_blk.as<float>(), part, nullptr,
                              0, s, partj, (int64_t)w * blk, (int64_t)c0 * blk));
        }
        hipLaunchKernelGGL(k_reduce_band, dim3(blocks_for(3 * np, kBlock)), dim3(kBlock), 0, s,
                           part, partj, (int)blk, c0, wb, w, np_pad, 3 * np, o.perm(), b == 0,
                           d_acc);
        GH_TRY(check_launch(c, "k_reduce_band"));
    }
    return 0;
}

// One-sided FP32 field at nt targets.  src64: the sources as (x, y, z, m), np_pad of
// them (massless padding); t: the targets' spatial order and frames in groups of
// 256 x the chosen targets per lane.
static int direct_f32_onesided(gh_ctx* c, const double4* src64, int64_t np_pad, OrderBufs& t,
                               int64_t nt, const F32Shape& sh, double eps, double* d_acc,
                               hipStream_t s)
{
    if (nt <= 0) return 0;
    if (np_pad % kBlock != 0) return fail(c, GH_EARG, "internal: source count not padded");
    const float eps2 = (float)(eps * eps);
    const bool guard = fp32_guard(eps);
    int64_t jps = pad_to((np_pad + sh.splits - 1) / sh.splits, kBlock);
    const int splits = (int)((np_pad + jps - 1) / jps);
    double* out = d_acc;
    if (splits > 1) {
        GH_TRY(ensure(c, c->scratch, sizeof(double) * 3 * (size_t)nt * splits));
        out = c->scratch.as<double>();
    }
    dim3 grid(blocks_for(nt, kBlock * sh.ipt), splits);
    {
        TimedScope ts(c, s, GH_T_DIRECT_F32);
        switch (sh.ipt) {
            case 1:
                launch_f32<1, false>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                launch_f32<1, true>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                break;
            case 2:
                launch_f32<2, false>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                launch_f32<2, true>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                break;
            case 4:
                launch_f32<4, false>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                launch_f32<4, true>(guard, grid, s, src64, np_pad, jps, t, nt, eps2, out);
                break;
            default: return fail(c, GH_EARG, "unsupported targets-per-lane %d", sh.ipt);
        }
        GH_TRY(check_launch(c, "k_direct_f32"));
    }
    if (splits > 1) {
        hipLaunchKernelGGL(k_reduce_splits, dim3(blocks_for(3 * nt, kBlock)), dim3(kBlock), 0,
                           s, out, splits, 3 * nt, t.perm(), d_acc);
        GH_TRY(check_launch(c, "k_reduce_splits"));
    }
    return 0;
}

int direct_dev(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np,
               const double* d_fpos, int64_t nf, double eps, int precision, double* d_acc,
               hipStream_t s)
{
    const bool self = (d_fpos == nullptr);
    const double* tgt = self ? d_pos : d_fpos;
    const int64_t nt = self ? np : nf;
    if (np < 0 || nt < 0) return fail(c, GH_EARG, "negative particle count");
    if (nt == 0) return 0;
    if (np == 0) {
        return zero_dev(c, d_acc, 3 * nt, s);
    }
    if (precision == GH_FP64) {
        const double eps2 = eps * eps;
        const int tpw = f64_targets_per_wave(c, nt);
        if (tpw < 0) {
            TimedScope ts(c, s, GH_T_DIRECT_F64);
            const int tc = -tpw;
            const dim3 gc((unsigned)((nt + tc - 1) / tc));
#define GH_CHAIN(T)                                                                       \
    do {                                                                                  \
        if (self)                                                                         \
            hipLaunchKernelGGL((k_direct_f64_chain<true, T>), gc, dim3(64), 0, s,         \
                               d_pos, d_mass, np, tgt, nt, eps2, d_acc);                  \
        else                                                                              \
            hipLaunchKernelGGL((k_direct_f64_chain<false, T>), gc, dim3(64), 0, s,        \
                               d_pos, d_mass, np, tgt, nt, eps2, d_acc);                  \
    } while (0)
            switch (tc) {
            case 1: GH_CHAIN(1); break;
            case 2: GH_CHAIN(2); break;
            case 4: GH_CHAIN(4); break;
            case 8: GH_CHAIN(8); break;
            default: GH_CHAIN(16); break;
            }
#undef GH_CHAIN
            return check_launch(c, "k_direct_f64_chain");
        }
        GH_TRY(ensure(c, c->src64, sizeof(double4) * (size_t)np));
        double4* src = c->src64.as<double4>();
        hipLaunchKernelGGL(k_pack_src64, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, s,
                           d_pos, d_mass, np, np, src);
        GH_TRY(check_launch(c, "k_pack_src64"));
        TimedScope ts(c, s, GH_T_DIRECT_F64);
        if (tpw == 64) {
            if (self)
                hipLaunchKernelGGL(k_direct_f64<true>, dim3(blocks_for(nt, kBlock)),
                                   dim3(kBlock), 0, s, src, np, tgt, nt, eps2, d_acc);
            else
                hipLaunchKernelGGL(k_direct_f64<false>, dim3(blocks_for(nt, kBlock)),
                                   dim3(kBlock), 0, s, src, np, tgt, nt, eps2, d_acc);
            return check_launch(c, "k_direct_f64");
        }
        const dim3 grid(blocks_for((nt + tpw - 1) / tpw, kBlock / 64));
#define GH_SPREAD(T)                                                                      \
    do {                                                                                  \
        if (self)                                                                         \
            hipLaunchKernelGGL((k_direct_f64_spread<true, T>), grid, dim3(kBlock), 0, s,  \
                               src, np, tgt, nt, eps2, d_acc);                            \
        else                                                                              \
            hipLaunchKernelGGL((k_direct_f64_spread<false, T>), grid, dim3(kBlock), 0, s, \
                               src, np, tgt, nt, eps2, d_acc);                            \
    } while (0)
        switch (tpw) {
        case 1: GH_SPREAD(1); break;
        case 2: GH_SPREAD(2); break;
        case 4: GH_SPREAD(4); break;
        case 8: GH_SPREAD(8); break;
        case 16: GH_SPREAD(16); break;
        default: GH_SPREAD(32); break;
        }
#undef GH_SPREAD
        return check_launch(c, "k_direct_f64_spread");
    }
    if (precision != GH_FP32) return fail(c, GH_EARG, "unknown precision %d", precision);
    const int sym = self ? sym_ipt(c, np) : 0;
    if (sym > 0) return direct_sym(c, d_pos, d_mass, np, sym, eps, d_acc, s);
    F32Shape sh = pick_shape(c, nt, pad_to(np, kBlock));
    const int forced = c->tune[GH_TUNE_DIRECT_IPT];
    if (forced == 1 || forced == 2 || forced == 4) sh.ipt = forced;
    const int group = kBlock * sh.ipt;
    const int64_t nt_pad = pad_to(nt, group);
    if (self) {
        // targets and sources are the particles: one order, the sorted FP64 particles
        // (massless padding at their group's origin) are the sources
        OrderBufs& o = c->ord_p;
        GH_TRY(spatial_order(c, o, d_pos, np, s));
        GH_TRY(gather_frames(c, o, d_pos, d_mass, o.perm(), np, nt_pad, group, nullptr, true, s));
        return direct_f32_onesided(c, o.s64.as<double4>(), nt_pad, o, nt, sh, eps, d_acc, s);
    }
    const int64_t np_pad = pad_to(np, kBlock);
    GH_TRY(ensure(c, c->src64, sizeof(double4) * (size_t)np_pad));
    hipLaunchKernelGGL(k_pack_src64, dim3(blocks_for(np_pad, kBlock)), dim3(kBlock), 0, s, d_pos,
                       d_mass, np, np_pad, c->src64.as<double4>());
    GH_TRY(check_launch(c, "k_pack_src64"));
    OrderBufs& t = c->ord_t;
    GH_TRY(spatial_order(c, t, tgt, nt, s));
    GH_TRY(gather_frames(c, t, tgt, nullptr, t.perm(), nt, nt_pad, group, nullptr, true, s));
    return direct_f32_onesided(c, c->src64.as<double4>(), np_pad, t, nt, sh, eps, d_acc, s);
}

}  // namespace gh
