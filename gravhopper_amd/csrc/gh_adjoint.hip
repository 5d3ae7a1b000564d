// Adjoint (reverse-mode) of the drift-kick-drift leapfrog with the direct-sum force.
//
// Forward step (code units; c = 0.5 dt D, D = km/s*Myr in kpc, A = G Msun/kpc^2 in
// km/s/Myr):   xh = x + c v;   v' = v + dt A F(xh, m);   x' = xh + c v'
// with F_k = sum_{j != k} m_j K(x_j - x_k), K(d) = d s^-3, s^2 = |d|^2 + eps^2.
//
// Reverse step, given lx = dL/dx', lv = dL/dv':
//   mv  = lv + c lx                 (total gradient of v')
//   mu  = dt A mv                   (gradient of F)
//   lxh = lx + J_x^T mu,  lm += J_m^T mu
//   lx <- lxh,  lv <- mv + c lxh
// with, for d = x_j - x_k and the softened tidal tensor T(d) = I s^-3 - 3 d d^T s^-5,
//   (J_x^T mu)_k = sum_{j != k} T(d) (m_k mu_j - m_j mu_k)
//   (J_m^T mu)_k = sum_{j != k} mu_j . K(x_k - x_j) = -sum_j s^-3 (mu_j . d)
//
// The pair terms are antisymmetric (T(-d) = T(d), w_kj = -w_jk), so the VJP kernel is
// a pair-once kernel like the force's: with w = m_k mu_j - m_j mu_k and
// u = -3 s^-2 (d . w), T w = s^-3 (w + u d), the i side accumulates
//   A += s^-3 mu_j,  B += s^-3 m_j,  C += s^-3 u d,  L += s^-3 (mu_j . d)
// (g_k = m_k A - B mu_k + C, lm_k = -L), and the j side, moving one lane per step,
//   A' += s^-3 m_k,  B' += s^-3 mu_k,  C' += s^-3 u d,  L' += s^-3 (mu_k . d)
// (g_j = -mu_j A' + m_j B' - C', lm_j = L').  Pair arithmetic in FP32, partial sums
// added in FP64.  The trajectory is re-run from checkpoints taken every k steps; the
// re-run keeps each step's half-step positions and, for FP32 forces, the spatial order
// its force evaluation made, which the VJP of the same positions reuses.
#include <algorithm>

#include "gh_internal.h"

namespace gh {


constexpr int kRotNext = 0x134;  // DPP wave_rol:1 -- lane l reads lane l+1 (mod 64)

__device__ __forceinline__ float rotn(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kRotNext, 0xf, 0xf, false));
}

// The particles are in the sorted order of gh_frames.hip and the workgroup works in
// block I's frame, as the force kernel does: src holds local FP32 coordinates and
// masses, block J's coordinates are shifted by fl(o_J - o_I).  The DS instance takes the
// items with a flagged block, with both sides as hi + lo words of FP64 differences
// from o_I (src64), and separations (hi_j - hi_k) + (lo_j - lo_k).
// j-side records in LDS: (x, y, z, m), (mu_x, mu_y, mu_z, 0) and, DS, (lo x, y, z, 0).
__device__ __forceinline__ void split_pair(double d, float& hi, float& lo)
{
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

template <int IPT, bool GUARD, bool DS>
__global__ __launch_bounds__(kBlock) void k_vjp_sym_f32(const float4* __restrict__ src,
                                                        const double4* __restrict__ src64,
                                                        const double4* __restrict__ frame,
                                                        const int* __restrict__ flag,
                                                        const float4* __restrict__ mu,
                                                        int64_t np_pad,
                                                        const int4* __restrict__ work,
                                                        float eps2, float4* __restrict__ part)
{
    constexpr int BLK = kBlock * IPT;
    constexpr int NCH = BLK / 64;
    __shared__ float4 tp[NCH * 128];
    __shared__ float4 tm[NCH * 128];
    __shared__ float4 tl[DS ? NCH * 128 : 1];
    __shared__ float jp[kBlock / 64][8][64];
    const int4 it = work[blockIdx.x];
    const int I = it.x, J = it.y;
    if ((flag[I] | flag[J]) != (DS ? 1 : 0)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ibase = (int64_t)I * BLK + wave * 64 * IPT + lane;
    const int64_t jbase = (int64_t)J * BLK;
    eps2 = in_vgpr(eps2);
    const double4 oI = frame[I];

    float xk[IPT], yk[IPT], zk[IPT], mk[IPT], ux[IPT], uy[IPT], uz[IPT];
    float xl[IPT], yl[IPT], zl[IPT];
    float Ax[IPT], Ay[IPT], Az[IPT], B[IPT], Cx[IPT], Cy[IPT], Cz[IPT], L[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        if (DS) {
            const double4 t = src64[ibase + k * 64];
            split_pair(t.x - oI.x, xk[k], xl[k]);
            split_pair(t.y - oI.y, yk[k], yl[k]);
            split_pair(t.z - oI.z, zk[k], zl[k]);
            mk[k] = (float)t.w;
        } else {
            const float4 t = src[ibase + k * 64];
            xk[k] = t.x;
            yk[k] = t.y;
            zk[k] = t.z;
            mk[k] = t.w;
            xl[k] = yl[k] = zl[k] = 0.f;
        }
        const float4 q = mu[ibase + k * 64];
        ux[k] = q.x;
        uy[k] = q.y;
        uz[k] = q.z;
        Ax[k] = Ay[k] = Az[k] = B[k] = Cx[k] = Cy[k] = Cz[k] = L[k] = 0.f;
    }
    {
        const double4 oJ = frame[J];
        const float offx = (float)(oJ.x - oI.x), offy = (float)(oJ.y - oI.y),
                    offz = (float)(oJ.z - oI.z);
        for (int q = threadIdx.x; q < NCH * 128; q += kBlock) {
            const int64_t g = jbase + (q >> 7) * 64 + (q & 63);
            if (DS) {
                const double4 v = src64[g];
                float4 h, l;
                split_pair(v.x - oI.x, h.x, l.x);
                split_pair(v.y - oI.y, h.y, l.y);
                split_pair(v.z - oI.z, h.z, l.z);
                h.w = (float)v.w;
                l.w = 0.f;
                tp[q] = h;
                tl[q] = l;
            } else {
                float4 v = src[g];
                v.x += offx;
                v.y += offy;
                v.z += offz;
                tp[q] = v;
            }
            tm[q] = mu[g];
        }
    }
    __syncthreads();

    if (I == J) {
        // one-sided over the block, self pair excluded
        for (int c = 0; c < NCH; c++) {
#pragma unroll 4
            for (int q = 0; q < 64; q++) {
                const float4 s = tp[c * 128 + q];
                const float4 m4 = tm[c * 128 + q];
                float4 sl = make_float4(0.f, 0.f, 0.f, 0.f);
                if (DS) sl = tl[c * 128 + q];
                const int jl = c * 64 + q;
#pragma unroll
                for (int k = 0; k < IPT; k++) {
                    const int kl = wave * 64 * IPT + k * 64 + lane;
                    float dx = s.x - xk[k], dy = s.y - yk[k], dz = s.z - zk[k];
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
                    float ri3 = ri2 * ri;
                    if (GUARD) ri3 = (r2 > 0.f) ? ri3 : 0.f;
                    ri3 = (jl == kl) ? 0.f : ri3;
                    const float pj = __builtin_fmaf(dx, m4.x, __builtin_fmaf(dy, m4.y, dz * m4.z));
                    const float pk = __builtin_fmaf(dx, ux[k], __builtin_fmaf(dy, uy[k], dz * uz[k]));
                    const float dw = __builtin_fmaf(mk[k], pj, -s.w * pk);
                    const float cu = ri3 * ((-3.f * ri2) * dw);
                    Ax[k] = __builtin_fmaf(ri3, m4.x, Ax[k]);
                    Ay[k] = __builtin_fmaf(ri3, m4.y, Ay[k]);
                    Az[k] = __builtin_fmaf(ri3, m4.z, Az[k]);
                    B[k] = __builtin_fmaf(ri3, s.w, B[k]);
                    Cx[k] = __builtin_fmaf(cu, dx, Cx[k]);
                    Cy[k] = __builtin_fmaf(cu, dy, Cy[k]);
                    Cz[k] = __builtin_fmaf(cu, dz, Cz[k]);
                    L[k] = __builtin_fmaf(ri3, pj, L[k]);
                }
            }
        }
    } else {
        for (int c = 0; c < NCH; c++) {
            const float4* rp = &tp[c * 128 + lane];
            const float4* rm = &tm[c * 128 + lane];
            const float4* rl = &tl[DS ? c * 128 + lane : 0];
            float a = 0.f, bx = 0.f, by = 0.f, bz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, l = 0.f;
            for (int s0 = 0; s0 < 64; s0 += 4) {
#pragma unroll
                for (int st = 0; st < 4; st++) {
                    const float4 s = rp[s0 + st];
                    const float4 m4 = rm[s0 + st];
                    float4 sl = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (DS) sl = rl[s0 + st];
#pragma unroll
                    for (int k = 0; k < IPT; k++) {
                        float dx = s.x - xk[k], dy = s.y - yk[k], dz = s.z - zk[k];
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
                        float ri3 = ri2 * ri;
                        if (GUARD) ri3 = (r2 > 0.f) ? ri3 : 0.f;
                        const float pj =
                            __builtin_fmaf(dx, m4.x, __builtin_fmaf(dy, m4.y, dz * m4.z));
                        const float pk =
                            __builtin_fmaf(dx, ux[k], __builtin_fmaf(dy, uy[k], dz * uz[k]));
                        const float dw = __builtin_fmaf(mk[k], pj, -s.w * pk);
                        const float cu = ri3 * ((-3.f * ri2) * dw);
                        Ax[k] = __builtin_fmaf(ri3, m4.x, Ax[k]);
                        Ay[k] = __builtin_fmaf(ri3, m4.y, Ay[k]);
                        Az[k] = __builtin_fmaf(ri3, m4.z, Az[k]);
                        B[k] = __builtin_fmaf(ri3, s.w, B[k]);
                        Cx[k] = __builtin_fmaf(cu, dx, Cx[k]);
                        Cy[k] = __builtin_fmaf(cu, dy, Cy[k]);
                        Cz[k] = __builtin_fmaf(cu, dz, Cz[k]);
                        L[k] = __builtin_fmaf(ri3, pj, L[k]);
                        a = __builtin_fmaf(ri3, mk[k], a);
                        bx = __builtin_fmaf(ri3, ux[k], bx);
                        by = __builtin_fmaf(ri3, uy[k], by);
                        bz = __builtin_fmaf(ri3, uz[k], bz);
                        cx = __builtin_fmaf(cu, dx, cx);
                        cy = __builtin_fmaf(cu, dy, cy);
                        cz = __builtin_fmaf(cu, dz, cz);
                        l = __builtin_fmaf(ri3, pk, l);
                    }
                    a = rotn(a);
                    bx = rotn(bx);
                    by = rotn(by);
                    bz = rotn(bz);
                    cx = rotn(cx);
                    cy = rotn(cy);
                    cz = rotn(cz);
                    l = rotn(l);
                }
            }
            // lane l holds this wavefront's sums for j = chunk particle l
            jp[wave][0][lane] = a;
            jp[wave][1][lane] = bx;
            jp[wave][2][lane] = by;
            jp[wave][3][lane] = bz;
            jp[wave][4][lane] = cx;
            jp[wave][5][lane] = cy;
            jp[wave][6][lane] = cz;
            jp[wave][7][lane] = l;
            __syncthreads();
            if (threadIdx.x < 64) {
                const int t = threadIdx.x;
                float r[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    float acc = jp[0][q][t];
#pragma unroll
                    for (int w = 1; w < kBlock / 64; w++) acc += jp[w][q][t];
                    r[q] = acc;
                }
                const float4 sj = tp[c * 128 + t];
                const float4 mj = tm[c * 128 + t];
                // g_j = -mu_j A' + m_j B' - C',  lm_j = L'
                const float4 o = make_float4(
                    __builtin_fmaf(sj.w, r[1], -mj.x * r[0]) - r[4],
                    __builtin_fmaf(sj.w, r[2], -mj.y * r[0]) - r[5],
                    __builtin_fmaf(sj.w, r[3], -mj.z * r[0]) - r[6], r[7]);
                part[(int64_t)it.w * np_pad + jbase + c * 64 + t] = o;
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        // g_k = m_k A - B mu_k + C,  lm_k = -L
        const float4 o = make_float4(__builtin_fmaf(mk[k], Ax[k], -B[k] * ux[k]) + Cx[k],
                                     __builtin_fmaf(mk[k], Ay[k], -B[k] * uy[k]) + Cy[k],
                                     __builtin_fmaf(mk[k], Az[k], -B[k] * uz[k]) + Cz[k], -L[k]);
        part[(int64_t)it.z * np_pad + ibase + k * 64] = o;
    }
}

// mv = lv + c lx; mu = dt A mv (FP32 copy for the pair kernel, in sorted order: mu[p]
// belongs to particle perm[p]); keeps mv
__global__ __launch_bounds__(kBlock) void k_adj_pre(const double* __restrict__ lx,
                                                    const double* __restrict__ lv, int64_t n,
                                                    int64_t n_pad, const u32* __restrict__ perm,
                                                    double cdrift, double dtA,
                                                    double* __restrict__ mv,
                                                    float4* __restrict__ mu)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_pad) return;
    if (p >= n) {
        mu[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const int64_t i = perm[p];
    double t[3];
    for (int k = 0; k < 3; k++) {
        t[k] = lv[3 * i + k] + cdrift * lx[3 * i + k];
        mv[3 * i + k] = t[k];
    }
    mu[p] = make_float4((float)(dtA * t[0]), (float)(dtA * t[1]), (float)(dtA * t[2]), 0.f);
}

// lxh = lx + g; lx = lxh; lv = mv + c lxh; lm += gm   (g, gm summed over slots in FP64;
// slot entries are in sorted order, sorted position p being particle perm[p])
__global__ __launch_bounds__(kBlock) void k_adj_post(const float4* __restrict__ part, int nb,
                                                     int64_t n_pad, int64_t n,
                                                     const u32* __restrict__ perm,
                                                     const double* __restrict__ mv,
                                                     double cdrift, double* __restrict__ lx,
                                                     double* __restrict__ lv,
                                                     double* __restrict__ lm)
{
    // 64 particles per workgroup, one per lane; wavefront w adds the slots w, w +
    // kSlotLanes, ... in order, and wavefront 0 adds the wavefronts' sums in a fixed order
    // and applies the update
    static_assert(kBlock == 64 * kSlotLanes, "one slice per wavefront");
    __shared__ double sh[kSlotLanes][4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t p = (int64_t)blockIdx.x * 64 + lane;
    double g4[4] = {0.0, 0.0, 0.0, 0.0};
    if (p < n)
        for (int b = w; b < nb; b += kSlotLanes) {
            const float4 q = part[(int64_t)b * n_pad + p];
            g4[0] += (double)q.x;
            g4[1] += (double)q.y;
            g4[2] += (double)q.z;
            g4[3] += (double)q.w;
        }
    for (int j = 0; j < 4; j++) sh[w][j][lane] = g4[j];
    __syncthreads();
    if (w != 0 || p >= n) return;
    double gs[4];
    for (int j = 0; j < 4; j++)
        gs[j] = (sh[0][j][lane] + sh[1][j][lane]) + (sh[2][j][lane] + sh[3][j][lane]);
    const double g0 = gs[0], g1 = gs[1], g2 = gs[2], gm = gs[3];
    const int64_t i = perm[p];
    const double g[3] = {g0, g1, g2};
    for (int k = 0; k < 3; k++) {
        const double lxh = lx[3 * i + k] + g[k];
        lx[3 * i + k] = lxh;
        lv[3 * i + k] = mv[3 * i + k] + cdrift * lxh;
    }
    lm[i] += gm;
}

__global__ __launch_bounds__(kBlock) void k_copy_d(const double* __restrict__ a,
                                                   double* __restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

static int copy_d(gh_ctx* c, const double* a, double* b, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_copy_d, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, a, b, n);
    return check_launch(c, "k_copy_d");
}

__global__ __launch_bounds__(kBlock) void k_copy_u32(const u32* __restrict__ a,
                                                     u32* __restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Puts back a spatial order that the forward replay saved for the same positions and
// clears the count of flagged groups as k_morton does: the order buffers are then as
// spatial_order leaves them, without sorting again.
__global__ __launch_bounds__(kBlock) void k_restore_order(const u32* __restrict__ saved,
                                                          u32* __restrict__ perm, int64_t n,
                                                          int* __restrict__ nflag)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) *nflag = 0;
    if (i < n) perm[i] = saved[i];
}

static int vjp_ipt(int64_t n)
{
    if (n >= 65536) return 4;
    if (n >= 32768) return 2;
    return 1;
}

template <int IPT, bool GUARD>
static void launch_vjp_g(unsigned items, hipStream_t s, const OrderBufs& o, const float4* mu,
                         int64_t np_pad, const int4* work, float eps2, float4* part)
{
    const float4* src = o.loc.as<float4>();
    const double4* s64 = o.s64.as<double4>();
    const double4* fr = o.frame.as<double4>();
    const int* fl = o.flag.as<int>();
    hipLaunchKernelGGL((k_vjp_sym_f32<IPT, GUARD, false>), dim3(items), dim3(kBlock), 0, s, src,
                       s64, fr, fl, mu, np_pad, work, eps2, part);
    hipLaunchKernelGGL((k_vjp_sym_f32<IPT, GUARD, true>), dim3(items), dim3(kBlock), 0, s, src,
                       s64, fr, fl, mu, np_pad, work, eps2, part);
}

template <int IPT>
static void launch_vjp_t(bool guard, unsigned items, hipStream_t s, const OrderBufs& o,
                         const float4* mu, int64_t np_pad, const int4* work, float eps2,
                         float4* part)
{
    if (guard)
        launch_vjp_g<IPT, true>(items, s, o, mu, np_pad, work, eps2, part);
    else
        launch_vjp_g<IPT, false>(items, s, o, mu, np_pad, work, eps2, part);
}

struct AdjWork {
    int ipt = 0, nb = 0;
    int64_t n_pad = 0, items = 0;
};

static int adj_prepare(gh_ctx* c, int64_t n, AdjWork& w, hipStream_t s)
{
    w.ipt = vjp_ipt(n);
    const int64_t blk = (int64_t)kBlock * w.ipt;
    w.n_pad = pad_to(n, blk);
    w.nb = (int)(w.n_pad / blk);
    w.items = (int64_t)w.nb * (w.nb + 1) / 2;
    if (w.items > 0x7fffffff) return fail(c, GH_EARG, "too many particles for the adjoint");
    if (c->adj_nb != w.nb || c->adj_ipt != w.ipt) {
        std::vector<int4> v((size_t)w.items);
        size_t k = 0;
        for (int I = 0; I < w.nb; I++)
            for (int J = I; J < w.nb; J++) v[k++] = make_int4(I, J, J, I);
        GH_TRY(ensure(c, c->adj_work, sizeof(int4) * v.size()));
        GH_HIP(c, hipMemcpyAsync(c->adj_work.p, v.data(), sizeof(int4) * v.size(),
                                 hipMemcpyHostToDevice, s));
        GH_HIP(c, hipStreamSynchronize(s));
        c->adj_nb = w.nb;
        c->adj_ipt = w.ipt;
    }
    return 0;
}

// One reverse step at half-step positions xh: (lx, lv) after the step -> before it;
// lm accumulates.  saved, if not null, is the spatial order of xh that the forward
// replay's force evaluation made.
static int reverse_step(gh_ctx* c, const AdjWork& w, const double* xh, const double* m,
                        int64_t n, double dt, double eps, double* lx, double* lv, double* lm,
                        const u32* saved, hipStream_t s)
{
    const double cdrift = 0.5 * dt * kDriftScale;
    const double dtA = dt * kAccelScale;
    float4* mu = c->adj_mu.as<float4>();
    double* mv = c->adj_mv.as<double>();
    float4* part = c->adj_part.as<float4>();
    OrderBufs& o = c->ord_p;
    if (saved && o.sorted_n == n) {
        hipLaunchKernelGGL(k_restore_order, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s,
                           saved, o.idx[o.sel].as<u32>(), n, o.nflag.as<int>());
        GH_TRY(check_launch(c, "k_restore_order"));
        o.sorted = xh;
    } else {
        GH_TRY(spatial_order(c, o, xh, n, s));
    }
    GH_TRY(gather_frames(c, o, xh, m, o.perm(), n, w.n_pad, kBlock * w.ipt, nullptr, true, s));
    hipLaunchKernelGGL(k_adj_pre, dim3(blocks_for(w.n_pad, kBlock)), dim3(kBlock), 0, s, lx, lv,
                       n, w.n_pad, o.perm(), cdrift, dtA, mv, mu);
    GH_TRY(check_launch(c, "k_adj_pre"));
    const float eps2 = (float)(eps * eps);
    const bool guard = fp32_guard(eps);
    {
        TimedScope ts(c, s, GH_T_VJP);
        switch (w.ipt) {
            case 1: launch_vjp_t<1>(guard, (unsigned)w.items, s, o, mu, w.n_pad, c->adj_work.as<int4>(), eps2, part); break;
            case 2: launch_vjp_t<2>(guard, (unsigned)w.items, s, o, mu, w.n_pad, c->adj_work.as<int4>(), eps2, part); break;
            case 4: launch_vjp_t<4>(guard, (unsigned)w.items, s, o, mu, w.n_pad, c->adj_work.as<int4>(), eps2, part); break;
            default: return fail(c, GH_EARG, "internal: vjp shape");
        }
        GH_TRY(check_launch(c, "k_vjp_sym_f32"));
    }
    hipLaunchKernelGGL(k_adj_post, dim3(blocks_for(n, 64)), dim3(kBlock), 0, s, part, w.nb,
                       w.n_pad, n, o.perm(), mv, cdrift, lx, lv, lm);
    return check_launch(c, "k_adj_post");
}

// forward over [0, len) steps from (x, v); or, when hx is not null, xh of every step saved
// to hx (len x n3), stopping at the last step's xh (x and v are then left part way) and,
// when hperm is not null and the force evaluation ordered the particles (FP32), that
// order to hperm (len x n), with have[st] = 1
static int forward_segment(gh_ctx* c, double* x, double* v, const double* m, int64_t n,
                           double dt, double eps, int precision, int64_t len, double* hx,
                           u32* hperm, std::vector<char>* have, hipStream_t s)
{
    const int64_t n3 = 3 * n;
    GH_TRY(ensure(c, c->xh, sizeof(double) * (size_t)n3));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * (size_t)n3));
    double* xh = c->xh.as<double>();
    double* acc = c->acc64.as<double>();
    if (have) have->assign((size_t)len, 0);
    GH_TRY(half_drift(c, x, v, n, dt, xh, s));
    for (int64_t st = 0; st < len; st++) {
        if (hx) {
            GH_TRY(copy_d(c, xh, hx + st * n3, n3, s));
            // the reverse sweep needs only the last step's half-step positions
            if (st == len - 1) break;
        }
        if (n > 1) {
            OrderBufs& o = c->ord_p;
            o.sorted = nullptr;
            GH_TRY(direct_dev(c, xh, m, n, nullptr, n, eps, precision, acc, s));
            if (hperm && have && o.sorted == xh && o.sorted_n == n) {
                hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s,
                                   o.perm(), hperm + st * n, n);
                GH_TRY(check_launch(c, "k_copy_u32"));
                (*have)[(size_t)st] = 1;
            }
        } else {
            GH_TRY(zero_dev(c, acc, n3, s));
        }
        GH_TRY(kick_drift(c, acc, n, dt, xh, v, x, nullptr, nullptr, s));
    }
    return 0;
}

}  // namespace gh

using namespace gh;

extern "C" {

int gh_adjoint_forward(gh_ctx* c, double* d_x, double* d_v, const double* d_m, int64_t n,
                       double dt, double eps, int precision, int64_t nsteps, int64_t every,
                       double* d_ckpt, void* stream)
{
    if (!c) return GH_EARG;
    if (n < 1 || nsteps < 0 || every < 1) return fail(c, GH_EARG, "bad adjoint sizes");
    if (precision != GH_FP64 && precision != GH_FP32)
        return fail(c, GH_EARG, "unknown precision %d", precision);
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = pick_stream(c, stream);
    StreamOrder order_(c, s);
    const int64_t n3 = 3 * n;
    for (int64_t s0 = 0, seg = 0; s0 < nsteps; s0 += every, seg++) {
        GH_TRY(copy_d(c, d_x, d_ckpt + (2 * seg) * n3, n3, s));
        GH_TRY(copy_d(c, d_v, d_ckpt + (2 * seg + 1) * n3, n3, s));
        const int64_t len = std::min(every, nsteps - s0);
        GH_TRY(forward_segment(c, d_x, d_v, d_m, n, dt, eps, precision, len, nullptr, nullptr,
                               nullptr, s));
    }
    return 0;
}

int gh_adjoint_backward(gh_ctx* c, const double* d_ckpt, const double* d_m, int64_t n,
                        double dt, double eps, int precision, int64_t nsteps, int64_t every,
                        double* d_lx, double* d_lv, double* d_lm, void* stream)
{
    if (!c) return GH_EARG;
    if (n < 1 || nsteps < 0 || every < 1) return fail(c, GH_EARG, "bad adjoint sizes");
    if (precision != GH_FP64 && precision != GH_FP32)
        return fail(c, GH_EARG, "unknown precision %d", precision);
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = pick_stream(c, stream);
    StreamOrder order_(c, s);
    const int64_t n3 = 3 * n;
    GH_TRY(zero_dev(c, d_lm, n, s));
    if (nsteps == 0) return 0;
    AdjWork w;
    GH_TRY(adj_prepare(c, n, w, s));
    GH_TRY(ensure(c, c->adj_mu, sizeof(float4) * (size_t)w.n_pad));
    GH_TRY(ensure(c, c->adj_mv, sizeof(double) * (size_t)n3));
    GH_TRY(ensure(c, c->adj_part, sizeof(float4) * (size_t)w.n_pad * w.nb));
    GH_TRY(ensure(c, c->adj_seg, sizeof(double) * (size_t)n3 * (size_t)(every + 2)));
    double* segx = c->adj_seg.as<double>();            // every x n3 half-step positions
    double* x = segx + every * n3;                      // replay state
    double* v = x + n3;
    u32* segp = nullptr;                                // every x n spatial orders
    if (precision == GH_FP32 && n > 1) {
        GH_TRY(ensure(c, c->adj_perm, sizeof(u32) * (size_t)n * (size_t)every));
        segp = c->adj_perm.as<u32>();
    }
    std::vector<char> have;
    const int64_t nseg = (nsteps + every - 1) / every;
    for (int64_t seg = nseg - 1; seg >= 0; seg--) {
        const int64_t s0 = seg * every;
        const int64_t len = std::min(every, nsteps - s0);
        GH_TRY(copy_d(c, d_ckpt + (2 * seg) * n3, x, n3, s));
        GH_TRY(copy_d(c, d_ckpt + (2 * seg + 1) * n3, v, n3, s));
        GH_TRY(forward_segment(c, x, v, d_m, n, dt, eps, precision, len, segx, segp, &have, s));
        for (int64_t st = len - 1; st >= 0; st--)
            GH_TRY(reverse_step(c, w, segx + st * n3, d_m, n, dt, eps, d_lx, d_lv, d_lm,
                                have[(size_t)st] ? segp + st * n : nullptr, s));
    }
    return 0;
}

}  // extern "C"
