// Leapfrog (drift-kick-drift) update kernels and FP64 diagnostics.
//
// The state of a device-resident run is kept as (n,3) float64 arrays: the half-step
// positions xh (where the force is evaluated), the velocities v and the full-step
// positions x.  One step of gravhopper/gravhopper.py:405-416 is
//   force(xh) -> acc;   k_kick_drift: v' = v + (acc*A + 0)*dt,
//                                     inc = ((0.5*v')*dt)*D,  x' = xh + inc,
//                                     xh' = x' + inc  (next step's half drift)
// so every step is one force evaluation plus one fused elementwise pass, which also
// writes the snapshot slots of the history when asked to.
#include "gh_internal.h"

namespace gh {

__global__ __launch_bounds__(kBlock) void k_half_drift(const double* __restrict__ x,
                                                       const double* __restrict__ v,
                                                       int64_t n3, double dt,
                                                       double* __restrict__ xh)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n3) xh[k] = x[k] + 0.5 * v[k] * dt * kDriftScale;
}

// hist_p / hist_v (nullable): snapshot slot of this step, or, when slot != nullptr,
// the base of the history, the slot index being read from *slot (graph replay).
// slot_next (nullable, another word than slot): receives *slot + 1 for the next step,
// so a replayed step needs no separate counter update.
// ext (nullable): external acceleration in km/s/Myr, added as the reference adds the
// summed external forces to the N-body acceleration (gravhopper.py:419-459); without
// external forces the sum is with 0.0, as in the reference.
__global__ __launch_bounds__(kBlock) void k_kick_drift(const double* __restrict__ acc,
                                                       int64_t n3, double dt,
                                                       double* __restrict__ xh,
                                                       double* __restrict__ v,
                                                       double* __restrict__ x,
                                                       double* __restrict__ hist_p,
                                                       double* __restrict__ hist_v,
                                                       const long long* __restrict__ slot,
                                                       const double* __restrict__ ext,
                                                       long long* __restrict__ slot_next)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n3) return;
    if (slot_next && k == 0) *slot_next = *slot + 1;
    if (slot) {
        const int64_t off = (int64_t)(*slot) * n3;
        if (hist_p) hist_p += off;
        if (hist_v) hist_v += off;
    }
    const double a = acc[k] * kAccelScale + (ext ? ext[k] : 0.0);
    const double vn = v[k] + a * dt;
    const double inc = 0.5 * vn * dt * kDriftScale;
    const double xn = xh[k] + inc;
    v[k] = vn;
    x[k] = xn;
    xh[k] = xn + inc;
    if (hist_p) hist_p[k] = xn;
    if (hist_v) hist_v[k] = vn;
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(double* __restrict__ p, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// Zero fill by kernel rather than hipMemsetAsync, so that captured step graphs hold
// only kernel nodes.
int zero_dev(gh_ctx* c, double* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, p, n);
    return check_launch(c, "k_fill_zero");
}

int half_drift(gh_ctx* c, const double* x, const double* v, int64_t n, double dt, double* xh,
               hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_half_drift, dim3(blocks_for(3 * n, kBlock)), dim3(kBlock), 0, s, x,
                       v, 3 * n, dt, xh);
    return check_launch(c, "k_half_drift");
}

int kick_drift(gh_ctx* c, const double* acc, int64_t n, double dt, double* xh, double* v,
               double* x, double* hp, double* hv, hipStream_t s, const long long* slot,
               const double* ext, long long* slot_next)
{
    if (n <= 0) return 0;
    if (slot_next && !slot) return fail(c, GH_EARG, "slot_next needs slot");
    hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(3 * n, kBlock)), dim3(kBlock), 0, s, acc,
                       3 * n, dt, xh, v, x, hp, hv, slot, ext, slot_next);
    return check_launch(c, "k_kick_drift");
}

__global__ void k_set_slot(long long* __restrict__ slot, long long value, int add)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *slot = add ? *slot + value : value;
}

// *slot = value (add == 0) or *slot += value (add != 0), by a one-thread kernel so
// that it can sit in a captured graph.
int set_slot(gh_ctx* c, long long* slot, long long value, int add, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(64), 0, s, slot, value, add);
    return check_launch(c, "k_set_slot");
}

// ------------------------------------------------------------------------------
// diagnostics: per-block partial sums, then one block adds them in block order, so
// the result does not depend on scheduling.
// ------------------------------------------------------------------------------
constexpr int kDiag = 5;  // kinetic, potential, px, py, pz

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// PACKED: particles as (x, y, z, m) records (src); otherwise read in place from the
// (n,3) positions and the masses.  The arithmetic is the same either way.
template <bool PACKED>
__global__ __launch_bounds__(kBlock) void k_energy_partials(const double4* __restrict__ src,
                                                            const double* __restrict__ pos,
                                                            const double* __restrict__ mass,
                                                            const double* __restrict__ vel,
                                                            int64_t n, double eps2,
                                                            double* __restrict__ partial)
{
    __shared__ double4 tile[kBlock];
    __shared__ double red[kBlock];
    auto load = [&](int64_t j) -> double4 {
        if (PACKED) return src[j];
        return make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j]);
    };
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < n;
    double4 me = make_double4(0, 0, 0, 0);
    if (live) me = load(i);
    double row = 0.0;
    for (int64_t base = 0; base < n; base += kBlock) {
        __syncthreads();
        if (base + threadIdx.x < n) tile[threadIdx.x] = load(base + threadIdx.x);
        __syncthreads();
        const int cnt = (int)((n - base) < kBlock ? (n - base) : kBlock);
        if (!live) continue;
        for (int jj = 0; jj < cnt; jj++) {
            if (base + jj == i) continue;
            const double4 s = tile[jj];
            const double dx = s.x - me.x, dy = s.y - me.y, dz = s.z - me.z;
            const double d = dx * dx + dy * dy + dz * dz + eps2;
            row += s.w / sqrt(d);
        }
    }
    double ke = 0.0, pe = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    if (live) {
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        ke = 0.5 * me.w * (vx * vx + vy * vy + vz * vz);
        pe = -0.5 * me.w * row;
        p0 = me.w * vx;
        p1 = me.w * vy;
        p2 = me.w * vz;
    }
    double r[kDiag] = {ke, pe, p0, p1, p2};
    for (int q = 0; q < kDiag; q++) {
        const double t = block_sum(r[q], red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * kDiag + q] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_energy_final(const double* __restrict__ partial,
                                                         int64_t nblocks,
                                                         double* __restrict__ out)
{
    if (threadIdx.x < kDiag) {
        double s = 0.0;
        for (int64_t b = 0; b < nblocks; b++) s += partial[b * kDiag + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_pack_src64_e(const double* __restrict__ pos,
                                                         const double* __restrict__ mass,
                                                         int64_t n, double4* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
}

int energy_dev(gh_ctx* c, const double* d_pos, const double* d_vel, const double* d_mass,
               int64_t n, double eps, double* d_out5, hipStream_t s)
{
    const int64_t nb = blocks_for(n, kBlock);
    GH_TRY(ensure(c, c->src64, sizeof(double4) * (size_t)(n > 0 ? n : 1)));
    GH_TRY(ensure(c, c->counters, sizeof(double) * kDiag * (size_t)(nb + 1)));
    double4* src = c->src64.as<double4>();
    double* part = c->counters.as<double>() + kDiag;
    if (n > 0) {
        hipLaunchKernelGGL(k_pack_src64_e, dim3(nb), dim3(kBlock), 0, s, d_pos, d_mass, n, src);
        GH_TRY(check_launch(c, "k_pack_src64_e"));
        hipLaunchKernelGGL(k_energy_partials<true>, dim3(nb), dim3(kBlock), 0, s, src,
                           (const double*)nullptr, (const double*)nullptr, d_vel, n, eps * eps,
                           part);
        GH_TRY(check_launch(c, "k_energy_partials"));
    }
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, s, part, nb, d_out5);
    return check_launch(c, "k_energy_final");
}

}  // namespace gh

// ------------------------------------------------------------------------------
// allocation-free diagnostics (gh_energy_dev_noalloc)
// ------------------------------------------------------------------------------
// The per-block partial sums live in this file-scope device array instead of an arena:
// kNoAllocBlocks blocks of 256 particles, i.e. up to 4194304 particles, 640 KiB held on
// every device the library runs on until the process ends.  Host code never passes the
// array's name to a kernel: each device's address of it comes from hipGetSymbolAddress,
// once per device, and is kept in g_scratch[device].
namespace gh {
namespace {
constexpr int64_t kNoAllocBlocks = 16384;
constexpr int kNoAllocDevices = 64;
__device__ __align__(256) double scratch[kNoAllocBlocks * kDiag];
double* g_scratch[kNoAllocDevices] = {nullptr};
}  // namespace

}  // namespace gh

using namespace gh;

extern "C" int gh_energy_dev_noalloc(const double* d_pos, const double* d_vel,
                                     const double* d_mass, int64_t np, double eps,
                                     double* d_out5, void* stream)
{
    if (np < 0 || np > kNoAllocBlocks * kBlock || !d_out5) return GH_EARG;
    if (np > 0 && (!d_pos || !d_vel || !d_mass)) return GH_EARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return GH_EHIP;
    double* part = __atomic_load_n(&g_scratch[dev], __ATOMIC_ACQUIRE);
    if (!part) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GH_EHIP;
        part = static_cast<double*>(p);
        __atomic_store_n(&g_scratch[dev], part, __ATOMIC_RELEASE);
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nb = blocks_for(np, kBlock);
    (void)hipGetLastError();  // a launch error below is this call's, not an earlier one
    if (np > 0) {
        // every partial slot read below (blocks 0 .. nb-1) is written first, so the
        // previous call's contents never show
        hipLaunchKernelGGL(k_energy_partials<false>, dim3((unsigned)nb), dim3(kBlock), 0, s,
                           (const double4*)nullptr, d_pos, d_mass, d_vel, np, eps * eps, part);
        if (hipGetLastError() != hipSuccess) return GH_EHIP;
    }
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, s, part, nb, d_out5);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}
