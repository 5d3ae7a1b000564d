// Batched ensembles: many independent direct-sum simulations of uneven N advanced
// together, one force launch per step for all of them.
//
// Layout: the simulations' particles are concatenated, each simulation starting at a
// multiple of the block size BLK = 256*IPT (padding particles are massless and never
// updated).  Every step sorts each simulation along a Morton curve in its own box
// (spatial_order_seg) and gathers the FP32 local frames of the blocks, exactly as a
// single run does.  The work list holds every block pair (I, J), I <= J,
// of every simulation, so one launch of the pair-once kernel covers the whole ensemble;
// all items are BLK x BLK pairs, so the load is balanced whatever the mix of N.  With
// `persistent` the launch is a fixed grid of workgroups that take items from an atomic
// counter.  Partial slots are per simulation (block index within the simulation), so
// a simulation's result does not depend on the others: it is bit-identical to running
// that simulation alone through the pair-once kernel with the same IPT.
//
// A step -- order, frames, zero the queue counter, forces (and their double-single
// instance), slot reduction, fused kick/drift -- can be captured once in a HIP graph
// and replayed.
#include <algorithm>
#include <cstring>

#include "gh_internal.h"

struct gh_ensemble {
    gh_ctx* c = nullptr;
    int nsims = 0, ipt = 4, blk = 1024;
    bool persistent = false;
    int64_t n_pad = 0, nb_total = 0, items = 0;
    int nb_max = 0;
    unsigned grid = 0;
    std::vector<int64_t> start, count;
    gh::DevBuf x, v, xh, m, acc, part, work, eps_blk, nb_blk, valid_blk, dt_blk, queue;
    gh::DevBuf seg_blk, seg_start, seg_count;  // simulation of each block, ranges
    gh::OrderBufs ord;                           // per-step order and frames
    bool guard = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace gh {

__global__ __launch_bounds__(kBlock) void k_ens_half_drift(
    const double* __restrict__ x, const double* __restrict__ v, int64_t n_pad, int blk,
    const int* __restrict__ valid_blk, const double* __restrict__ dt_blk, double* __restrict__ xh)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_pad) return;
    const int64_t b = p / blk;
    if ((int)(p - b * blk) >= valid_blk[b]) return;  // padding: never updated
    const double dt = dt_blk[b];
    for (int k = 0; k < 3; k++) xh[3 * p + k] = x[3 * p + k] + 0.5 * v[3 * p + k] * dt * kDriftScale;
}

// v' = v + (a A + 0) dt; x' = xh + ((0.5 v') dt) D; xh' = x' + ((0.5 v') dt) D
__global__ __launch_bounds__(kBlock) void k_ens_kick_drift(
    const double* __restrict__ acc, int64_t n_pad, int blk, const int* __restrict__ valid_blk,
    const double* __restrict__ dt_blk, double* __restrict__ xh, double* __restrict__ v,
    double* __restrict__ x)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_pad) return;
    const int64_t b = p / blk;
    if ((int)(p - b * blk) >= valid_blk[b]) return;
    const double dt = dt_blk[b];
    for (int k = 0; k < 3; k++) {
        const double a = acc[3 * p + k] * kAccelScale + 0.0;
        const double vn = v[3 * p + k] + a * dt;
        const double inc = 0.5 * vn * dt * kDriftScale;
        const double xn = xh[3 * p + k] + inc;
        v[3 * p + k] = vn;
        x[3 * p + k] = xn;
        xh[3 * p + k] = xn + inc;
    }
}

__global__ void k_zero_int(int* __restrict__ p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

static int ens_step(gh_ensemble* e, hipStream_t s)
{
    gh_ctx* c = e->c;
    OrderBufs& o = e->ord;
    GH_TRY(spatial_order_seg(c, o, e->xh.as<double>(), e->n_pad, e->blk, e->nsims,
                             e->seg_blk.as<int>(), e->seg_start.as<int64_t>(),
                             e->seg_count.as<int64_t>(), s));
    GH_TRY(gather_frames(c, o, e->xh.as<double>(), e->m.as<double>(), o.perm(), e->n_pad,
                         e->n_pad, e->blk, e->valid_blk.as<int>(), true, s));
    if (e->persistent) {
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, e->queue.as<int>());
        GH_TRY(check_launch(c, "k_zero_int"));
    }
    {
        TimedScope ts(c, s, GH_T_DIRECT_F32);
        GH_TRY(launch_sym(c, e->ipt, e->guard, o.loc.as<float4>(), o.s64.as<double4>(),
                          o.frame.as<double4>(), o.flag.as<int>(), o.nflag.as<int>(), e->n_pad,
                          e->work.as<int4>(),
                          e->items, e->eps_blk.as<float>(), e->part.as<float>(),
                          e->persistent ? e->queue.as<int>() : nullptr, e->grid, s));
    }
    GH_TRY(reduce_sym(c, e->part.as<float>(), 0, e->nb_blk.as<int>(), e->blk, e->n_pad,
                      e->n_pad, o.perm(), e->acc.as<double>(), s));
    TimedScope tu(c, s, GH_T_UPDATE);
    hipLaunchKernelGGL(k_ens_kick_drift, dim3(blocks_for(e->n_pad, kBlock)), dim3(kBlock), 0, s,
                       e->acc.as<double>(), e->n_pad, e->blk, e->valid_blk.as<int>(),
                       e->dt_blk.as<double>(), e->xh.as<double>(), e->v.as<double>(),
                       e->x.as<double>());
    return check_launch(c, "k_ens_kick_drift");
}

static void ens_free(gh_ensemble* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    gh_ctx* c = e->c;
    OrderBufs& o = e->ord;
    DevBuf* bufs[] = {&e->x,         &e->v,         &e->xh,       &e->m,         &e->acc,
                      &e->part,      &e->work,      &e->eps_blk,  &e->nb_blk,    &e->valid_blk,
                      &e->dt_blk,    &e->queue,     &e->seg_blk,  &e->seg_start, &e->seg_count,
                      &o.keys[0],    &o.keys[1],    &o.idx[0],    &o.idx[1],     &o.hist,
                      &o.box,        &o.boxpart,    &o.loc,       &o.s64,        &o.frame,
                      &o.flag,       &o.nflag};
    for (DevBuf* b : bufs) release(c, *b);
}

}  // namespace gh

using namespace gh;

extern "C" {

int gh_ensemble_create(gh_ctx* c, int n_sims, const int64_t* counts, const double* pos,
                       const double* vel, const double* mass, const double* eps,
                       const double* dt, int ipt, int persistent, gh_ensemble** out)
{
    if (!c || !out) return GH_EARG;
    *out = nullptr;
    if (n_sims <= 0) return fail(c, GH_EARG, "an ensemble needs at least one simulation");
    if (ipt == 0) ipt = 4;
    if (ipt != 1 && ipt != 2 && ipt != 4)
        return fail(c, GH_EARG, "particles per lane must be 1, 2 or 4 (got %d)", ipt);
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    gh_ensemble* e = new gh_ensemble();
    e->c = c;
    e->nsims = n_sims;
    e->ipt = ipt;
    e->blk = kBlock * ipt;
    e->persistent = persistent != 0;
    int64_t total = 0, pad = 0;
    std::vector<int> nb_blk, valid_blk, seg_blk;
    std::vector<float> eps_blk;
    std::vector<double> dt_blk;
    std::vector<int4> work;
    for (int s = 0; s < n_sims; s++) {
        const int64_t n = counts[s];
        if (n < 1) {
            delete e;
            return fail(c, GH_EARG, "simulation %d has %lld particles", s, (long long)n);
        }
        e->start.push_back(pad);
        e->count.push_back(n);
        const int64_t nb = (n + e->blk - 1) / e->blk;
        const int64_t b0 = pad / e->blk;
        for (int64_t I = 0; I < nb; I++) {
            nb_blk.push_back((int)nb);
            seg_blk.push_back(s);
            valid_blk.push_back((int)std::min<int64_t>(e->blk, n - I * e->blk));
            eps_blk.push_back((float)(eps[s] * eps[s]));
            dt_blk.push_back(dt[s]);
            for (int64_t J = I; J < nb; J++)
                work.push_back(make_int4((int)(b0 + I), (int)(b0 + J), (int)J, (int)I));
        }
        if (fp32_guard(eps[s])) e->guard = true;
        e->nb_max = std::max<int>(e->nb_max, (int)nb);
        pad += nb * e->blk;
        total += n;
    }
    e->n_pad = pad;
    e->nb_total = pad / e->blk;
    e->items = (int64_t)work.size();
    e->grid = (unsigned)std::min<int64_t>(e->items, 256 * 8);
    const size_t np = (size_t)pad;
    int rc = 0;
    auto al = [&](DevBuf& b, size_t bytes) {
        if (!rc) rc = ensure(c, b, bytes);
    };
    al(e->x, 24 * np);
    al(e->v, 24 * np);
    al(e->xh, 24 * np);
    al(e->m, 8 * np);
    al(e->acc, 24 * np);
    al(e->seg_blk, sizeof(int) * seg_blk.size());
    al(e->seg_start, sizeof(int64_t) * (size_t)n_sims);
    al(e->seg_count, sizeof(int64_t) * (size_t)n_sims);
    al(e->part, 12 * np * (size_t)e->nb_max);
    al(e->work, sizeof(int4) * work.size());
    al(e->eps_blk, sizeof(float) * eps_blk.size());
    al(e->nb_blk, sizeof(int) * nb_blk.size());
    al(e->valid_blk, sizeof(int) * valid_blk.size());
    al(e->dt_blk, sizeof(double) * dt_blk.size());
    al(e->queue, 256);
    if (rc) {
        ens_free(e);
        delete e;
        return rc;
    }
    // padded host staging
    std::vector<double> hx(3 * np, 0.0), hv(3 * np, 0.0), hm(np, 0.0);
    int64_t off = 0;
    for (int s = 0; s < n_sims; s++) {
        const int64_t st = e->start[s], n = e->count[s];
        std::memcpy(&hx[3 * st], pos + 3 * off, sizeof(double) * 3 * n);
        std::memcpy(&hv[3 * st], vel + 3 * off, sizeof(double) * 3 * n);
        std::memcpy(&hm[st], mass + off, sizeof(double) * n);
        off += n;
    }
    hipStream_t s = c->stream;
    auto up = [&](DevBuf& b, const void* h, size_t bytes) {
        if (!rc && bytes)
            rc = hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s) == hipSuccess
                     ? 0
                     : fail(c, GH_EHIP, "ensemble upload failed");
    };
    up(e->x, hx.data(), 24 * np);
    up(e->v, hv.data(), 24 * np);
    up(e->xh, hx.data(), 24 * np);
    up(e->m, hm.data(), 8 * np);
    up(e->work, work.data(), sizeof(int4) * work.size());
    up(e->eps_blk, eps_blk.data(), sizeof(float) * eps_blk.size());
    up(e->nb_blk, nb_blk.data(), sizeof(int) * nb_blk.size());
    up(e->valid_blk, valid_blk.data(), sizeof(int) * valid_blk.size());
    up(e->dt_blk, dt_blk.data(), sizeof(double) * dt_blk.size());
    up(e->seg_blk, seg_blk.data(), sizeof(int) * seg_blk.size());
    up(e->seg_start, e->start.data(), sizeof(int64_t) * (size_t)n_sims);
    up(e->seg_count, e->count.data(), sizeof(int64_t) * (size_t)n_sims);
    // size the order and frame buffers now, so that a captured step allocates nothing
    if (!rc)
        rc = spatial_order_seg(c, e->ord, e->xh.as<double>(), e->n_pad, e->blk, e->nsims,
                               e->seg_blk.as<int>(), e->seg_start.as<int64_t>(),
                               e->seg_count.as<int64_t>(), s);
    if (!rc)
        rc = gather_frames(c, e->ord, e->xh.as<double>(), e->m.as<double>(), e->ord.perm(),
                           e->n_pad, e->n_pad, e->blk, e->valid_blk.as<int>(), true, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(c, GH_EHIP, "ensemble upload");
    if (rc) {
        ens_free(e);
        delete e;
        return rc;
    }
    *out = e;
    return 0;
}

int gh_ensemble_run(gh_ensemble* e, int64_t nsteps, int use_graph)
{
    if (!e) return GH_EARG;
    gh_ctx* c = e->c;
    if (nsteps < 0) return fail(c, GH_EARG, "negative step count");
    if (nsteps == 0) return 0;
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    hipLaunchKernelGGL(k_ens_half_drift, dim3(blocks_for(e->n_pad, kBlock)), dim3(kBlock), 0, s,
                       e->x.as<double>(), e->v.as<double>(), e->n_pad, e->blk,
                       e->valid_blk.as<int>(), e->dt_blk.as<double>(), e->xh.as<double>());
    GH_TRY(check_launch(c, "k_ens_half_drift"));
    if (!use_graph || c->timing) {
        for (int64_t k = 0; k < nsteps; k++) GH_TRY(ens_step(e, s));
        return 0;
    }
    if (!e->exec) {
        GH_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = ens_step(e, s);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ee != hipSuccess) return fail(c, GH_EHIP, "graph capture failed: %s", hipGetErrorString(ee));
        e->graph = g;
        GH_HIP(c, hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
    }
    for (int64_t k = 0; k < nsteps; k++) GH_HIP(c, hipGraphLaunch(e->exec, s));
    return 0;
}

int gh_ensemble_get(gh_ensemble* e, double* pos, double* vel)
{
    if (!e) return GH_EARG;
    gh_ctx* c = e->c;
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    int64_t off = 0;
    for (int k = 0; k < e->nsims; k++) {
        const int64_t st = e->start[k], n = e->count[k];
        if (pos)
            GH_HIP(c, hipMemcpyAsync(pos + 3 * off, e->x.as<double>() + 3 * st,
                                     sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        if (vel)
            GH_HIP(c, hipMemcpyAsync(vel + 3 * off, e->v.as<double>() + 3 * st,
                                     sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        off += n;
    }
    GH_HIP(c, hipStreamSynchronize(s));
    return 0;
}

int gh_ensemble_info(const gh_ensemble* e, int64_t* n_pad, int64_t* block_pairs,
                     int64_t* pairs_per_step)
{
    if (!e) return GH_EARG;
    int64_t pairs = 0;
    for (int64_t n : e->count) pairs += n * (n - 1);
    if (n_pad) *n_pad = e->n_pad;
    if (block_pairs) *block_pairs = e->items;
    if (pairs_per_step) *pairs_per_step = pairs;
    return 0;
}

void gh_ensemble_destroy(gh_ensemble* e)
{
    if (!e) return;
    (void)hipSetDevice(e->c->device);
    (void)hipStreamSynchronize(e->c->stream);
    ens_free(e);
    delete e;
}

}  // extern "C"
