// Exported C ABI (include/gravhop.h): contexts, host-pointer wrappers, the
// device-resident leapfrog loop and diagnostics.
#include <algorithm>
#include <cstring>
#include <utility>

#include "gh_internal.h"

namespace gh {

int set_slot(gh_ctx* c, long long* slot, long long value, int add, hipStream_t s);
int energy_dev(gh_ctx* c, const double* d_pos, const double* d_vel, const double* d_mass,
               int64_t n, double eps, double* d_out5, hipStream_t s);

int fail(gh_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

int ensure(gh_ctx* c, DevBuf& b, size_t bytes)
{
    if (bytes <= b.cap) return 0;
    if (b.p) {
        // the buffer may still be read by queued work, on the context's stream or on a
        // stream a _dev caller passed in (growth is rare, so wait for the whole device)
        GH_HIP(c, hipDeviceSynchronize());
        GH_HIP(c, hipFree(b.p));
        c->arena_bytes -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;  // some headroom for slowly growing inputs
    want = (want + 255) & ~size_t(255);
    const int limit_mb = c->tune[GH_TUNE_ALLOC_LIMIT_MB];
    if (limit_mb > 0 && c->arena_bytes + want > ((size_t)limit_mb << 20))
        return fail(c, GH_ENOMEM,
                    "device arena limit of %d MiB reached: %zu bytes requested, %zu held",
                    limit_mb, want, c->arena_bytes);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, GH_ENOMEM, "device allocation of %zu bytes failed: %s", want,
                    hipGetErrorString(e));
    }
    b.cap = want;
    c->arena_bytes += want;
    return 0;
}

int check_launch(gh_ctx* c, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(c, GH_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return 0;
}

static hipEvent_t take_event(gh_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedScope::TimedScope(gh_ctx* c_, hipStream_t s_, int cls_) : c(c_), s(s_), cls(cls_)
{
    if (!c->timing) return;
    b = take_event(c);
    if (b) (void)hipEventRecord(b, s);
}

TimedScope::~TimedScope()
{
    if (!b) return;
    hipEvent_t e = take_event(c);
    if (!e) return;
    (void)hipEventRecord(e, s);
    c->spans.push_back({b, e, cls});
}

StreamOrder::StreamOrder(gh_ctx* c_, hipStream_t s_) : c(c_), s(s_)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return;
    live = true;
    if (c->last_event && c->last_stream != s) (void)hipStreamWaitEvent(s, c->last_event, 0);
}

StreamOrder::~StreamOrder()
{
    if (!live) return;
    if (!c->last_event &&
        hipEventCreateWithFlags(&c->last_event, hipEventDisableTiming) != hipSuccess) {
        c->last_event = nullptr;
        return;
    }
    if (hipEventRecord(c->last_event, s) == hipSuccess) c->last_stream = s;
}

static void release_spans(gh_ctx* c)
{
    for (auto& sp : c->spans) {
        c->event_pool.push_back(sp.a);
        c->event_pool.push_back(sp.b);
    }
    c->spans.clear();
}

void release(gh_ctx* c, DevBuf& b)
{
    if (b.p) {
        (void)hipFree(b.p);
        c->arena_bytes -= b.cap;
    }
    b.p = nullptr;
    b.cap = 0;
}

// Host <-> device staging for the host-pointer entry points.
static int upload(gh_ctx* c, DevBuf& b, const double* h, int64_t count)
{
    GH_TRY(ensure(c, b, sizeof(double) * (size_t)(count > 0 ? count : 1)));
    if (count > 0)
        GH_HIP(c, hipMemcpyAsync(b.p, h, sizeof(double) * (size_t)count,
                                 hipMemcpyHostToDevice, c->stream));
    return 0;
}

static int download(gh_ctx* c, double* h, const DevBuf& b, int64_t count)
{
    if (count > 0)
        GH_HIP(c, hipMemcpyAsync(h, b.p, sizeof(double) * (size_t)count,
                                 hipMemcpyDeviceToHost, c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int check_common(gh_ctx* c, int64_t np, int64_t nf, double eps, int precision)
{
    if (!c) return GH_EARG;
    if (np < 0 || nf < 0) return fail(c, GH_EARG, "negative particle count");
    if (np > (int64_t)1 << 31 || nf > (int64_t)1 << 31)
        return fail(c, GH_EARG, "more than 2^31 particles or targets");
    if (precision != GH_FP64 && precision != GH_FP32)
        return fail(c, GH_EARG, "unknown precision %d", precision);
    (void)eps;
    return 0;
}

// One step: force at xh, then the fused kick/drift/half-drift.  slot != nullptr: the
// history slot is read on the device (graph replay), and *slot_next becomes *slot + 1;
// otherwise hp/hv point at the step's snapshot.
static int one_step(gh_ctx* c, double* x, double* v, const double* m, int64_t np,
                    double dt, double eps, int algorithm, double theta, int precision,
                    double* hp, double* hv, const long long* slot, hipStream_t s,
                    long long* slot_next = nullptr)
{
    const int64_t n3 = 3 * np;
    double* xh = c->xh.as<double>();
    double* acc = c->acc64.as<double>();
    if (np > 1) {
        if (algorithm == GH_DIRECT)
            GH_TRY(direct_dev(c, xh, m, np, nullptr, np, eps, precision, acc, s));
        else
            GH_TRY(tree_dev(c, xh, m, np, nullptr, np, eps, theta, precision, acc, s));
    } else {
        // gravhopper.py:442,449-450: a single particle feels no N-body force
        GH_TRY(zero_dev(c, acc, n3, s));
    }
    TimedScope ts(c, s, GH_T_UPDATE);
    return kick_drift(c, acc, np, dt, xh, v, x, hp, hv, s, slot, nullptr, slot_next);
}

// Copies a run's device history to host memory in slices of steps, on the copy stream,
// each once an event recorded after its last step has fired.  A copy to pageable host
// memory holds the calling thread until it is done, so callers issue a slice's copy
// only after later steps are queued (or after the steps have finished).
struct HistCopier {
    gh_ctx* c;
    hipStream_t s, cs;
    int64_t n3, nsteps, slice;
    const double *dp, *dv;  // device history (nullable)
    double *hp, *hv;        // host destinations (nullable)
    bool on;
    int64_t marked = 0;  // steps [0, marked) belong to recorded slices
    size_t issued = 0;   // slices whose copies have been issued
    std::vector<std::pair<int64_t, int64_t>> slices;
    std::vector<hipEvent_t> events;

    HistCopier(gh_ctx* c_, hipStream_t s_, int64_t n3_, int64_t nsteps_, int64_t slice_,
               const double* dp_, const double* dv_, double* hp_, double* hv_)
        : c(c_), s(s_), cs(c_->copy_stream), n3(n3_), nsteps(nsteps_), slice(slice_), dp(dp_),
          dv(dv_), hp(dp_ ? hp_ : nullptr), hv(dv_ ? hv_ : nullptr),
          on((hp || hv) && c_->copy_stream != nullptr)
    {
    }
    ~HistCopier()
    {
        if (on) (void)hipStreamSynchronize(cs);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
    // steps [marked, done) are queued on s: record them as a slice once it is long
    // enough or the run is complete
    int mark(int64_t done)
    {
        if (!on || (done - marked < slice && done < nsteps)) return 0;
        hipEvent_t e = nullptr;
        GH_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        events.push_back(e);
        GH_HIP(c, hipEventRecord(e, s));
        slices.emplace_back(marked, done);
        marked = done;
        return 0;
    }
    // issue the copies of the recorded slices [issued, upto)
    int issue(size_t upto)
    {
        for (; issued < upto && issued < slices.size(); issued++) {
            const int64_t a = slices[issued].first, b = slices[issued].second;
            const size_t bytes = sizeof(double) * (size_t)(n3 * (b - a));
            GH_HIP(c, hipStreamWaitEvent(cs, events[issued], 0));
            if (hp)
                GH_HIP(c, hipMemcpyAsync(hp + a * n3, dp + a * n3, bytes, hipMemcpyDeviceToHost,
                                         cs));
            if (hv)
                GH_HIP(c, hipMemcpyAsync(hv + a * n3, dv + a * n3, bytes, hipMemcpyDeviceToHost,
                                         cs));
        }
        return 0;
    }
    int finish()
    {
        GH_TRY(issue(slices.size()));
        GH_HIP(c, hipStreamSynchronize(cs));
        return 0;
    }
    bool complete() const { return on && marked == nsteps && issued == slices.size(); }
};

// nsteps DKD steps on device state; d_hist_* (nullable) receive (nsteps, n, 3).
// Direct-sum runs replay two captured steps as one HIP graph after a first eager step
// (which sizes every buffer); an odd last step runs eagerly.  In the graph the history
// slot comes from two device counters that the steps' update kernels advance in turn
// (step A reads slot[0] and writes slot[1] = slot[0] + 1, step B the reverse), so a step
// is two launches.  The tree's build reads its node count back to the host, so tree
// runs stay eager.
// out_p / out_v (nullable, host memory; needs the context's copy stream): the run copies
// its history there itself while later steps still run, and sets *copied; otherwise the
// caller copies the device history afterwards.  Replayed runs queue every step first and
// then copy slices of at least 8 steps (at most about 16); eager runs copy slices of
// about 16 MiB, each after the next step has been queued.
static int run_steps(gh_ctx* c, double* x, double* v, const double* m, int64_t np, double dt,
                     double eps, int algorithm, double theta, int precision, int64_t nsteps,
                     double* hist_p, double* hist_v, hipStream_t s, double* out_p = nullptr,
                     double* out_v = nullptr, bool* copied = nullptr)
{
    if (copied) *copied = false;
    if (np == 0 || nsteps == 0) return 0;
    const int64_t n3 = 3 * np;
    GH_TRY(ensure(c, c->xh, sizeof(double) * (size_t)n3));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * (size_t)n3));
    GH_TRY(half_drift(c, x, v, np, dt, c->xh.as<double>(), s));
    GH_TRY(one_step(c, x, v, m, np, dt, eps, algorithm, theta, precision, hist_p, hist_v,
                    nullptr, s));
    const bool graph = algorithm == GH_DIRECT && !c->timing && nsteps >= 4 &&
                       c->tune[GH_TUNE_RUN_GRAPH] >= 0;
    if (!graph) {
        const int64_t step_bytes = (int64_t)sizeof(double) * n3 * 2;
        HistCopier hc(c, s, n3, nsteps, std::max<int64_t>(1, ((int64_t)16 << 20) / step_bytes),
                      hist_p, hist_v, out_p, out_v);
        GH_TRY(hc.mark(1));
        for (int64_t st = 1; st < nsteps; st++) {
            GH_TRY(one_step(c, x, v, m, np, dt, eps, algorithm, theta, precision,
                            hist_p ? hist_p + st * n3 : nullptr,
                            hist_v ? hist_v + st * n3 : nullptr, nullptr, s));
            // step st is queued: copy every slice recorded before it
            GH_TRY(hc.issue(hc.slices.size()));
            GH_TRY(hc.mark(st + 1));
        }
        GH_TRY(hc.finish());
        if (copied) *copied = hc.complete();
        return 0;
    }
    GH_TRY(ensure(c, c->slot, 256));
    long long* slot = c->slot.as<long long>();
    GH_TRY(set_slot(c, slot, 1, 0, s));
    GH_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = one_step(c, x, v, m, np, dt, eps, algorithm, theta, precision, hist_p, hist_v,
                      slot, s, slot + 1);
    if (!rc)
        rc = one_step(c, x, v, m, np, dt, eps, algorithm, theta, precision, hist_p, hist_v,
                      slot + 1, s, slot);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ee != hipSuccess)
        return fail(c, GH_EHIP, "step capture failed: %s", hipGetErrorString(ee));
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(c, GH_EHIP, "step graph instantiation failed: %s", hipGetErrorString(ie));
    {
        HistCopier hc(c, s, n3, nsteps, std::max<int64_t>(8, (nsteps + 15) / 16), hist_p,
                      hist_v, out_p, out_v);
        if (!rc) rc = hc.mark(1);
        int64_t st = 1;
        for (; st + 1 < nsteps && !rc; st += 2) {
            if (hipGraphLaunch(ex, s) != hipSuccess) rc = fail(c, GH_EHIP, "graph launch failed");
            if (!rc) rc = hc.mark(st + 2);
        }
        if (!rc && st < nsteps) {
            rc = one_step(c, x, v, m, np, dt, eps, algorithm, theta, precision,
                          hist_p ? hist_p + st * n3 : nullptr,
                          hist_v ? hist_v + st * n3 : nullptr, nullptr, s);
            if (!rc) rc = hc.mark(nsteps);
        }
        // every step is queued: copy the slices out as they complete
        if (!rc) rc = hc.finish();
        if (!rc && copied) *copied = hc.complete();
    }
    // the executable graph may be destroyed once its launches are enqueued
    (void)hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(ex);
    return rc;
}

// nsteps DKD steps on device state with external accelerations from a host callback
// (gravhopper.py:405-473).  Per step, on the context stream s: the N-body force of the
// half-step positions, then the kick and drifts with a = F A + ext.  Meanwhile, on a
// copy stream: x_half (and the pre-kick v when asked) go to pinned host memory, `fn`
// fills ext there at the kick time t + dt/2, and ext goes back to the device.  Times
// advance as t <- t + dt, the reference's running sum (gravhopper.py:320).
static int run_ext_steps(gh_ctx* c, double* x, double* v, const double* m, int64_t np,
                         double dt, double eps, int algorithm, double theta, int precision,
                         int64_t nsteps, double* t, gh_external_fn fn, void* user, int pass_v,
                         double* hist_p, double* hist_v, hipStream_t s, double* out_p = nullptr,
                         double* out_v = nullptr, bool* copied = nullptr)
{
    if (copied) *copied = false;
    const int64_t n3 = 3 * np;
    double* xh = c->xh.as<double>();
    double* acc = c->acc64.as<double>();
    double* ext = c->ext.as<double>();
    double* h_xh = c->pinned;
    double* h_v = c->pinned + n3;
    double* h_ext = c->pinned + 2 * n3;
    hipStream_t cs = c->copy_stream;
    hipEvent_t ready = c->ext_events[0], host = c->ext_events[1], back = c->ext_events[2];
    const size_t bytes = sizeof(double) * (size_t)n3;
    // history to host memory (out_p / out_v) in slices of about 16 MiB, each copied on the
    // copy stream after the next step's force is queued, so that it overlaps that force
    HistCopier hc(c, s, n3, nsteps,
                  std::max<int64_t>(1, ((int64_t)16 << 20) / (int64_t)(2 * bytes)), hist_p,
                  hist_v, out_p, out_v);
    GH_TRY(half_drift(c, x, v, np, dt, xh, s));
    for (int64_t st = 0; st < nsteps; st++) {
        GH_HIP(c, hipEventRecord(ready, s));
        GH_HIP(c, hipStreamWaitEvent(cs, ready, 0));
        GH_HIP(c, hipMemcpyAsync(h_xh, xh, bytes, hipMemcpyDeviceToHost, cs));
        if (pass_v) GH_HIP(c, hipMemcpyAsync(h_v, v, bytes, hipMemcpyDeviceToHost, cs));
        GH_HIP(c, hipEventRecord(host, cs));
        if (np > 1) {
            if (algorithm == GH_DIRECT)
                GH_TRY(direct_dev(c, xh, m, np, nullptr, np, eps, precision, acc, s));
            else
                GH_TRY(tree_dev(c, xh, m, np, nullptr, np, eps, theta, precision, acc, s));
        } else {
            GH_TRY(zero_dev(c, acc, n3, s));
        }
        GH_TRY(hc.issue(hc.slices.size()));
        GH_HIP(c, hipEventSynchronize(host));
        const int rc = fn(user, np, h_xh, pass_v ? h_v : nullptr, *t + 0.5 * dt, h_ext);
        if (rc != 0)
            return fail(c, GH_ECALLBACK, "external force callback returned %d at step %lld",
                        rc, (long long)st);
        GH_HIP(c, hipMemcpyAsync(ext, h_ext, bytes, hipMemcpyHostToDevice, cs));
        GH_HIP(c, hipEventRecord(back, cs));
        GH_HIP(c, hipStreamWaitEvent(s, back, 0));
        {
            TimedScope ts(c, s, GH_T_UPDATE);
            GH_TRY(kick_drift(c, acc, np, dt, xh, v, x, hist_p ? hist_p + st * n3 : nullptr,
                              hist_v ? hist_v + st * n3 : nullptr, s, nullptr, ext));
        }
        *t = *t + dt;
        GH_TRY(hc.mark(st + 1));
    }
    GH_TRY(hc.finish());
    if (copied) *copied = hc.complete();
    return 0;
}

}  // namespace gh

using namespace gh;

extern "C" {

int gh_device_count(int* count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return 0;
}

int gh_ctx_create(int device, gh_ctx** out)
{
    if (!out) return GH_EARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GH_EHIP;
    if (device < 0 || device >= n) return GH_EARG;
    if (hipSetDevice(device) != hipSuccess) return GH_EHIP;
    gh_ctx* c = new gh_ctx();
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        c->n_cu <= 0)
        c->n_cu = 256;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return GH_EHIP;
    }
    *out = c;
    return 0;
}

void gh_ctx_destroy(gh_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->last_event) (void)hipEventSynchronize(c->last_event);  // a call on a caller's stream
    tree_free(c);
    DevBuf* bufs[] = {&c->pos64, &c->mass64, &c->fpos64, &c->acc64, &c->src64,
                      &c->src32a, &c->src32b, &c->tgt32,  &c->xs,    &c->vs,
                      &c->xh,     &c->hist_p, &c->hist_v, &c->scratch, &c->counters,
                      &c->work,   &c->eps_blk, &c->adj_work, &c->adj_src, &c->adj_mu,
                      &c->adj_mv, &c->adj_part, &c->adj_seg, &c->adj_perm, &c->slot};
    for (DevBuf* b : bufs) release(c, *b);
    release(c, c->ext);
    for (OrderBufs* o : {&c->ord_p, &c->ord_t}) {
        DevBuf* ob[] = {&o->keys[0], &o->keys[1], &o->idx[0], &o->idx[1], &o->hist, &o->box,
                        &o->boxpart, &o->loc,     &o->s64,    &o->frame,  &o->flag,
                        &o->nflag};
        for (DevBuf* b : ob) release(c, *b);
    }
    release_spans(c);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ext_events)
        if (e) (void)hipEventDestroy(e);
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->last_event) (void)hipEventDestroy(c->last_event);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* gh_last_error(const gh_ctx* c) { return c ? c->err.c_str() : "no context"; }

int gh_sync(gh_ctx* c)
{
    if (!c) return GH_EARG;
    GH_HIP(c, hipSetDevice(c->device));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    if (c->last_event) GH_HIP(c, hipEventSynchronize(c->last_event));
    return 0;
}

int gh_direct_summation_dev(gh_ctx* c, const double* d_pos, const double* d_mass,
                            int64_t np, const double* d_fpos, int64_t nf, double eps,
                            int precision, double* d_acc, void* stream)
{
    GH_TRY(check_common(c, np, nf, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, pick_stream(c, stream));
    return direct_dev(c, d_pos, d_mass, np, d_fpos, nf, eps, precision, d_acc,
                      pick_stream(c, stream));
}

int gh_direct_summation(gh_ctx* c, const double* pos, const double* mass, int64_t np,
                        double eps, int precision, double* acc)
{
    GH_TRY(check_common(c, np, 0, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    GH_TRY(upload(c, c->pos64, pos, 3 * np));
    GH_TRY(upload(c, c->mass64, mass, np));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 3 * (size_t)(np > 0 ? np : 1)));
    GH_TRY(direct_dev(c, c->pos64.as<double>(), c->mass64.as<double>(), np, nullptr, np, eps,
                      precision, c->acc64.as<double>(), c->stream));
    return download(c, acc, c->acc64, 3 * np);
}

int gh_direct_summation_position(gh_ctx* c, const double* pos, const double* mass,
                                 int64_t np, const double* fpos, int64_t nf, double eps,
                                 int precision, double* acc)
{
    GH_TRY(check_common(c, np, nf, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    GH_TRY(upload(c, c->pos64, pos, 3 * np));
    GH_TRY(upload(c, c->mass64, mass, np));
    GH_TRY(upload(c, c->fpos64, fpos, 3 * nf));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 3 * (size_t)(nf > 0 ? nf : 1)));
    GH_TRY(direct_dev(c, c->pos64.as<double>(), c->mass64.as<double>(), np,
                      c->fpos64.as<double>(), nf, eps, precision, c->acc64.as<double>(),
                      c->stream));
    return download(c, acc, c->acc64, 3 * nf);
}

int gh_tree_force_dev(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np,
                      const double* d_fpos, int64_t nf, double eps, double theta,
                      int precision, double* d_acc, void* stream)
{
    GH_TRY(check_common(c, np, nf, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, pick_stream(c, stream));
    return tree_dev(c, d_pos, d_mass, np, d_fpos, d_fpos ? nf : np, eps, theta, precision,
                    d_acc, pick_stream(c, stream));
}

int gh_tree_force(gh_ctx* c, const double* pos, const double* mass, int64_t np, double eps,
                  double theta, int precision, double* acc)
{
    GH_TRY(check_common(c, np, 0, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    GH_TRY(upload(c, c->pos64, pos, 3 * np));
    GH_TRY(upload(c, c->mass64, mass, np));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 3 * (size_t)(np > 0 ? np : 1)));
    GH_TRY(tree_dev(c, c->pos64.as<double>(), c->mass64.as<double>(), np, nullptr, np, eps,
                    theta, precision, c->acc64.as<double>(), c->stream));
    return download(c, acc, c->acc64, 3 * np);
}

int gh_tree_force_position(gh_ctx* c, const double* pos, const double* mass, int64_t np,
                           const double* fpos, int64_t nf, double eps, double theta,
                           int precision, double* acc)
{
    GH_TRY(check_common(c, np, nf, eps, precision));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    GH_TRY(upload(c, c->pos64, pos, 3 * np));
    GH_TRY(upload(c, c->mass64, mass, np));
    GH_TRY(upload(c, c->fpos64, fpos, 3 * nf));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 3 * (size_t)(nf > 0 ? nf : 1)));
    GH_TRY(tree_dev(c, c->pos64.as<double>(), c->mass64.as<double>(), np,
                    c->fpos64.as<double>(), nf, eps, theta, precision, c->acc64.as<double>(),
                    c->stream));
    return download(c, acc, c->acc64, 3 * nf);
}

int gh_run_dev(gh_ctx* c, double* d_pos, double* d_vel, const double* d_mass, int64_t np,
               double dt, double eps, int algorithm, double theta, int precision,
               int64_t nsteps, double* d_hist_pos, double* d_hist_vel, void* stream)
{
    GH_TRY(check_common(c, np, 0, eps, precision));
    if (algorithm != GH_DIRECT && algorithm != GH_TREE)
        return fail(c, GH_EARG, "unknown algorithm %d", algorithm);
    if (nsteps < 0) return fail(c, GH_EARG, "negative step count");
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, pick_stream(c, stream));
    return run_steps(c, d_pos, d_vel, d_mass, np, dt, eps, algorithm, theta, precision, nsteps,
                     d_hist_pos, d_hist_vel, pick_stream(c, stream));
}

int gh_run(gh_ctx* c, double* pos, double* vel, const double* mass, int64_t np, double dt,
           double eps, int algorithm, double theta, int precision, int64_t nsteps,
           double* hist_pos, double* hist_vel)
{
    GH_TRY(check_common(c, np, 0, eps, precision));
    if (algorithm != GH_DIRECT && algorithm != GH_TREE)
        return fail(c, GH_EARG, "unknown algorithm %d", algorithm);
    if (nsteps < 0) return fail(c, GH_EARG, "negative step count");
    if (np == 0 || nsteps == 0) return 0;
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    const int64_t n3 = 3 * np;
    GH_TRY(upload(c, c->xs, pos, n3));
    GH_TRY(upload(c, c->vs, vel, n3));
    GH_TRY(upload(c, c->mass64, mass, np));
    double* x = c->xs.as<double>();
    double* v = c->vs.as<double>();
    const double* m = c->mass64.as<double>();

    // History goes through a device ring of `chunk` snapshots, copied out per chunk.
    // A new chunk recomputes the half drift from (x, v), which reproduces the carried
    // value bit for bit, so chunking does not change the trajectory.
    const bool hist = hist_pos || hist_vel;
    int64_t chunk = nsteps;
    if (hist) {
        const size_t per = sizeof(double) * (size_t)n3;
        const int64_t cap = (int64_t)((size_t)1 << 31) / (int64_t)per;
        chunk = cap < 1 ? 1 : (cap < nsteps ? cap : nsteps);
        if (hist_pos) GH_TRY(ensure(c, c->hist_p, per * (size_t)chunk));
        if (hist_vel) GH_TRY(ensure(c, c->hist_v, per * (size_t)chunk));
    }
    if (hist && !c->copy_stream)
        GH_HIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    for (int64_t s0 = 0; s0 < nsteps; s0 += chunk) {
        const int64_t s1 = (s0 + chunk < nsteps) ? s0 + chunk : nsteps;
        bool copied = false;
        GH_TRY(run_steps(c, x, v, m, np, dt, eps, algorithm, theta, precision, s1 - s0,
                         hist_pos ? c->hist_p.as<double>() : nullptr,
                         hist_vel ? c->hist_v.as<double>() : nullptr, s,
                         hist_pos ? hist_pos + s0 * n3 : nullptr,
                         hist_vel ? hist_vel + s0 * n3 : nullptr, &copied));
        if (copied) continue;
        if (hist_pos)
            GH_HIP(c, hipMemcpyAsync(hist_pos + s0 * n3, c->hist_p.p,
                                     sizeof(double) * (size_t)(n3 * (s1 - s0)),
                                     hipMemcpyDeviceToHost, s));
        if (hist_vel)
            GH_HIP(c, hipMemcpyAsync(hist_vel + s0 * n3, c->hist_v.p,
                                     sizeof(double) * (size_t)(n3 * (s1 - s0)),
                                     hipMemcpyDeviceToHost, s));
        if (hist) GH_HIP(c, hipStreamSynchronize(s));
    }
    GH_HIP(c, hipMemcpyAsync(pos, x, sizeof(double) * (size_t)n3, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(vel, v, sizeof(double) * (size_t)n3, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipStreamSynchronize(s));
    return 0;
}

int gh_run_ext(gh_ctx* c, double* pos, double* vel, const double* mass, int64_t np, double dt,
               double eps, int algorithm, double theta, int precision, int64_t nsteps,
               double t0, gh_external_fn fn, void* user, int pass_velocity, double* hist_pos,
               double* hist_vel)
{
    GH_TRY(check_common(c, np, 0, eps, precision));
    if (algorithm != GH_DIRECT && algorithm != GH_TREE)
        return fail(c, GH_EARG, "unknown algorithm %d", algorithm);
    if (nsteps < 0) return fail(c, GH_EARG, "negative step count");
    if (!fn) return fail(c, GH_EARG, "no external force callback");
    if (np == 0 || nsteps == 0) return 0;
    GH_HIP(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    StreamOrder order_(c, s);
    const int64_t n3 = 3 * np;
    const size_t bytes = sizeof(double) * (size_t)n3;
    if (!c->copy_stream) GH_HIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    for (hipEvent_t& e : c->ext_events)
        if (!e) GH_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (c->pinned_bytes < 3 * bytes) {
        if (c->pinned) {
            GH_HIP(c, hipStreamSynchronize(c->copy_stream));
            GH_HIP(c, hipHostFree(c->pinned));
            c->pinned = nullptr;
            c->pinned_bytes = 0;
        }
        GH_HIP(c, hipHostMalloc((void**)&c->pinned, 3 * bytes, hipHostMallocDefault));
        c->pinned_bytes = 3 * bytes;
    }
    GH_TRY(ensure(c, c->xh, bytes));
    GH_TRY(ensure(c, c->acc64, bytes));
    GH_TRY(ensure(c, c->ext, bytes));
    GH_TRY(upload(c, c->xs, pos, n3));
    GH_TRY(upload(c, c->vs, vel, n3));
    GH_TRY(upload(c, c->mass64, mass, np));
    double* x = c->xs.as<double>();
    double* v = c->vs.as<double>();
    const double* m = c->mass64.as<double>();
    // history through a device ring of `chunk` snapshots, as in gh_run
    const bool hist = hist_pos || hist_vel;
    int64_t chunk = nsteps;
    if (hist) {
        const int64_t cap = (int64_t)((size_t)1 << 31) / (int64_t)bytes;
        chunk = cap < 1 ? 1 : (cap < nsteps ? cap : nsteps);
        if (hist_pos) GH_TRY(ensure(c, c->hist_p, bytes * (size_t)chunk));
        if (hist_vel) GH_TRY(ensure(c, c->hist_v, bytes * (size_t)chunk));
    }
    double t = t0;
    for (int64_t s0 = 0; s0 < nsteps; s0 += chunk) {
        const int64_t s1 = (s0 + chunk < nsteps) ? s0 + chunk : nsteps;
        bool copied = false;
        const int rc = run_ext_steps(c, x, v, m, np, dt, eps, algorithm, theta, precision,
                                     s1 - s0, &t, fn, user, pass_velocity,
                                     hist_pos ? c->hist_p.as<double>() : nullptr,
                                     hist_vel ? c->hist_v.as<double>() : nullptr, s,
                                     hist_pos ? hist_pos + s0 * n3 : nullptr,
                                     hist_vel ? hist_vel + s0 * n3 : nullptr, &copied);
        if (rc) {
            // leave both streams idle before returning the error
            (void)hipStreamSynchronize(c->copy_stream);
            (void)hipStreamSynchronize(s);
            return rc;
        }
        if (copied) continue;
        if (hist_pos)
            GH_HIP(c, hipMemcpyAsync(hist_pos + s0 * n3, c->hist_p.p, bytes * (size_t)(s1 - s0),
                                     hipMemcpyDeviceToHost, s));
        if (hist_vel)
            GH_HIP(c, hipMemcpyAsync(hist_vel + s0 * n3, c->hist_v.p, bytes * (size_t)(s1 - s0),
                                     hipMemcpyDeviceToHost, s));
        if (hist) GH_HIP(c, hipStreamSynchronize(s));
    }
    GH_HIP(c, hipMemcpyAsync(pos, x, bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipMemcpyAsync(vel, v, bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(c, hipStreamSynchronize(s));
    return 0;
}

int gh_set_tuning(gh_ctx* c, int key, int value)
{
    if (!c) return GH_EARG;
    if (key < 0 || key >= GH_TUNE_COUNT) return fail(c, GH_EARG, "unknown tuning key %d", key);
    c->tune[key] = value;
    return 0;
}

int gh_timing_enable(gh_ctx* c, int on)
{
    if (!c) return GH_EARG;
    GH_HIP(c, hipSetDevice(c->device));
    GH_HIP(c, hipDeviceSynchronize());
    release_spans(c);
    c->timing = (on != 0);
    return 0;
}

int gh_timing_read(gh_ctx* c, int cls, double* total_ms, int64_t* launches)
{
    if (!c) return GH_EARG;
    if (cls < 0 || cls >= GH_T_COUNT) return fail(c, GH_EARG, "unknown kernel class %d", cls);
    GH_HIP(c, hipSetDevice(c->device));
    double tot = 0.0;
    int64_t n = 0;
    for (auto& sp : c->spans) {
        if (sp.cls != cls) continue;
        GH_HIP(c, hipEventSynchronize(sp.b));
        float ms = 0.f;
        GH_HIP(c, hipEventElapsedTime(&ms, sp.a, sp.b));
        tot += ms;
        n++;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return 0;
}

int gh_energy(gh_ctx* c, const double* pos, const double* vel, const double* mass,
              int64_t np, double eps, double* kinetic, double* potential, double* momentum)
{
    GH_TRY(check_common(c, np, 0, eps, GH_FP64));
    GH_HIP(c, hipSetDevice(c->device));
    StreamOrder order_(c, c->stream);
    GH_TRY(upload(c, c->pos64, pos, 3 * np));
    GH_TRY(upload(c, c->vs, vel, 3 * np));
    GH_TRY(upload(c, c->mass64, mass, np));
    GH_TRY(ensure(c, c->acc64, sizeof(double) * 8));
    GH_TRY(energy_dev(c, c->pos64.as<double>(), c->vs.as<double>(), c->mass64.as<double>(), np,
                      eps, c->acc64.as<double>(), c->stream));
    double out[5];
    GH_HIP(c, hipMemcpyAsync(out, c->acc64.p, sizeof(out), hipMemcpyDeviceToHost, c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    if (kinetic) *kinetic = out[0];
    if (potential) *potential = out[1];
    if (momentum) std::memcpy(momentum, out + 2, 3 * sizeof(double));
    return 0;
}

}  // extern "C"
