// Internal declarations shared by the libgravhop.so translation units.
// Everything here is C++/HIP; the exported C ABI is in include/gravhop.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gravhop.h"

namespace gh {

// Reference unit scales (jbgrav.py:40-48, gravhopper.py:409,416): G*Msun/kpc^2 in
// km/s/Myr and (km/s)*Myr in kpc, from CODATA-2018 G, IAU-2015 GM_sun, Julian year.
constexpr double kAccelScale = 4.398600412921223e-09;
constexpr double kDriftScale = 1.022712165045695e-03;

constexpr int kBlock = 256;  // threads per workgroup for every 1-D kernel
// Wavefronts that share the FP64 sums of 64 particles' partial slots (k_reduce_sym,
// k_adj_post): each adds every kSlotLanes-th slot, then their sums are combined.
constexpr int kSlotLanes = 4;

// A device buffer that only grows; reused across calls so the hot loop never allocates.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct TreeBufs;  // gh_tree.hip

typedef unsigned long long u64;
typedef unsigned int u32;

// Spatial order and local frames of the FP32 pair kernels (gh_frames.hip).  Points are
// sorted along a Morton curve and cut into groups of consecutive sorted points; each
// group has an FP64 origin (one of its points), and FP32 coordinates are only ever
// formed relative to such an origin, so their rounding error scales with the group's
// extent rather than with the distance from the coordinate origin.
struct OrderBufs {
    DevBuf keys[2], idx[2], hist, box, boxpart;  // the sort
    DevBuf loc;    // float4 (x - o, y - o, z - o, m) per sorted point, o its group's origin
    DevBuf s64;    // double4 (x, y, z, m) per sorted point
    DevBuf frame;  // double4 per group: origin (x, y, z), largest |x - o| (max norm)
    DevBuf flag;   // int per group: 1 if the group is not compact (see k_frames)
    DevBuf nflag;  // int: number of flagged groups (cleared by the key kernels)
    int sel = 0;   // which of idx[0] / idx[1] holds the order
    const double* sorted = nullptr;  // positions spatial_order last ordered (n of them)
    int64_t sorted_n = 0;
    const u32* perm() const { return idx[sel].as<u32>(); }
};

}  // namespace gh

struct gh_ctx {
    int device = 0;
    int n_cu = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    std::string err;
    // generic arenas
    gh::DevBuf pos64, mass64, fpos64, acc64, src64, src32a, src32b, tgt32;
    gh::DevBuf xs, vs, xh, hist_p, hist_v, scratch, counters;
    gh::TreeBufs* tree = nullptr;
    int tune[GH_TUNE_COUNT] = {0};  // gh_set_tuning values, 0 = automatic
    size_t arena_bytes = 0;         // device memory held by the context's arenas
    gh::DevBuf work;                // block-pair list of the pair-once direct kernel
    gh::DevBuf eps_blk;             // its per-block squared softening
    int sym_nb = -1, sym_ipt = -1, sym_w = -1;  // shape of the block-pair list
    std::vector<int64_t> sym_band_start;        // its bands' first items
    float sym_eps2 = -1.f;          // the softening eps_blk holds
    // adjoint workspaces (gh_adjoint.hip)
    gh::DevBuf adj_work, adj_src, adj_mu, adj_mv, adj_part, adj_seg, adj_perm;
    int adj_nb = -1, adj_ipt = -1;
    gh::DevBuf slot;  // history slot counter of graph-replayed runs
    gh::OrderBufs ord_p, ord_t;  // spatial order of sources and of targets (FP32 paths)
    // the stream of the last entry-point call and an event recorded at its end: a call
    // on another stream waits for it first (StreamOrder), since all calls share the
    // context's buffers
    hipStream_t last_stream = nullptr;
    hipEvent_t last_event = nullptr;
    // gh_run_ext: copy stream, pinned host staging (x_half, v, external acceleration)
    // and the device copy of the external acceleration
    hipStream_t copy_stream = nullptr;
    double* pinned = nullptr;
    size_t pinned_bytes = 0;
    gh::DevBuf ext;
    hipEvent_t ext_events[3] = {nullptr, nullptr, nullptr};
    int64_t last_tree_nodes = 0;
    int last_tree_depth = 0;
    // kernel timing (gh_timing_enable / gh_timing_read)
    bool timing = false;
    struct Span {
        hipEvent_t a, b;
        int cls;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> event_pool;
};

namespace gh {

// Error plumbing: every exported function returns 0 or a GH_E* code and leaves a
// message in ctx->err.
int fail(gh_ctx* c, int code, const char* fmt, ...);

#define GH_HIP(ctx, call)                                                           \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return gh::fail((ctx), GH_EHIP, "%s failed: %s (%s:%d)", #call,         \
                            hipGetErrorString(e_), __FILE__, __LINE__);             \
    } while (0)

#define GH_TRY(expr)            \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != 0) return rc_; \
    } while (0)

int ensure(gh_ctx* c, DevBuf& b, size_t bytes);
// Frees an arena buffer and takes its capacity off the context's arena total.
void release(gh_ctx* c, DevBuf& b);

// Brackets the kernels launched between construction and destruction with HIP events
// of class `cls` on stream s when timing is enabled on the context.
struct TimedScope {
    gh_ctx* c;
    hipStream_t s;
    int cls;
    hipEvent_t b = nullptr;
    TimedScope(gh_ctx* c_, hipStream_t s_, int cls_);
    ~TimedScope();
};
// Orders an entry point's work on stream s after the previous call's, whatever stream
// that one used (the _dev entry points take the caller's stream, the others the
// context's).  Construct after selecting the device; the destructor records the end of
// this call.  Inactive on a stream under graph capture.
struct StreamOrder {
    gh_ctx* c;
    hipStream_t s;
    bool live = false;
    StreamOrder(gh_ctx* c_, hipStream_t s_);
    ~StreamOrder();
};
int check_launch(gh_ctx* c, const char* what);
int zero_dev(gh_ctx* c, double* p, int64_t n, hipStream_t s);
inline hipStream_t pick_stream(gh_ctx* c, void* s) {
    return s ? static_cast<hipStream_t>(s) : c->stream;
}
// A wave-uniform value held in a vector register.  On gfx950 a wave64 VALU instruction
// that reads a scalar register issues about every 4 cycles, one that reads only vector
// registers or literals about every 2.2 (measured with independent FMA, MUL and SUB
// chains), so loop-invariant scalars that the pair loops use are copied once into a
// VGPR.  The empty asm makes the copy opaque to the compiler, which would otherwise
// fold the scalar back into the instructions.
__device__ __forceinline__ float in_vgpr(float x)
{
    float r;
    asm volatile("" : "=v"(r) : "0"(x));
    return r;
}

// The zero-separation guard of the FP32 kernels: a pair's scale factor m r^-3 made
// finite (+inf and NaN become FLT_MAX), so that a zero separation (the self pair, a
// coincident pair, a massless padding particle) contributes exactly 0 instead of
// 0 * inf.  Applied when fp32_guard(eps) holds.
__device__ __forceinline__ float clamp_finite(float x) { return fminf(x, 3.402823466e38f); }

inline unsigned blocks_for(int64_t n, int per_block) {
    return static_cast<unsigned>((n + per_block - 1) / per_block);
}

// direct sum (gh_direct.hip)
int direct_dev(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np,
               const double* d_fpos, int64_t nf, double eps, int precision, double* d_acc,
               hipStream_t s);
int64_t pad_to(int64_t n, int64_t m);
int sym_ipt(gh_ctx* c, int64_t np);
// true when the FP32 pair kernels need the zero-separation guard for this softening
bool fp32_guard(double eps);
int direct_sym(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np, int ipt,
               double eps, double* d_acc, hipStream_t s);
// The pair-once kernel over a block-pair list (src: local FP32 coordinates per sorted
// particle, s64: the same particles in FP64, frame/flag: per block), followed by its
// double-single instance for the items with a flagged block.
// jpart (nullable): the j-side partials go to jpart[slot_j][p - joff] with a stride of
// jstride particles instead of part.
int launch_sym(gh_ctx* c, int ipt, bool guard, const float4* src, const double4* s64,
               const double4* frame, const int* flag, const int* nflag, int64_t np_pad,
               const int4* work, int64_t items, const float* eps2_blk, float* part, int* queue,
               unsigned persistent_grid, hipStream_t s, float* jpart = nullptr,
               int64_t jstride = 0, int64_t joff = 0);
// acc[perm[p]] (perm nullable) = the FP64 sum of particle p's partial slots
int reduce_sym(gh_ctx* c, const float* part, int nb, const int* nb_blk, int blk,
               int64_t np_pad, int64_t n, const u32* perm, double* acc, hipStream_t s);

// leapfrog update (gh_step.hip).  kick_drift: ext (nullable) is an extra acceleration
// in km/s/Myr added to the N-body one, as a = acc*A + ext; slot (nullable) is the
// history slot counter of graph-replayed runs, and slot_next (nullable) receives
// *slot + 1.
int half_drift(gh_ctx* c, const double* x, const double* v, int64_t n, double dt, double* xh,
               hipStream_t s);
int kick_drift(gh_ctx* c, const double* acc, int64_t n, double dt, double* xh, double* v,
               double* x, double* hp, double* hv, hipStream_t s,
               const long long* slot = nullptr, const double* ext = nullptr,
               long long* slot_next = nullptr);

// spatial order and frames (gh_frames.hip)
constexpr int kFrameLog2 = 10;  // a group is flagged when R > 2^kFrameLog2 x its mean gap
// Sorts n points (pos (n,3)) along a Morton curve (about 11 levels) in their root box; the
// order is o.perm().
int spatial_order(gh_ctx* c, OrderBufs& o, const double* d_pos, int64_t n, hipStream_t s);
// The same order for a padded layout of nseg segments (segment g: positions
// [start[g], start[g] + count[g]) of pos (n_pad,3), blocks of blk positions, seg[b] the
// segment of block b, all device arrays): each segment sorted in its own box exactly as
// spatial_order sorts it alone, segments in place, padding last in its segment.
int spatial_order_seg(gh_ctx* c, OrderBufs& o, const double* d_pos, int64_t n_pad, int blk,
                      int nseg, const int* seg, const int64_t* start, const int64_t* count,
                      hipStream_t s);
// Fills o.loc, o.s64 (nullable output: skipped if want_s64 is false), o.frame and
// o.flag for n_pad / group points in groups of `group` (256, 512 or 1024) consecutive
// sorted points; perm == nullptr means the points are already in order.  valid
// (nullable) gives each group's count of real points; otherwise groups are full up to
// n.  Padding entries are massless at their group's origin.  mass may be null (w = 0).
int gather_frames(gh_ctx* c, OrderBufs& o, const double* d_pos, const double* d_mass,
                  const u32* perm, int64_t n, int64_t n_pad, int group, const int* valid,
                  bool want_s64, hipStream_t s);

// tree helpers shared with the spatial order (gh_tree.hip)
constexpr int kRsItems = 8;
constexpr int kRsChunk = kBlock * kRsItems;  // elements per radix-sort workgroup
// Sorts of at most this many chunks may take radix_sort_direct.
constexpr int kRsDirectChunks = 32;
// have_hist0: hist already holds the first pass's digit counts (k_rs_hist's layout).
int radix_sort(gh_ctx* c, DevBuf* keys, DevBuf* idx, DevBuf& hist, int64_t n, int words,
               int lo_shift, hipStream_t s, int* sel_out, int hi_shift = 64,
               bool have_hist0 = false);
// One-word stable LSD sort of (keys[0], idx[0]) by bits [lo_shift, hi_shift), 8 per pass,
// one kernel per pass, for at most kRsDirectChunks chunks.  hist holds, per pass p, the
// digit counts of each chunk of that pass's input at hist + p * 256 * nb (digit d, chunk
// b at d * nb + b, nb = the chunk count): pass 0's counted by the caller, the later ones
// zero on entry (each pass counts the next pass's).
int radix_sort_direct(gh_ctx* c, DevBuf* keys, DevBuf* idx, DevBuf& hist, int64_t n,
                      int lo_shift, int hi_shift, hipStream_t s, int* sel_out);
int bbox_dev(gh_ctx* c, const double* d_pos, int64_t n, double eps, DevBuf& part, double* box,
             hipStream_t s);
int bbox_partial_dev(gh_ctx* c, const double* d_pos, int64_t n, DevBuf& part, u32* zero,
                     int64_t nzero, hipStream_t s, int* nparts);

// tree (gh_tree.hip)
int tree_dev(gh_ctx* c, const double* d_pos, const double* d_mass, int64_t np,
             const double* d_fpos, int64_t nf, double eps, double theta, int precision,
             double* d_acc, hipStream_t s);
void tree_free(gh_ctx* c);

}  // namespace gh
