/*
 * gravhop.h -- C ABI of libgravhop.so, the MI355X (gfx950) gravity engine behind
 * gravhopper_amd's drop-in replacements of GravHopper's force entry points and
 * Simulation.run().
 *
 * Conventions
 *   - Code units as in the reference C engine: positions in kpc, masses in Msun,
 *     G = 1; accelerations come back in G*Msun/kpc^2 (gravhopper/_jbgrav.c works in
 *     these units and gravhopper/jbgrav.py:38-48 scales them).
 *   - Position arrays are (n,3) row-major float64 ("AoS"), like the C-contiguous
 *     numpy arrays the reference receives (_jbgrav.c:79-80).
 *   - Every function returns 0 on success or a negative GH_E* code; the message of
 *     the last failure on a context is gh_last_error(ctx).  Nothing calls exit().
 *   - A context owns one HIP stream and device arenas that grow on demand and are
 *     reused.  Calls on one context must not overlap; separate contexts may.
 *   - "_dev" entry points take device pointers and a hipStream_t (passed as void*;
 *     NULL selects the context's stream) and return without synchronising.  The
 *     others take host pointers and return when the result is in host memory.
 *   - Calls on one context run in call order whatever streams they use: a call whose
 *     stream differs from the previous call's first makes its stream wait for the end
 *     of that call (an event), since both use the context's arenas.  On a stream under
 *     graph capture this ordering is left to the caller.  gh_sync waits for the
 *     context's stream and for the last call.
 */
#ifndef GRAVHOP_H
#define GRAVHOP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gh_ctx gh_ctx;

/* error codes */
#define GH_OK 0
#define GH_EARG -1        /* invalid argument (shape, size, value)                   */
#define GH_EHIP -2        /* HIP runtime error                                       */
#define GH_ENOMEM -3      /* device allocation failed                                */
#define GH_ECOINCIDENT -4 /* tree: two particles at the same position                */
#define GH_EDEPTH -5      /* tree: particles closer than the deepest supported level */
#define GH_ECALLBACK -6   /* gh_run_ext: the external force callback failed          */

/* precision of the pair arithmetic */
#define GH_FP64 0 /* IEEE FP64 in the reference's operation order                  */
#define GH_FP32 1 /* FP32 pair forces, FP32 tile sums, two-sum across tiles; FP64 out */

/* force algorithm of gh_run */
#define GH_DIRECT 0
#define GH_TREE 1

/* Context management (no reference counterpart: the reference keeps no state). */
int gh_ctx_create(int device, gh_ctx **out);
void gh_ctx_destroy(gh_ctx *ctx);
const char *gh_last_error(const gh_ctx *ctx);
int gh_device_count(int *count);
int gh_sync(gh_ctx *ctx);

/* Direct summation.  Replaces _jbgrav.direct_summation(pos, mass, eps)
 * (gravhopper/_jbgrav.c:68-135, workhorse :140-193).  acc is (np,3).  Self pairs are
 * excluded by index; with GH_FP64 the result is bit-identical to the reference. */
int gh_direct_summation(gh_ctx *ctx, const double *pos, const double *mass, int64_t np,
                        double eps, int precision, double *acc);

/* Field at force_pos from all particles.  Replaces
 * _jbgrav.direct_summation_position(pos, mass, force_pos, eps)
 * (gravhopper/_jbgrav.c:200-294, workhorse :299-353).  A pair whose softened
 * squared distance is exactly 0 contributes 0 (:327-330).  acc is (nf,3). */
int gh_direct_summation_position(gh_ctx *ctx, const double *pos, const double *mass,
                                 int64_t np, const double *force_pos, int64_t nf,
                                 double eps, int precision, double *acc);

/* Device-pointer form of both: force_pos == NULL means "the particles themselves"
 * (index self-exclusion), otherwise nf targets. */
int gh_direct_summation_dev(gh_ctx *ctx, const double *d_pos, const double *d_mass,
                            int64_t np, const double *d_force_pos, int64_t nf, double eps,
                            int precision, double *d_acc, void *stream);

/* Barnes-Hut octree force.  Replaces _jbgrav.tree_force(pos, mass, eps, theta)
 * (gravhopper/_jbgrav.c:564-630, workhorse :737-806).  The octree is the reference's
 * (one particle per leaf, strict '>' octants, FP64 centre recurrence, bounding-box
 * size rule of :764-769), built on the device; the walk opens a node unless it is a
 * leaf or size/|centre - x| < theta (:496-502) and adds monopoles (:506-524). */
int gh_tree_force(gh_ctx *ctx, const double *pos, const double *mass, int64_t np,
                  double eps, double theta, int precision, double *acc);

/* Replaces _jbgrav.tree_force_position(pos, mass, force_pos, eps, theta)
 * (gravhopper/_jbgrav.c:634-727). */
int gh_tree_force_position(gh_ctx *ctx, const double *pos, const double *mass,
                           int64_t np, const double *force_pos, int64_t nf, double eps,
                           double theta, int precision, double *acc);

/* Device-pointer form of both (force_pos == NULL: targets are the particles). */
int gh_tree_force_dev(gh_ctx *ctx, const double *d_pos, const double *d_mass, int64_t np,
                      const double *d_force_pos, int64_t nf, double eps, double theta,
                      int precision, double *d_acc, void *stream);

/* Canonical serialisation of the device-built octree, depth-first pre-order with
 * children in branch order 0..7 (the visiting order of gravoct_calc_accel,
 * gravhopper/_jbgrav.c:530-537).  One record per node. */
typedef struct {
    int32_t level;   /* root = 0 */
    int32_t branch;  /* octant within the parent (bit0 x, bit1 y, bit2 z); -1 root */
    int64_t leaf;    /* particle index for a leaf, -1 for an internal node */
    double center[3];
    double size;
    double mass;
    double com[3];
} gh_node_rec;

/* Writes min(n_nodes, cap) records to out (host memory, may be NULL when cap == 0)
 * and the node count to *n_nodes. */
int gh_tree_dump(gh_ctx *ctx, const double *pos, const double *mass, int64_t np,
                 double eps, gh_node_rec *out, int64_t cap, int64_t *n_nodes);

/* FP64 tree force together with the number of nodes each target accepts (leaves
 * included): the walk's interaction count, comparable with the reference walk
 * gravoct_calc_accel (gravhopper/_jbgrav.c:487-541).  force_pos == NULL: targets are
 * the particles.  acc (nt,3), counts (nt). */
int gh_tree_interactions(gh_ctx *ctx, const double *pos, const double *mass, int64_t np,
                         const double *force_pos, int64_t nf, double eps, double theta,
                         double *acc, int64_t *counts);

/* FP32 tree force (the group walk) together with each target's interaction-list length:
 * the nodes the group opening test accepted for the target's group, at least as many as
 * the reference's walk accepts for the target.  No reference counterpart. */
int gh_tree_interactions_fp32(gh_ctx *ctx, const double *pos, const double *mass,
                              int64_t np, const double *force_pos, int64_t nf, double eps,
                              double theta, double *acc, int64_t *counts);

/* Device-resident drift-kick-drift leapfrog: nsteps of Simulation.perform_timestep
 * (gravhopper/gravhopper.py:405-416) with the force of calculate_acceleration
 * (:419-459) and no external forces, in code units (kpc, km/s, Msun, Myr):
 *   xh = x + ((0.5 v) dt) D;  a = F(xh) A;  v' = v + (a + 0) dt;  x' = xh + ((0.5 v') dt) D
 * with D = 1.022712165045695e-3 kpc per km/s*Myr and A = 4.398600412921223e-9 km/s/Myr
 * per G*Msun/kpc^2.  pos/vel (np,3) host arrays are advanced in place.  hist_pos /
 * hist_vel (nullable) receive nsteps snapshots of (np,3) after every step. */
int gh_run(gh_ctx *ctx, double *pos, double *vel, const double *mass, int64_t np,
           double dt, double eps, int algorithm, double theta, int precision,
           int64_t nsteps, double *hist_pos, double *hist_vel);

/* External accelerations for gh_run_ext, called once per step from the calling thread
 * with host arrays: the half-step positions x_half (n,3, kpc), the velocities before
 * the kick v (n,3, km/s; NULL unless pass_velocity), and the kick time t (Myr).  Writes
 * the summed external acceleration in km/s/Myr into acc (n,3).  The arrays are
 * library staging buffers, valid during the call only.  Returns 0, or nonzero to stop
 * the run (gh_run_ext then returns GH_ECALLBACK). */
typedef int (*gh_external_fn)(void *user, int64_t n, const double *x_half, const double *v,
                              double t, double *acc);

/* gh_run with external accelerations (gravhopper/gravhopper.py:405-473): the kick uses
 *   a = F(xh) A + ext,  ext = fn(xh, v, t_k + dt/2)
 * where t_0 = t0 and t_{k+1} = t_k + dt, as Simulation.perform_timestep adds the
 * position-, time- and velocity-dependent external forces at the half step (the
 * velocity-dependent ones see the pre-kick velocity).  The state stays on the device;
 * each step sends x_half (and v) to the host and ext back, overlapped with the N-body
 * force.  Arguments otherwise as gh_run. */
int gh_run_ext(gh_ctx *ctx, double *pos, double *vel, const double *mass, int64_t np,
               double dt, double eps, int algorithm, double theta, int precision,
               int64_t nsteps, double t0, gh_external_fn fn, void *user, int pass_velocity,
               double *hist_pos, double *hist_vel);

/* Device-pointer form of gh_run: d_pos/d_vel (np,3) are advanced in place on the
 * device; d_hist_pos/d_hist_vel (nullable) are device arrays of nsteps*(np,3). */
int gh_run_dev(gh_ctx *ctx, double *d_pos, double *d_vel, const double *d_mass, int64_t np,
               double dt, double eps, int algorithm, double theta, int precision,
               int64_t nsteps, double *d_hist_pos, double *d_hist_vel, void *stream);

/* Adjoint of gh_run's direct-sum leapfrog (no reference counterpart).
 * gh_adjoint_forward advances device arrays d_x, d_v (n,3) by nsteps, saving (x, v)
 * every `every` steps into d_ckpt (ceil(nsteps/every) x 2 x (n,3) doubles).
 * gh_adjoint_backward takes dL/dx and dL/dv at the end of the run in d_lx, d_lv and
 * overwrites them with dL/dx and dL/dv at the start; d_lm (n) receives dL/dm.  Each
 * segment is re-run from its checkpoint with the forward kernels of `precision`
 * (bit-identical to the forward pass); the vector-Jacobian products through the
 * softened tidal tensor use FP32 pairs (pair-once kernel) and FP64 sums. */
int gh_adjoint_forward(gh_ctx *ctx, double *d_x, double *d_v, const double *d_m, int64_t n,
                       double dt, double eps, int precision, int64_t nsteps, int64_t every,
                       double *d_ckpt, void *stream);
int gh_adjoint_backward(gh_ctx *ctx, const double *d_ckpt, const double *d_m, int64_t n,
                        double dt, double eps, int precision, int64_t nsteps, int64_t every,
                        double *d_lx, double *d_lv, double *d_lm, void *stream);

/* Batched ensembles (no reference counterpart): n_sims independent direct-sum
 * simulations of uneven size, each advanced by the leapfrog of gh_run with its own
 * softening and time step, all with FP32 pair forces (pair-once kernel, `ipt`
 * particles per lane: 1, 2 or 4; 0 selects 4) and FP64 state.  One force launch per
 * step covers every simulation; `persistent` makes it a fixed grid fed from an atomic
 * work counter.  pos/vel/mass are the simulations' (count[s],3) / (count[s]) arrays
 * concatenated in order; eps and dt have one entry per simulation.  Each simulation's
 * trajectory is bit-identical to running it alone with the same kernel shape. */
typedef struct gh_ensemble gh_ensemble;
int gh_ensemble_create(gh_ctx *ctx, int n_sims, const int64_t *counts, const double *pos,
                       const double *vel, const double *mass, const double *eps,
                       const double *dt, int ipt, int persistent, gh_ensemble **out);
/* nsteps steps; use_graph != 0 replays one captured step as a HIP graph. */
int gh_ensemble_run(gh_ensemble *ens, int64_t nsteps, int use_graph);
/* current full-step positions and velocities (concatenated, host memory) */
int gh_ensemble_get(gh_ensemble *ens, double *pos, double *vel);
int gh_ensemble_info(const gh_ensemble *ens, int64_t *n_padded, int64_t *block_pairs,
                     int64_t *pairs_per_step);
void gh_ensemble_destroy(gh_ensemble *ens);

/* Kernel timing (no reference counterpart).  While enabled, the library brackets its
 * kernels with HIP events on the stream they run on; gh_timing_read synchronises and
 * returns the summed duration [ms] and launch count of one kernel class since the
 * last enable.  Classes: */
#define GH_T_DIRECT_F32 0 /* FP32 direct-sum pair kernel                       */
#define GH_T_DIRECT_F64 1 /* FP64 direct-sum pair kernel                       */
#define GH_T_TREE_BUILD 2 /* octree build (keys, sort, linking, moments)       */
#define GH_T_TREE_WALK 3  /* octree walk                                       */
#define GH_T_UPDATE 4     /* fused leapfrog update                             */
#define GH_T_VJP 5        /* adjoint pair kernel (vector-Jacobian product)     */
#define GH_T_COUNT 6
int gh_timing_enable(gh_ctx *ctx, int on);
int gh_timing_read(gh_ctx *ctx, int kernel_class, double *total_ms, int64_t *launches);

/* Kernel-shape tuning (no reference counterpart).  Keys and values:
 *   GH_TUNE_DIRECT_SELF  FP32 self field: 0 automatic, -1 one-sided kernel,
 *                        1/2/4 pair-once kernel with that many particles per lane
 *   GH_TUNE_DIRECT_IPT   FP32 one-sided kernel: 0 automatic, 1/2/4 targets per lane
 *   GH_TUNE_WALK_GROUP   FP32 tree walk: 0 automatic, 1/2/4/8/16/32/64 targets per
 *                        group (the targets that share one opening test and interaction
 *                        list)
 *   GH_TUNE_RUN_GRAPH    direct-sum gh_run / gh_run_dev: 0 replay two captured steps as
 *                        one HIP graph (from 4 steps on), -1 launch every kernel eagerly
 *   GH_TUNE_WALK_F64EVAL FP32 tree walk: 1 evaluates the same interaction lists in FP64
 *                        (a check of the FP32 arithmetic), 0 FP32
 *   GH_TUNE_DIRECT_F64_TPW  FP64 direct sum: 0 automatic, 1/2/4/8/16/32/64 targets per
 *                        wavefront (64: one target per lane; fewer: the terms of a
 *                        target are computed by 64/value lanes and then added in order),
 *                        -1/-2/-4/-8/-16: that many targets per wavefront, every lane
 *                        computing one source's terms for all of them and each of the
 *                        3 x targets sums added in order by a lane of its own
 *   GH_TUNE_ALLOC_LIMIT_MB  cap, in MiB, on the device memory the context's arenas may
 *                        hold (0: no cap); a call that would exceed it fails with
 *                        GH_ENOMEM and leaves the context usable
 *   GH_TUNE_SYM_BAND     FP32 pair-once self field: 0 automatic (one pass while one FP32
 *                        partial per particle and block fits 13 GiB, else bands of
 *                        block columns, one pass each, whose partials fit it), n > 0
 *                        bands of n block columns
 *   GH_TUNE_SORT_DIRECT  spatial order of the FP32 direct sum: 0 automatic (up to 32
 *                        chunks of 2048 points, one kernel per radix pass), -1 always
 *                        the histogram, scan and scatter kernels per pass
 *   GH_TUNE_TREE_SORT    octree key sort: 0 automatic (radix passes over the digits of
 *                        levels 0-12, then each run of keys equal there sorted on its
 *                        own; a run of more than 64 keys makes the build sort over all
 *                        digits, and so do the context's next 64 builds), 1 the same
 *                        without the 64 builds, -1 always over all digits
 *   GH_TUNE_WALK_PREFIX  FP32 tree walk: 0 each block of 64 groups first decides once the
 *                        nodes decided alike for all of its groups (k_walk_prefix), and
 *                        the groups start from there; -1 every group from the root
 * The FP32 direct-sum keys and GH_TUNE_WALK_F64EVAL change results only by rounding;
 * GH_TUNE_WALK_GROUP changes which nodes the FP32 walk accepts (never fewer than the
 * reference's opening test accepts for any target); GH_TUNE_DIRECT_F64_TPW and
 * GH_TUNE_RUN_GRAPH do not change results, nor do GH_TUNE_SORT_DIRECT and
 * GH_TUNE_TREE_SORT (each of their forms gives the same order); GH_TUNE_WALK_PREFIX keeps
 * every target's accepted nodes and changes only their order, so the FP32 rounding of
 * the sums; GH_TUNE_ALLOC_LIMIT_MB only decides whether a call may allocate. */
#define GH_TUNE_DIRECT_SELF 0
#define GH_TUNE_DIRECT_IPT 1
#define GH_TUNE_WALK_GROUP 2
#define GH_TUNE_RUN_GRAPH 3
#define GH_TUNE_WALK_F64EVAL 4
#define GH_TUNE_ALLOC_LIMIT_MB 5
#define GH_TUNE_DIRECT_F64_TPW 6
#define GH_TUNE_SYM_BAND 7
#define GH_TUNE_SORT_DIRECT 8
#define GH_TUNE_TREE_SORT 9
#define GH_TUNE_WALK_PREFIX 10
#define GH_TUNE_COUNT 11
int gh_set_tuning(gh_ctx *ctx, int key, int value);

/* FP64 diagnostics (no reference counterpart): kinetic energy sum 0.5 m v^2
 * [Msun (km/s)^2], potential energy -sum_{i<j} m_i m_j / sqrt(r^2 + eps^2)
 * [Msun^2/kpc, G = 1] and momentum sum m v [Msun km/s] (3 doubles). */
int gh_energy(gh_ctx *ctx, const double *pos, const double *vel, const double *mass,
              int64_t np, double eps, double *kinetic, double *potential,
              double *momentum);

/* gh_energy on device pointers that allocates nothing: no hipMalloc, no context (no
 * reference counterpart; for callers that cannot allocate, such as code under HIP graph
 * capture or serving loops that forbid it).  Runs on the current device and on stream
 * (NULL: the default stream); out5 (device, 5 doubles) receives kinetic, potential and
 * the 3 momentum components, bitwise gh_energy's.  Its per-block partial sums live in a
 * fixed device array of the library, 640 KiB per device: np is at most 4194304
 * (GH_EARG beyond).  Calls that overlap on one device race on that array. */
int gh_energy_dev_noalloc(const double *d_pos, const double *d_vel, const double *d_mass,
                          int64_t np, double eps, double *d_out5, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GRAVHOP_H */
