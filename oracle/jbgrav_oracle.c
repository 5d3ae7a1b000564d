/*
 * oracle/jbgrav_oracle.c -- CPU restatement of GravHopper's force engine and leapfrog.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check and
 * the cpu_baseline leg of bench.py load it through oracle/oracle.py.  The product
 * library (gravhopper_amd, libgravhop.so) never links, loads or calls it.
 *
 * Every function restates an algorithm of the reference (jbailinua/gravhopper v1.2.0)
 * with the reference's FP64 operation order, so that, built with gcc on x86-64
 * without FMA contraction, its results are bit-for-bit those of the reference:
 *
 *   orc_direct            gravhopper/_jbgrav.c:140-193  (directsummation_workhorse)
 *   orc_direct_position   gravhopper/_jbgrav.c:299-353  (directsummation_position_workhorse)
 *   orc_tree              gravhopper/_jbgrav.c:360-558, 737-806 (octree build + walk)
 *   orc_tree_dump         canonical pre-order serialisation of the same octree
 *   orc_dkd_direct        gravhopper/gravhopper.py:405-416 (DKD leapfrog) in code units
 *   orc_energy            no reference counterpart: FP64 kinetic/potential energy
 *
 * Memory: the reference stores N^2 scratch arrays; the restatement computes the same
 * per-pair quantities on the fly (O(N) memory), which does not change a single bit
 * because each stored value is used exactly as it was computed.
 *
 * Threads: the outer per-target loops may run under OpenMP (nthreads > 1).  Each
 * target's sum is still computed by one thread in the reference order, so the result
 * does not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ENOMEM -1
#define ORC_ECOINCIDENT -2
#define ORC_EDEPTH -3
#define ORC_EARG -4

static void orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
    (void)nthreads;
#endif
}

/* ---------------------------------------------------------------------------------
 * Direct summation, gravhopper/_jbgrav.c:140-193.
 *
 * The reference fills dpos[i][j][k] and invdpos3[i][j] for i<j (:155-171) and then
 * sums, for every i and every k, over j ascending with j != i (:174-185):
 *   diff   = x_i[k] - x_j[k]                    (:159, always the lower index first)
 *   dpos[i][j][k] = -diff, dpos[j][i][k] = diff  (:161-162)
 *   d2     = ((0.0 + diff0^2) + diff1^2) + diff2^2   (:157,:160,:163)
 *   inv    = 1.0 / (d2 + eps*eps) / sqrt(d2 + eps*eps)   (:152,:165-166)
 *   acc[i][k] = 0.0; acc[i][k] += (m_j * dpos[i][j][k]) * inv   (:177,:181-182)
 * ------------------------------------------------------------------------------- */
int orc_direct(const double *pos, const double *mass, int64_t np, double eps,
               double *acc, int nthreads)
{
    const double eps2 = eps * eps;
    orc_set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < np; i++) {
        double f[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < np; j++) {
            if (j == i) continue;
            const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
            double dp[3], d2 = 0.0;
            for (int k = 0; k < 3; k++) {
                const double diff = pos[3 * lo + k] - pos[3 * hi + k];
                /* row lo stores -diff, row hi stores +diff */
                dp[k] = (i == lo) ? -diff : diff;
                d2 += diff * diff;
            }
            const double d = d2 + eps2;
            const double inv = 1.0 / d / sqrt(d);
            for (int k = 0; k < 3; k++) f[k] += mass[j] * dp[k] * inv;
        }
        for (int k = 0; k < 3; k++) acc[3 * i + k] = f[k];
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------------------
 * Field at arbitrary targets, gravhopper/_jbgrav.c:299-353.  No index exclusion;
 * a pair with d2 + eps^2 == 0 contributes exactly zero (:327-330).
 * ------------------------------------------------------------------------------- */
int orc_direct_position(const double *pos, const double *mass, int64_t np,
                        const double *fpos, int64_t nf, double eps, double *acc,
                        int nthreads)
{
    const double eps2 = eps * eps;
    orc_set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nf; i++) {
        double f[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < np; j++) {
            double dp[3], d2 = 0.0;
            for (int k = 0; k < 3; k++) {
                const double diff = fpos[3 * i + k] - pos[3 * j + k];
                dp[k] = -diff;
                d2 += diff * diff;
            }
            const double d = d2 + eps2;
            const double inv = (d == 0.0) ? 0.0 : 1.0 / d / sqrt(d);
            for (int k = 0; k < 3; k++) f[k] += mass[j] * dp[k] * inv;
        }
        for (int k = 0; k < 3; k++) acc[3 * i + k] = f[k];
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------------------
 * Barnes-Hut octree, gravhopper/_jbgrav.c:360-558 and 737-806.
 *
 * Nodes live in one growable arena and refer to each other by index; insertion is a
 * loop instead of the reference's recursion, but every node receives the same values
 * in the same order:
 *   - an empty node becomes a leaf holding one particle; mass = m, moment = m*x (:392-400)
 *   - a leaf that receives a second particle moves its particle into a new child and
 *     then treats the newcomer as an internal node would (:401-415)
 *   - an internal node adds the particle's mass and moment and descends (:416-436)
 *   - octant bit k = (x_k > c_k) strictly, branch = bit0 + 2 bit1 + 4 bit2 (:441-462)
 *   - child centre = c + (s * 0.5) * halfsize, s = +-1; child size = parent halfsize,
 *     child halfsize = 0.5 * child size (:360-384, :406, :409, :426, :428)
 * The reference recurses forever on coincident particles; the restatement reports
 * ORC_ECOINCIDENT instead, and ORC_EDEPTH when the centre recurrence stops separating
 * two distinct particles.
 * ------------------------------------------------------------------------------- */
typedef struct {
    double center[3];
    double size, halfsize, mass;
    double fm[3], com[3];
    int32_t child[8];
    int64_t leaf; /* particle index, or -1 */
    int8_t empty, comvalid;
    int16_t level;
} onode;

typedef struct {
    onode *n;
    int64_t count, cap;
    const double *pos, *mass;
} otree;

#define ORC_MAX_LEVEL 4000

static int64_t otree_new_node(otree *t, const double *center, double size, int level)
{
    if (t->count == t->cap) {
        int64_t ncap = t->cap ? 2 * t->cap : 1024;
        onode *nn = (onode *)realloc(t->n, (size_t)ncap * sizeof(onode));
        if (!nn) return -1;
        t->n = nn;
        t->cap = ncap;
    }
    onode *r = &t->n[t->count];
    r->size = size;
    r->halfsize = 0.5 * size;
    for (int i = 0; i < 3; i++) {
        r->center[i] = center[i];
        r->fm[i] = 0.0;
        r->com[i] = 0.0;
    }
    for (int i = 0; i < 8; i++) r->child[i] = -1;
    r->mass = 0.0;
    r->empty = 1;
    r->comvalid = 0;
    r->leaf = -1;
    r->level = (int16_t)level;
    return t->count++;
}

static int octant_of(const onode *nd, const double *x, int *sgn)
{
    int b = 0;
    for (int i = 0; i < 3; i++) {
        if (x[i] > nd->center[i]) {
            sgn[i] = 1;
            b += (1 << i);
        } else {
            sgn[i] = -1;
        }
    }
    return b;
}

static int64_t make_child(otree *t, int64_t node, int b, const int *sgn)
{
    double sub[3];
    onode *nd = &t->n[node];
    for (int i = 0; i < 3; i++) sub[i] = nd->center[i] + sgn[i] * 0.5 * nd->halfsize;
    const double hs = nd->halfsize;
    const int lvl = nd->level + 1;
    int64_t c = otree_new_node(t, sub, hs, lvl);
    if (c < 0) return -1;
    t->n[node].child[b] = (int32_t)c;
    return c;
}

static int otree_insert(otree *t, int64_t p)
{
    const double *x = &t->pos[3 * p];
    const double m = t->mass[p];
    int64_t node = 0;
    for (;;) {
        onode *nd = &t->n[node];
        if (nd->empty) {
            nd->empty = 0;
            nd->leaf = p;
            nd->mass = m;
            for (int i = 0; i < 3; i++) nd->fm[i] = m * x[i];
            return ORC_OK;
        }
        if (nd->leaf >= 0) {
            const int64_t a = nd->leaf;
            const double *xa = &t->pos[3 * a];
            if (xa[0] == x[0] && xa[1] == x[1] && xa[2] == x[2]) return ORC_ECOINCIDENT;
            if (nd->level >= ORC_MAX_LEVEL) return ORC_EDEPTH;
            int sgn[3];
            const int b = octant_of(nd, xa, sgn);
            const int64_t c = make_child(t, node, b, sgn);
            if (c < 0) return ORC_ENOMEM;
            nd = &t->n[node];
            onode *cn = &t->n[c];
            cn->empty = 0;
            cn->leaf = a;
            cn->mass = t->mass[a];
            for (int i = 0; i < 3; i++) cn->fm[i] = t->mass[a] * xa[i];
            nd->leaf = -1;
        }
        int sgn[3];
        const int b = octant_of(nd, x, sgn);
        int64_t c = nd->child[b];
        if (c < 0) {
            c = make_child(t, node, b, sgn);
            if (c < 0) return ORC_ENOMEM;
        }
        nd = &t->n[node];
        nd->mass += m;
        for (int i = 0; i < 3; i++) nd->fm[i] += m * x[i];
        node = c;
    }
}

/* gravhopper/_jbgrav.c:747-772: bounding box with the size quirk */
static void root_box(const double *pos, int64_t np, double eps, double *center,
                     double *size)
{
    double mn[3], mx[3];
    for (int i = 0; i < 3; i++) mn[i] = mx[i] = pos[i];
    for (int64_t i = 1; i < np; i++) {
        for (int j = 0; j < 3; j++) {
            const double q = pos[3 * i + j];
            if (q < mn[j]) mn[j] = q;
            if (q > mx[j]) mx[j] = q;
        }
    }
    double bs = mx[0] - mn[0] + eps;
    for (int i = 1; i < 3; i++)
        if ((mx[i] - mn[i]) > bs) bs = mx[i] - mn[i] + eps;
    for (int i = 0; i < 3; i++) center[i] = 0.5 * (mn[i] + mx[i]);
    *size = bs;
}

static int otree_build(otree *t, const double *pos, const double *mass, int64_t np,
                       double eps)
{
    memset(t, 0, sizeof(*t));
    t->pos = pos;
    t->mass = mass;
    double c[3], s;
    root_box(pos, np, eps, c, &s);
    if (otree_new_node(t, c, s, 0) < 0) return ORC_ENOMEM;
    for (int64_t i = 0; i < np; i++) {
        int rc = otree_insert(t, i);
        if (rc != ORC_OK) return rc;
    }
    return ORC_OK;
}

/* gravhopper/_jbgrav.c:467-483 */
static void otree_finalize(otree *t, onode *nd)
{
    if (!nd->comvalid) {
        if (nd->leaf >= 0) {
            for (int i = 0; i < 3; i++) nd->com[i] = t->pos[3 * nd->leaf + i];
        } else {
            for (int i = 0; i < 3; i++) nd->com[i] = nd->fm[i] / nd->mass;
        }
        nd->comvalid = 1;
    }
}

/* gravhopper/_jbgrav.c:487-541.  The COM of every node is finalised before the walk,
 * which gives the same values as the reference's lazy finalisation and lets several
 * threads walk one tree. */
static void otree_accel(const otree *t, int64_t node, const double *x, double eps,
                        double theta, double *force, int64_t *count)
{
    const onode *nd = &t->n[node];
    const double eps2 = eps * eps;
    double node_dist = 0.0;
    for (int i = 0; i < 3; i++)
        node_dist += (nd->center[i] - x[i]) * (nd->center[i] - x[i]);
    node_dist = sqrt(node_dist);
    if ((nd->leaf >= 0) || ((nd->size / node_dist) < theta)) {
        double dp[3], d2 = 0.0;
        for (int i = 0; i < 3; i++) {
            const double diff = nd->com[i] - x[i];
            dp[i] = diff;
            d2 += diff * diff;
        }
        const double d = d2 + eps2;
        const double inv = (d == 0.0) ? 0.0 : 1.0 / d / sqrt(d);
        for (int i = 0; i < 3; i++) force[i] = dp[i] * nd->mass * inv;
        (*count)++;
    } else {
        double bf[3];
        for (int i = 0; i < 3; i++) force[i] = 0.0;
        for (int j = 0; j < 8; j++) {
            if (nd->child[j] >= 0) {
                otree_accel(t, nd->child[j], x, eps, theta, bf, count);
                for (int i = 0; i < 3; i++) force[i] += bf[i];
            }
        }
    }
}

int orc_tree(const double *pos, const double *mass, int64_t np, const double *fpos,
             int64_t nf, double eps, double theta, double *acc, int64_t *ninteract,
             int nthreads)
{
    if (np < 1) return ORC_EARG;
    otree t;
    int rc = otree_build(&t, pos, mass, np, eps);
    if (rc != ORC_OK) {
        free(t.n);
        return rc;
    }
    for (int64_t k = 0; k < t.count; k++)
        if (!t.n[k].empty) otree_finalize(&t, &t.n[k]);
    orc_set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nf; i++) {
        double f[3];
        int64_t cnt = 0;
        otree_accel(&t, 0, &fpos[3 * i], eps, theta, f, &cnt);
        for (int k = 0; k < 3; k++) acc[3 * i + k] = f[k];
        if (ninteract) ninteract[i] = cnt;
    }
    free(t.n);
    return ORC_OK;
}

/* Canonical serialisation: depth-first pre-order with children visited in branch
 * order 0..7 (the order of gravoct_calc_accel's loop, _jbgrav.c:530-537). */
typedef struct {
    int32_t level;
    int32_t branch;  /* octant within the parent, -1 for the root */
    int64_t leaf;    /* particle index for a leaf, -1 for an internal node */
    double center[3];
    double size;
    double mass;
    double com[3];
} orc_node_rec;

static int64_t dump_rec(const otree *t, int64_t node, int branch, orc_node_rec *out,
                        int64_t cap, int64_t k)
{
    const onode *nd = &t->n[node];
    if (k < cap) {
        orc_node_rec *r = &out[k];
        r->level = nd->level;
        r->branch = branch;
        r->leaf = nd->leaf;
        for (int i = 0; i < 3; i++) {
            r->center[i] = nd->center[i];
            r->com[i] = nd->com[i];
        }
        r->size = nd->size;
        r->mass = nd->mass;
    }
    k++;
    for (int j = 0; j < 8; j++)
        if (nd->child[j] >= 0) k = dump_rec(t, nd->child[j], j, out, cap, k);
    return k;
}

/* Returns the number of nodes (records beyond cap are counted, not written), or a
 * negative error code. */
int64_t orc_tree_dump(const double *pos, const double *mass, int64_t np, double eps,
                      orc_node_rec *out, int64_t cap)
{
    if (np < 1) return ORC_EARG;
    otree t;
    int rc = otree_build(&t, pos, mass, np, eps);
    if (rc != ORC_OK) {
        free(t.n);
        return rc;
    }
    for (int64_t k = 0; k < t.count; k++)
        if (!t.n[k].empty) otree_finalize(&t, &t.n[k]);
    int64_t n = dump_rec(&t, 0, -1, out, cap, 0);
    free(t.n);
    return n;
}

/* ---------------------------------------------------------------------------------
 * DKD leapfrog, gravhopper/gravhopper.py:405-416, in code units (kpc, km/s, Msun,
 * Myr).  astropy turns the drift (km/s * Myr) into kpc by multiplying with
 * `drift` = 1.022712165045695e-3, and jbgrav.py:48 turns G*Msun/kpc^2 into km/s/Myr
 * by multiplying with `ascale` = 4.398600412921223e-9:
 *   xh  = x + ((0.5 * v) * dt) * drift                         (:409)
 *   a   = direct(xh) * ascale + 0.0      (zero external acceleration, :455-457)
 *   v'  = v + a * dt                                           (:414)
 *   x'  = xh + ((0.5 * v') * dt) * drift                       (:416)
 * pos/vel are (np,3) and are advanced in place.  hist_pos/hist_vel, when given, hold
 * (nsteps, np, 3) snapshots after every step.
 * ------------------------------------------------------------------------------- */
int orc_dkd_direct(double *pos, double *vel, const double *mass, int64_t np, double dt,
                   double eps, int64_t nsteps, double drift, double ascale,
                   double *hist_pos, double *hist_vel, int nthreads)
{
    double *xh = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    double *a = (double *)malloc(sizeof(double) * 3 * (size_t)np);
    if (!xh || !a) {
        free(xh);
        free(a);
        return ORC_ENOMEM;
    }
    for (int64_t s = 0; s < nsteps; s++) {
        for (int64_t k = 0; k < 3 * np; k++) xh[k] = pos[k] + 0.5 * vel[k] * dt * drift;
        if (np > 1) {
            orc_direct(xh, mass, np, eps, a, nthreads);
        } else {
            for (int64_t k = 0; k < 3 * np; k++) a[k] = 0.0;
        }
        for (int64_t k = 0; k < 3 * np; k++) {
            const double acc = a[k] * ascale + 0.0;
            vel[k] = vel[k] + acc * dt;
            pos[k] = xh[k] + 0.5 * vel[k] * dt * drift;
        }
        if (hist_pos) memcpy(hist_pos + 3 * np * s, pos, sizeof(double) * 3 * (size_t)np);
        if (hist_vel) memcpy(hist_vel + 3 * np * s, vel, sizeof(double) * 3 * (size_t)np);
    }
    free(xh);
    free(a);
    return ORC_OK;
}

/* ---------------------------------------------------------------------------------
 * Energy and momentum in FP64 (no reference counterpart).
 *   kinetic   = sum_i 0.5 m_i |v_i|^2                       [Msun (km/s)^2]
 *   potential = -sum_{i<j} m_i m_j / sqrt(|x_i-x_j|^2+eps^2) [Msun^2 / kpc, G = 1]
 *   momentum  = sum_i m_i v_i                               [Msun km/s]
 * The potential is the one whose gradient is orc_direct's force.
 * ------------------------------------------------------------------------------- */
int orc_energy(const double *pos, const double *vel, const double *mass, int64_t np,
               double eps, double *kinetic, double *potential, double *momentum,
               int nthreads)
{
    const double eps2 = eps * eps;
    double ke = 0.0, pe = 0.0, p[3] = {0.0, 0.0, 0.0};
    for (int64_t i = 0; i < np; i++) {
        const double *v = &vel[3 * i];
        ke += 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) p[k] += mass[i] * v[k];
    }
    orc_set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : pe)
    for (int64_t i = 0; i < np; i++) {
        double row = 0.0;
        for (int64_t j = i + 1; j < np; j++) {
            double d2 = 0.0;
            for (int k = 0; k < 3; k++) {
                const double diff = pos[3 * i + k] - pos[3 * j + k];
                d2 += diff * diff;
            }
            row += mass[j] / sqrt(d2 + eps2);
        }
        pe -= mass[i] * row;
    }
    *kinetic = ke;
    *potential = pe;
    for (int k = 0; k < 3; k++) momentum[k] = p[k];
    return ORC_OK;
}
