/* Sanitizer driver for the CPU restatement (test infrastructure only).
 *
 * Compiled together with jbgrav_oracle.c under -fsanitize=address,undefined by
 * tests/test_oracle_sanitize.py, it runs every entry point of the restatement on small
 * seeded inputs and on the edge cases the tree has to handle: one particle, a pair
 * closer than 21 and 42 octant levels, coincident particles (an error, not a hang),
 * targets outside the root box, theta = 0 and a huge theta.  It prints one line per
 * case and returns nonzero on an unexpected return code; the sanitizers report
 * memory and undefined-behaviour errors themselves.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "jbgrav_oracle.c"

static unsigned long long rng_state = 2026;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(rng_state >> 11) / 9007199254740992.0;
}

static int expect(const char *what, long long got, long long want)
{
    printf("%-32s rc %lld\n", what, got);
    if (got != want) {
        fprintf(stderr, "%s: return code %lld, expected %lld\n", what, got, want);
        return 1;
    }
    return 0;
}

int main(void)
{
    int bad = 0;
    enum { N = 300, NF = 40 };
    double *pos = malloc(sizeof(double) * 3 * N), *vel = malloc(sizeof(double) * 3 * N);
    double *mass = malloc(sizeof(double) * N), *acc = malloc(sizeof(double) * 3 * N);
    double *fpos = malloc(sizeof(double) * 3 * NF), *facc = malloc(sizeof(double) * 3 * NF);
    int64_t *cnt = malloc(sizeof(int64_t) * N);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] = urand() - 0.5;
            vel[3 * i + k] = urand() - 0.5;
        }
        mass[i] = 1.0 + urand();
    }
    /* a pair closer than 42 levels below the root box, and a third one nearby */
    for (int k = 0; k < 3; k++) {
        pos[3 * 1 + k] = pos[k] + 1e-14 * (k + 1);
        pos[3 * 2 + k] = pos[k] + 3e-9 * (k + 1);
    }
    for (int i = 0; i < NF; i++)
        for (int k = 0; k < 3; k++) fpos[3 * i + k] = 4.0 * (urand() - 0.5);

    bad |= expect("direct", orc_direct(pos, mass, N, 0.01, acc, 1), ORC_OK);
    bad |= expect("direct_position", orc_direct_position(pos, mass, N, fpos, NF, 0.0, facc, 1),
                  ORC_OK);
    bad |= expect("tree theta 0.7", orc_tree(pos, mass, N, NULL, 0, 0.01, 0.7, acc, cnt, 1),
                  ORC_OK);
    bad |= expect("tree theta 0", orc_tree(pos, mass, N, NULL, 0, 0.01, 0.0, acc, cnt, 1),
                  ORC_OK);
    bad |= expect("tree theta 1e30", orc_tree(pos, mass, N, NULL, 0, 0.01, 1e30, acc, cnt, 1),
                  ORC_OK);
    bad |= expect("tree_position", orc_tree(pos, mass, N, fpos, NF, 0.01, 0.7, facc, NULL, 1),
                  ORC_OK);
    const int64_t nn = orc_tree_dump(pos, mass, N, 0.01, NULL, 0);
    printf("%-32s nodes %lld\n", "tree_dump (count)", (long long)nn);
    if (nn <= N) bad = 1;
    orc_node_rec *nodes = malloc(sizeof(orc_node_rec) * (size_t)(nn > 0 ? nn : 1));
    bad |= expect("tree_dump (records)", orc_tree_dump(pos, mass, N, 0.01, nodes, nn), nn);
    bad |= expect("tree one particle", orc_tree(pos, mass, 1, NULL, 0, 0.0, 0.7, acc, cnt, 1),
                  ORC_OK);
    bad |= expect("dkd 5 steps", orc_dkd_direct(pos, vel, mass, N, 0.001, 0.01, 5,
                                                1.022712165045695e-3, 4.398600412921223e-9,
                                                NULL, NULL, 1),
                  ORC_OK);
    double ke, pe, mom[3];
    bad |= expect("energy", orc_energy(pos, vel, mass, N, 0.01, &ke, &pe, mom, 1), ORC_OK);
    /* coincident particles: the reference recurses forever, the restatement errs */
    for (int k = 0; k < 3; k++) pos[3 * 5 + k] = pos[3 * 4 + k];
    bad |= expect("tree coincident", orc_tree(pos, mass, N, NULL, 0, 0.01, 0.7, acc, cnt, 1),
                  ORC_ECOINCIDENT);
    free(nodes);
    free(pos);
    free(vel);
    free(mass);
    free(acc);
    free(fpos);
    free(facc);
    free(cnt);
    printf("%s\n", bad ? "FAILED" : "ok");
    return bad;
}
