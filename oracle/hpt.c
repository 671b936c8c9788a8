/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Parallel-tempering bookkeeping of the PAR_TEMP build, single rank:
 *   or_init_k                  OpenAcc/HPT_utilities.c:30-310
 *   or_metro_swap              OpenAcc/HPT_utilities.c:704-719 (uniform draw passed in)
 *   or_manage_replica_swaps    OpenAcc/HPT_utilities.c:722-1003 (labels and counters;
 *                              actions and uniforms come from callbacks)
 * Pins: tests/test_oracle_hpt.py (brute-force defect box, closed-form
 * counts, hand-worked swap sequences, permutation invariants).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const int or_perp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

/* def_out (TLSM layout, 100 ints): min[4][4], max[4][4], minT[2][4][4],
 * maxT[2][4][4], def_axis_mapped, perp[3] */
void or_init_k(const or_geom *g, const int map[4], void *conf, double c_r, int def_axis, const int def_vec[3],
               int *def_out, int defect_info_config) {
    const long sizeh = g->sizeh;
    const int dam = map[def_axis];
    int nd[4], i, mu, nu, j;
    int (*mn)[4] = (int (*)[4])(def_out);
    int (*mx)[4] = (int (*)[4])(def_out + 16);
    int (*mnT)[4][4] = (int (*)[4][4])(def_out + 32);
    int (*mxT)[4][4] = (int (*)[4][4])(def_out + 64);
    for (i = 0; i < 4; i++) nd[i] = g->nd[i] - 1;
    if (defect_info_config == 0) {
        def_out[96] = dam;
        for (nu = 0; nu < 4; nu++) {
            for (i = 0; i < 3; i++) {
                mx[nu][i] = 0;
                mn[nu][i] = nd[i];
                for (j = 0; j < 2; j++) { mxT[j][nu][i] = 0; mnT[j][nu][i] = nd[i]; }
            }
            mx[nu][3] = 0;           /* D3_HALO = 0 on one rank */
            mn[nu][3] = nd[3];
            for (j = 0; j < 2; j++) { mxT[j][nu][3] = 0; mnT[j][nu][3] = nd[3]; }
        }
        for (i = 0; i < 3; i++) def_out[97 + i] = or_perp[dam][i];
    }
    /* physical extent of the defect: 1 along def_axis, def_vec on the others */
    int ext[4] = {0, 0, 0, 0};
    ext[def_axis] = 1;
    for (i = 0; i < 3; i++) ext[or_perp[def_axis][i]] = def_vec[i];
    const int gn[4] = {g->nd[map[0]], g->nd[map[1]], g->nd[map[2]], g->nd[map[3]]};
    int lo[4], hi[4];
    for (i = 0; i < 4; i++) {
        lo[i] = (map[i] == dam) ? gn[i] - 1 : 0;
        hi[i] = lo[i] + ext[i];
    }
    const size_t stride = or_su3_bytes(g);
    int count = 0, d[4];
    for (d[3] = 0; d[3] < g->nd[3]; d[3]++)
        for (d[2] = 0; d[2] < g->nd[2]; d[2]++)
            for (d[1] = 0; d[1] < g->nd[1]; d[1]++)
                for (d[0] = 0; d[0] < g->nd[0]; d[0]++) {
                    const int idxh = or_snum(g, d[0], d[1], d[2], d[3]);
                    const int x = d[map[0]], y = d[map[1]], z = d[map[2]], t = d[map[3]];
                    const int parity = (d[0] + d[1] + d[2] + d[3]) % 2;
                    for (mu = 0; mu < 4; mu++) {
                        double *K = (double *)((char *)conf + (2 * mu + parity) * stride + 144 * sizeh);
                        K[idxh] = 1.0;
                    }
                    const int on = x >= lo[0] && y >= lo[1] && z >= lo[2] && t >= lo[3] && x < hi[0] && y < hi[1] &&
                                   z < hi[2] && t < hi[3];
                    if (!on) continue;
                    if (defect_info_config == 0)
                        for (nu = 0; nu < 4; nu++)
                            for (i = 0; i < 4; i++) {
                                if (d[i] < mn[nu][i]) {
                                    mn[nu][i] = d[i];
                                    for (j = 0; j < 2; j++) mnT[j][nu][i] = d[i];
                                }
                                if (d[i] > mx[nu][i]) {
                                    mx[nu][i] = d[i];
                                    for (j = 0; j < 2; j++) mxT[j][nu][i] = d[i];
                                }
                            }
                    count++;
                    double *Kd = (double *)((char *)conf + (2 * dam + parity) * stride + 144 * sizeh);
                    Kd[idxh] = c_r;
                }
    if (defect_info_config != 0) return;
    if (count == 0) {
        for (nu = 0; nu < 4; nu++)
            for (i = 0; i < 4; i++) {
                mn[nu][i] = 0; mx[nu][i] = 0;
                for (j = 0; j < 2; j++) { mnT[j][nu][i] = 0; mxT[j][nu][i] = 0; }
            }
        return;
    }
    for (nu = 0; nu < 4; nu++)
        for (i = 0; i < 4; i++) {
            mx[nu][i] += 1;
            for (j = 0; j < 2; j++) mxT[j][nu][i] += 1;
        }
    mn[0][0] -= 1; mn[1][1] -= 1; mn[2][2] -= 1; mn[3][3] -= 1;     /* :206-209 */
    mnT[0][0][0] -= 2; mnT[0][1][1] -= 2; mnT[0][2][2] -= 2; mnT[0][3][3] -= 2;   /* :214-218 */
    mnT[1][0][0] -= 1; mnT[1][1][1] -= 1; mnT[1][2][2] -= 1; mnT[1][3][3] -= 1;   /* :221-224 */
    mnT[1][0][dam] -= 1; mnT[1][1][dam] -= 1; mnT[1][2][dam] -= 1; mnT[1][3][dam] -= 1; /* :226-229 */
}

int or_metro_swap(double delta_s, double u) {
    if (delta_s < 0) return 1;
    return (u < exp(-delta_s)) ? 1 : 0;
}

static void or_swap_labels(int n, int *label, int a, int b) {
    int ii = -1, jj = -1, idx;
    for (idx = 0; idx < n; ++idx) {
        if (label[idx] == a) ii = idx;
        else if (label[idx] == b) jj = idx;
    }
    int aux = label[ii];
    label[ii] = label[jj];
    label[jj] = aux;
}

static int or_twin(int lab, int order) {
    return ((order == 0 && lab % 2 == 0) || (order == 1 && lab % 2 == 1)) ? lab + 1 : lab - 1;
}

void or_manage_replica_swaps(int n, int is_evenodd, const double *cr_vec, int *label, or_pt_action_fn action,
                             or_pt_uniform_fn uniform, void *user, int *swap_num, int *all_swap_vet,
                             int *acceptance_vet) {
    double *S_prev = calloc(n, sizeof(double)), *S_next = calloc(n, sizeof(double));
    double *cr = calloc(n, sizeof(double));
    int *acc = calloc(n, sizeof(int));
    int r, lbs;
    int swap_order = (uniform(user) < 0.5) ? 1 : 0;
    if (is_evenodd) {
        for (int itr = 0; itr < 2; ++itr) {
            swap_order = (itr + swap_order) % 2;
            for (r = 0; r < n; r++) { S_prev[r] = 0.0; cr[r] = cr_vec[label[r]]; }
            action(cr, label, S_prev, user);
            for (r = 0; r < n; r++) {          /* every replica takes its twin's coefficient */
                int tw = or_twin(label[r], swap_order);
                if (!(tw < 0 || tw >= n)) cr[r] = cr_vec[tw];
            }
            for (r = 0; r < n; r++) S_next[r] = 0.0;
            action(cr, label, S_next, user);
            for (r = 0; r < n; r++) acc[r] = 0;
            for (lbs = swap_order; lbs < n; lbs += 2) {
                int tw = or_twin(lbs, swap_order);
                if (tw < 0 || tw >= n) { acc[lbs] = 0; continue; }
                int lo = tw > lbs ? lbs : tw, hi = tw > lbs ? tw : lbs;
                double ds = -(S_next[lo] + S_next[hi] - S_prev[lo] - S_prev[hi]);
                /* metro_SWAP_worldmaster draws only when Delta S >= 0 */
                acc[lo] = or_metro_swap(ds, ds < 0 ? 0.0 : uniform(user));
                acc[hi] = acc[lo];
                if (acc[lo]) or_swap_labels(n, label, lo, hi);
            }
            *swap_num += 1;
            for (lbs = swap_order; lbs < n; lbs += 2) {
                int tw = or_twin(lbs, swap_order);
                if (tw < 0 || tw >= n) continue;
                int fixed = tw < lbs ? tw : lbs;
                all_swap_vet[fixed]++;
                if (acc[lbs] == 1) acceptance_vet[fixed]++;
            }
        }
    } else {
        for (int ic = 0; ic < n - 1; ic++) {
            for (r = 0; r < n; r++) { S_prev[r] = 0.0; cr[r] = cr_vec[label[r]]; }
            action(cr, label, S_prev, user);
            int l1, l2;
            if (swap_order == 1) { l1 = ic; l2 = ic + 1; }
            else { l1 = n - ic - 1; l2 = n - ic - 2; }
            for (r = 0; r < n; r++) {
                if (l1 == label[r]) cr[r] = cr_vec[l2];
                if (l2 == label[r]) cr[r] = cr_vec[l1];
            }
            for (r = 0; r < n; r++) S_next[r] = 0.0;
            action(cr, label, S_next, user);
            double ds = -(S_next[l1] + S_next[l2] - S_prev[l1] - S_prev[l2]);
            int a = or_metro_swap(ds, ds < 0 ? 0.0 : uniform(user));
            if (a) or_swap_labels(n, label, l1, l2);
            *swap_num += 1;
            int fixed = (swap_order == 1) ? l1 : l2;
            all_swap_vet[fixed]++;
            if (a == 1) acceptance_vet[fixed]++;
        }
    }
    free(S_prev); free(S_next); free(cr); free(acc);
}
