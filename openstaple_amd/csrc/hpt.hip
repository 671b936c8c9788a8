// Parallel-tempering (Hasenbusch defect) bookkeeping of the PAR_TEMP build.
//
//  init_k                        OpenAcc/HPT_utilities.c:30-310
//  manage_replica_swaps labels   OpenAcc/HPT_utilities.c:722-1003
//  metro_SWAP_worldmaster        OpenAcc/HPT_utilities.c:704-719
//  get_index_of_pbc_replica      Include/rep_info.c:8-14
//  label initialisation          OpenAcc/main.c:341-342
//
// K (the per-link defect coefficient, the double plane after row 2 of each
// su3_soa) enters only the gauge staples and action, never Dslash; the
// library writes it on the device, one thread per half-site and parity.
// The label manager is host code: it permutes small int arrays exactly as
// the reference's world master does, with the replica actions and the
// uniform draws supplied by the caller (all replicas live in one process
// here, so the MPI broadcasts and reductions of the reference collapse).
#include "ostp_blas.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace ostp {

struct KGeom {
    int nd[4];
    int map[4];          // logical direction of physical x, y, z, t
    int lo[4], hi[4];    // defect box in physical coordinates: lo <= x < hi
    int axis_mapped;
    int sizeh;
};

__global__ __launch_bounds__(kBlock) void k_init_k(double* __restrict__ conf, size_t su3_stride_d, KGeom g,
                                                   double c_r) {
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= 2L * g.sizeh) return;
    const int parity = (int)(gid / g.sizeh), idxh = (int)(gid - (long)parity * g.sizeh);
    const int nd0h = g.nd[0] >> 1;
    int t = idxh / nd0h;
    const int hd0 = idxh - t * nd0h;
    int t2 = t / g.nd[1];
    int d[4];
    d[1] = t - t2 * g.nd[1];
    d[3] = t2 / g.nd[2];
    d[2] = t2 - d[3] * g.nd[2];
    d[0] = 2 * hd0 + ((d[1] + d[2] + d[3] + parity) & 1);
    bool on = true;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int x = d[g.map[p]];
        on = on && x >= g.lo[p] && x < g.hi[p];
    }
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        // K plane of link array 2*mu+parity: 18 doubles (9 complex planes) after its start
        double* K = conf + (size_t)(2 * mu + parity) * su3_stride_d + (size_t)18 * g.sizeh;
        K[idxh] = (on && mu == g.axis_mapped) ? c_r : 1.0;
    }
}

// physical perpendicular directions (HPT_utilities.c:34)
static const int kPerp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

static KGeom k_geom(int def_axis, const int* def_vec) {
    const Lattice& L = ctx().lat;
    KGeom g;
    for (int i = 0; i < 4; i++) {
        g.nd[i] = L.nd[i];
        g.map[i] = L.map[i];
    }
    g.axis_mapped = L.map[def_axis];
    int ext[4];
    ext[def_axis] = 1;
    for (int i = 0; i < 3; i++) ext[kPerp[def_axis][i]] = def_vec[i];
    for (int p = 0; p < 4; p++) {
        // lower corner: last site along the defect axis, 0 elsewhere
        const int gn = L.nd[L.map[p]];
        g.lo[p] = (L.map[p] == g.axis_mapped) ? gn - 1 : 0;
        g.hi[p] = g.lo[p] + ext[p];
    }
    g.sizeh = (int)L.sizeh;
    return g;
}

// Delta-S swap ranges of one rank holding the whole lattice
// (HPT_utilities.c:42-71, 143-176, 181-241).  Writes the TLSM or the Wilson
// layout of defect_info.
static void fill_defect_info(const KGeom& g, void* def_out, bool tlsm) {
    int mn[4][4], mx[4][4], mnT[2][4][4], mxT[2][4][4];
    const int ndm1[4] = {g.nd[0] - 1, g.nd[1] - 1, g.nd[2] - 1, g.nd[3] - 1};
    for (int nu = 0; nu < 4; nu++)
        for (int i = 0; i < 4; i++) {
            mx[nu][i] = 0;
            mn[nu][i] = ndm1[i];
            for (int j = 0; j < 2; j++) {
                mxT[j][nu][i] = 0;
                mnT[j][nu][i] = ndm1[i];
            }
        }
    int count = 0;
    int d[4];
    for (d[3] = 0; d[3] < g.nd[3]; d[3]++)
        for (d[2] = 0; d[2] < g.nd[2]; d[2]++)
            for (d[1] = 0; d[1] < g.nd[1]; d[1]++)
                for (d[0] = 0; d[0] < g.nd[0]; d[0]++) {
                    bool on = true;
                    for (int p = 0; p < 4; p++) {
                        const int x = d[g.map[p]];
                        on = on && x >= g.lo[p] && x < g.hi[p];
                    }
                    if (!on) continue;
                    count++;
                    for (int nu = 0; nu < 4; nu++)
                        for (int i = 0; i < 4; i++) {
                            if (d[i] < mn[nu][i]) {
                                mn[nu][i] = d[i];
                                for (int j = 0; j < 2; j++) mnT[j][nu][i] = d[i];
                            }
                            if (d[i] > mx[nu][i]) {
                                mx[nu][i] = d[i];
                                for (int j = 0; j < 2; j++) mxT[j][nu][i] = d[i];
                            }
                        }
                }
    if (verbosity_lv > 0) printf("MPI%02d: num of defect sites on the bulk = %d\n", 0, count);
    if (count == 0) {
        for (int nu = 0; nu < 4; nu++)
            for (int i = 0; i < 4; i++) {
                mn[nu][i] = mx[nu][i] = 0;
                for (int j = 0; j < 2; j++) mnT[j][nu][i] = mxT[j][nu][i] = 0;
            }
    } else {
        for (int nu = 0; nu < 4; nu++)
            for (int i = 0; i < 4; i++) {
                mx[nu][i] += 1;  // used as an exclusive loop bound
                for (int j = 0; j < 2; j++) mxT[j][nu][i] += 1;
            }
        // 1x1 plaquettes on (mu = defect axis, nu): include x+nu in the defect
        for (int k = 0; k < 4; k++) mn[k][k] -= 1;
        // 1x2 rectangles: x+nu and x+2nu; 2x1: x+nu and x+mu
        for (int k = 0; k < 4; k++) {
            mnT[0][k][k] -= 2;
            mnT[1][k][k] -= 1;
        }
        for (int k = 0; k < 4; k++) mnT[1][k][g.axis_mapped] -= 1;
    }
    if (tlsm) {
        defect_info* o = (defect_info*)def_out;
        for (int nu = 0; nu < 4; nu++)
            for (int i = 0; i < 4; i++) {
                o->defect_swap_min[nu][i] = mn[nu][i];
                o->defect_swap_max[nu][i] = mx[nu][i];
                for (int j = 0; j < 2; j++) {
                    o->defect_swap_min_TLSM[j][nu][i] = mnT[j][nu][i];
                    o->defect_swap_max_TLSM[j][nu][i] = mxT[j][nu][i];
                }
            }
        o->def_axis_mapped = g.axis_mapped;
        for (int i = 0; i < 3; i++) o->def_mapped_perp_dir[i] = kPerp[g.axis_mapped][i];
    } else {
        ostp_defect_info_wilson* o = (ostp_defect_info_wilson*)def_out;
        for (int nu = 0; nu < 4; nu++)
            for (int i = 0; i < 4; i++) {
                o->defect_swap_min[nu][i] = mn[nu][i];
                o->defect_swap_max[nu][i] = mx[nu][i];
            }
        o->def_axis_mapped = g.axis_mapped;
        for (int i = 0; i < 3; i++) o->def_mapped_perp_dir[i] = kPerp[g.axis_mapped][i];
    }
}

// ------------------------------------------------------------ label manager
static int twin_label(int lab, int swap_order, int n) {
    const int twin = ((swap_order == 0 && lab % 2 == 0) || (swap_order == 1 && lab % 2 == 1)) ? lab + 1 : lab - 1;
    return (twin < 0 || twin >= n) ? -1 : twin;
}

static int metropolis(double ds, double u) {
    if (ds < 0) return 1;
    return u < std::exp(-ds) ? 1 : 0;
}

// exchange the replicas that carry labels a and b
static void swap_labels(int* label, int n, int a, int b) {
    int ii = -1, jj = -1;
    for (int idx = 0; idx < n; idx++) {
        if (label[idx] == a) ii = idx;
        else if (label[idx] == b) jj = idx;
    }
    if (ii < 0 || jj < 0) return;
    const int t = label[ii];
    label[ii] = label[jj];
    label[jj] = t;
}

}  // namespace ostp

using namespace ostp;

extern "C" {

void init_k(su3_soa* conf, double c_r, int def_axis, int* def_vec, defect_info* def, int defect_info_config) {
    const char* nm = "init_k";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    if (!L.par_temp) {
        fail("%s: the lattice was initialised without the K plane (ostp_init(..., par_temp = 1))", nm);
        return;
    }
    if (def_axis < 0 || def_axis > 3 || !def_vec) {
        fail("%s: invalid defect axis %d or null extents", nm, def_axis);
        return;
    }
    double* d = (double*)resolve(conf, 8 * L.su3_bytes(), nm);
    if (!d) return;
    const KGeom g = k_geom(def_axis, def_vec);
    const size_t stride_d = L.su3_bytes() / sizeof(double);
    hipLaunchKernelGGL(k_init_k, dim3(blocks_for(2 * L.sizeh)), dim3(kBlock), 0, ctx().stream, d, stride_d, g, c_r);
    if (!hip_ok(hipGetLastError(), nm)) return;
    if (defect_info_config == 0 && def) fill_defect_info(g, def, gauge_action_tlsm());
    hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

void ostp_pt_init_labels(rep_info* hpt_params) {
    if (!hpt_params || !hpt_params->label) return;
    for (int r = 0; r < hpt_params->replicas_total_number; r++) hpt_params->label[r] = r;
}

int ostp_pt_index_of_label0(const rep_info* hpt_params) {
    if (!hpt_params || !hpt_params->label) return -1;
    for (int r = 0; r < hpt_params->replicas_total_number; r++)
        if (hpt_params->label[r] == 0) return r;
    return -1;
}

int ostp_pt_metropolis(double delta_s_swap, double u) { return metropolis(delta_s_swap, u); }

int ostp_pt_twin_label(int lab, int swap_order, int nreplicas) { return twin_label(lab, swap_order, nreplicas); }

void ostp_pt_manage_replica_swaps(rep_info* hp, ostp_pt_action_fn action, ostp_pt_uniform_fn uniform, void* user,
                                  int* swap_num, int* all_swap_vet, int* acceptance_vet) {
    if (!hp || !hp->label || !hp->cr_vec || !action || !uniform || !swap_num || !all_swap_vet || !acceptance_vet) {
        fail("ostp_pt_manage_replica_swaps: null argument");
        return;
    }
    const int n = hp->replicas_total_number;
    int* label = hp->label;
    std::vector<double> cr(n), s_prev(n), s_next(n);
    auto current_cr = [&]() {
        for (int r = 0; r < n; r++) cr[r] = hp->cr_vec[label[r]];
    };
    int swap_order = (uniform(user) < 0.5) ? 1 : 0;
    if (verbosity_lv > 1) printf("swap order: %d\n", swap_order);

    if (hp->is_evenodd) {
        for (int itr = 0; itr < 2; itr++) {
            swap_order = (itr + swap_order) % 2;
            std::fill(s_prev.begin(), s_prev.end(), 0.0);
            current_cr();
            action(cr.data(), label, s_prev.data(), user);
            // each replica takes its twin's coefficient for the trial
            for (int r = 0; r < n; r++) {
                const int twin = twin_label(label[r], swap_order, n);
                if (twin >= 0) cr[r] = hp->cr_vec[twin];
            }
            std::fill(s_next.begin(), s_next.end(), 0.0);
            action(cr.data(), label, s_next.data(), user);
            std::vector<int> accepted(n, 0);
            for (int lbs = swap_order; lbs < n; lbs += 2) {
                const int twin = twin_label(lbs, swap_order, n);
                if (twin < 0) continue;
                const int lo = twin > lbs ? lbs : twin, hi = twin > lbs ? twin : lbs;
                const double ds = -(s_next[lo] + s_next[hi] - s_prev[lo] - s_prev[hi]);
                const double u = ds < 0 ? 0.0 : uniform(user);
                accepted[lo] = accepted[hi] = metropolis(ds, u);
                if (verbosity_lv > 1) printf("DELTA_S_SWAP(lab1=%d,lab2=%d):%.15lg\n", lo, hi, ds);
                if (accepted[lo]) swap_labels(label, n, lo, hi);
            }
            *swap_num += 1;
            for (int lbs = swap_order; lbs < n; lbs += 2) {
                const int twin = twin_label(lbs, swap_order, n);
                if (twin < 0) continue;
                const int fixed = twin < lbs ? twin : lbs;
                all_swap_vet[fixed]++;
                if (accepted[lbs] == 1) acceptance_vet[fixed]++;
            }
        }
    } else {
        for (int i = 0; i < n - 1; i++) {
            std::fill(s_prev.begin(), s_prev.end(), 0.0);
            current_cr();
            action(cr.data(), label, s_prev.data(), user);
            const int lab1 = swap_order == 1 ? i : n - i - 1;
            const int lab2 = swap_order == 1 ? i + 1 : n - i - 2;
            if (verbosity_lv > 1) printf("proposing swap %d %d\n", lab1, lab2);
            for (int r = 0; r < n; r++) {
                if (label[r] == lab1) cr[r] = hp->cr_vec[lab2];
                if (label[r] == lab2) cr[r] = hp->cr_vec[lab1];
            }
            std::fill(s_next.begin(), s_next.end(), 0.0);
            action(cr.data(), label, s_next.data(), user);
            const double ds = -(s_next[lab1] + s_next[lab2] - s_prev[lab1] - s_prev[lab2]);
            const double u = ds < 0 ? 0.0 : uniform(user);
            const int acc = metropolis(ds, u);
            if (verbosity_lv > 1) printf("DELTA_S_SWAP(lab1=%d,lab2=%d):%.15lg\n", lab1, lab2, ds);
            if (acc) swap_labels(label, n, lab1, lab2);
            *swap_num += 1;
            const int fixed = swap_order == 1 ? lab1 : lab2;
            all_swap_vet[fixed]++;
            if (acc) acceptance_vet[fixed]++;
        }
    }
}

}  // extern "C"
