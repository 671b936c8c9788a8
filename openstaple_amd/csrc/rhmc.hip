// One RHMC update: UPDATE_SOLOACC_UNOSTEP_VERSATILE (OpenAcc/update_versatile.c:44-580)
// for the build with STOUT_FERMIONS and a double-precision molecular dynamics.
//
// The steps and formulas are the reference's: backup of the links, Gaussian
// momenta and pseudofermions phi, stouting, power-method bounds and rescaling
// of the first-inversion approximation, initial action, heat-bath
// chi = r_fi(M^+M) phi, rescaled MD approximation, the 2MN trajectory,
// stouting, bounds and the last-inversion approximation, phi' = r_li(M^+M) chi,
// final action and the Metropolis test
//
//   dS = -(-S_g,in + S_p,in + |phi|^2) + (-S_g,fin + S_p,fin + <chi, phi'>)
//
// accepted if dS < 0 or u < exp(-dS).  Everything runs on the device through the
// library's own entry points; the only host work is the bookkeeping and the
// random draws' counters.
//
// Random numbers: momenta, phi and the power-method starts come from the
// library's counter-based generator, one stream per field and call, from the
// seed set by ostp_set_rng_seed; the Metropolis uniform comes from a callback
// (ostp_set_metropolis_uniform) or, by default, from the same seed.
#include "ostp_blas.h"

#include <vector>
#include <cmath>
#include <cstdio>
#include <cstring>

extern "C" {
// OpenAcc/alloc_vars.c / alloc_settings.h / fermion_parameters.h globals that
// the update reads; null until the host (or a test) points them at its fields.
// A host that defines them itself takes precedence.
su3_soa* conf_acc_bkp;
su3_soa* aux_conf_acc;
su3_soa* auxbis_conf_acc;
su3_soa* gstout_conf_acc_arr;
su3_soa* gconf_as_fermionmatrix;
thmat_soa* momenta;
tamat_soa* ipdot_acc;
vec3_soa* ferm_chi_acc;
vec3_soa* ferm_phi_acc;
vec3_soa* ferm_shiftmulti_acc;
vec3_soa* kloc_r;
vec3_soa* kloc_h;
vec3_soa* kloc_s;
vec3_soa* kloc_p;
vec3_soa* k_p_shiftferm;
dcomplex_soa* local_sums;
double_soa* d_local_sums;
ferm_param* fermions_parameters;
alloc_settings alloc_info;
}

static_assert(sizeof(alloc_settings) == 44, "alloc_settings layout (alloc_settings.h:4-20)");

namespace ostp {
// solvers.hip: the rational solves of several sources on one operator together,
// and the recombination of a rational function's shifted solutions
bool multishift_fp64_many(const void* u, ferm_param* const* pars, const RationalApprox* const* ras,
                          const void* const* in, void* const* out, int n, double res, int max_cg,
                          int convergence_importance);
bool recombine(const void* xs, const void* in, void* out, const RationalApprox* ra);
bool find_max_eigenvalues_many(const void* u, ferm_param* const* pars, void* const* r, void* const* h,
                               void* const* p, int n, double* maxv);
}  // namespace ostp

namespace {

using namespace ostp;

uint64_t g_seed = 0x2545F4914F6CDD1DULL;
uint32_t g_stream = 0;
ostp_uniform_fn g_uniform = nullptr;
void* g_uniform_ctx = nullptr;
uint64_t g_uniform_count = 0;
ostp_rhmc_info g_info;

// splitmix64 (Steele, Lea, Flood, OOPSLA 2014) of seed and a counter: the
// library's default Metropolis uniform in [0, 1)
double default_uniform() {
    uint64_t z = g_seed + 0x9E3779B97F4A7C15ULL * (++g_uniform_count);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

char* at(void* base, int i, size_t bytes) { return (char*)base + (size_t)i * bytes; }

struct DevScratch {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t need) {
        if (bytes >= need) return p;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (!hip_ok(hipMalloc(&p, need), "rhmc scratch")) return nullptr;
        bytes = need;
        return p;
    }
};
DevScratch g_conf_f, g_shift_f, g_loc_f, g_xs_many, g_pm_many;

// gauge action S = beta/3 (c0 sum Re tr P + c1 sum Re tr R) as the reference
// sums it (update_versatile.c:215-221)
bool gauge_action(const void* u, double* s) {
    const double beta_by_three = act_params.beta * 0.33333333333333333333333;
    const double c0 = stout_c_zero();
    double p = 0.0, r = 0.0;
    if (!gauge_action_sum(u, 0, &p)) return false;
    *s = c0 * beta_by_three * p;
    if (stout_tlsm()) {
        if (!gauge_action_sum(u, 1, &r)) return false;
        *s += -0.25 * 0.33333333333333333333333 * beta_by_three * r;
    }
    return true;
}

}  // namespace

extern "C" {

void ostp_set_rng_seed(uint64_t seed, uint32_t first_stream) {
    g_seed = seed;
    g_stream = first_stream;
    g_uniform_count = 0;
}

void ostp_set_metropolis_uniform(ostp_uniform_fn fn, void* ctx) {
    g_uniform = fn;
    g_uniform_ctx = ctx;
}

int ostp_rhmc_last_info(ostp_rhmc_info* out) {
    if (!out) return 1;
    *out = g_info;
    return 0;
}

int UPDATE_SOLOACC_UNOSTEP_VERSATILE(su3_soa* tconf_acc, double res_metro, double res_md, int id_iter, int acc,
                                     int metro, int max_cg) {
    const char* nm = "UPDATE_SOLOACC_UNOSTEP_VERSATILE";
    if (!require_init(nm)) return acc;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), vb = L.vec3_bytes();
    const int nflav = alloc_info.NDiffFlavs;
    const int steps = ostp_stout_steps();
    memset(&g_info, 0, sizeof g_info);
    g_info.iteration = id_iter + 1;
    if (md_parameters.singlePrecMD) {
        fail("%s: the single-precision molecular dynamics is not supported", nm);
        return acc;
    }
    if (act_params.topo_action == 1) {
        fail("%s: the topological action is not supported", nm);
        return acc;
    }
    if (nflav > 0 && (!fermions_parameters || !ferm_phi_acc || !ferm_chi_acc || !ferm_shiftmulti_acc ||
                      !k_p_shiftferm || !kloc_r || !kloc_h || !kloc_s || !kloc_p)) {
        fail("%s: fermion fields (fermions_parameters, ferm_*_acc, kloc_*, k_p_shiftferm) are not set", nm);
        return acc;
    }
    if (!momenta || !ipdot_acc || !aux_conf_acc || !auxbis_conf_acc || (metro && !conf_acc_bkp) ||
        (steps > 0 && nflav > 0 && !gstout_conf_acc_arr)) {
        fail("%s: gauge fields (momenta, ipdot_acc, aux_conf_acc, auxbis_conf_acc, conf_acc_bkp, "
             "gstout_conf_acc_arr) are not set",
             nm);
        return acc;
    }
    void* du = resolve(tconf_acc, gb, nm);
    void* dbkp = metro ? resolve(conf_acc_bkp, gb, nm) : nullptr;
    if (!du || (metro && !dbkp)) return acc;
    hipStream_t s = ctx().stream;
    const unsigned long e0 = error_count();
    auto failed = [&]() { return error_count() != e0; };

    // 1. backup; 2. momenta and pseudofermions
    if (metro && !hip_ok(hipMemcpyAsync(dbkp, du, gb, hipMemcpyDeviceToDevice, s), nm)) return acc;
    ostp_generate_Momenta_gauss(momenta, g_seed, g_stream++);
    for (int f = 0; f < nflav; f++)
        for (int ips = 0; ips < fermions_parameters[f].number_of_ps; ips++)
            ostp_generate_vec3_soa_gauss(
                (vec3_soa*)at(ferm_phi_acc, fermions_parameters[f].index_of_the_first_ps + ips, vb), g_seed,
                g_stream++);
    if (failed()) return acc;

    // the Dirac operator's links: the last stout level, or the links themselves
    auto stout_links = [&]() {
        if (steps > 0 && nflav > 0) {
            stout_wrapper(tconf_acc, gstout_conf_acc_arr, 0);
            gconf_as_fermionmatrix = (su3_soa*)at(gstout_conf_acc_arr, steps - 1, gb);
        } else {
            gconf_as_fermionmatrix = tconf_acc;
        }
        return !failed();
    };
    // bounds of M^+M and the rescaled approximation `which` (0 fi, 1 md, 2 li)
    double minmax[64][2];
    if (nflav > 64) {
        fail("%s: %d flavours exceed 64", nm, nflav);
        return acc;
    }
    auto bounds = [&](int f) {
        ostp_generate_vec3_soa_gauss(kloc_p, g_seed, g_stream++);
        ostp_generate_vec3_soa_gauss(kloc_s, g_seed, g_stream++);
        find_min_max_eigenvalue_soloopenacc(gconf_as_fermionmatrix, &fermions_parameters[f], kloc_r, kloc_h, kloc_p,
                                            kloc_s, minmax[f]);
        return !failed();
    };
    // the bounds of every flavour: the start vectors drawn as bounds(f) draws
    // them, then the power iterations of all flavours together where their
    // phases agree (one by one otherwise)
    auto bounds_all = [&]() {
        if (nflav < 2 || nflav > 4) {
            for (int f = 0; f < nflav; f++)
                if (!bounds(f)) return false;
            return true;
        }
        const unsigned long mark = error_count();
        char* w = (char*)g_pm_many.get((size_t)3 * nflav * vb);
        if (!w) {
            (void)hipGetLastError();
            forgive_errors(mark);
            for (int f = 0; f < nflav; f++)
                if (!bounds(f)) return false;
            return true;
        }
        void *r[4], *h[4], *pv[4];
        ferm_param* fps[4];
        for (int f = 0; f < nflav; f++) {
            pv[f] = w + (size_t)(3 * f) * vb;
            r[f] = w + (size_t)(3 * f + 1) * vb;
            h[f] = w + (size_t)(3 * f + 2) * vb;
            fps[f] = &fermions_parameters[f];
            ostp_generate_vec3_soa_gauss((vec3_soa*)pv[f], g_seed, g_stream++);
            ostp_generate_vec3_soa_gauss(kloc_s, g_seed, g_stream++);
        }
        if (failed()) return false;
        const void* du_fm = resolve(gconf_as_fermionmatrix, 8 * L.su3_bytes(), nm);
        if (!du_fm) return false;
        double maxv[4];
        const bool together = find_max_eigenvalues_many(du_fm, fps, r, h, pv, nflav, maxv);
        if (!together && failed()) return false;
        for (int f = 0; f < nflav; f++) {
            minmax[f][0] = fermions_parameters[f].ferm_mass * fermions_parameters[f].ferm_mass;
            minmax[f][1] = together ? maxv[f]
                                    : ker_find_max_eigenvalue_openacc(gconf_as_fermionmatrix, &fermions_parameters[f],
                                                                      (vec3_soa*)r[f], (vec3_soa*)h[f],
                                                                      (vec3_soa*)pv[f]);
        }
        return !failed();
    };
    auto rescale = [&](int f, int which) {
        ferm_param& fp = fermions_parameters[f];
        RationalApprox* mother = which == 0 ? &fp.approx_fi_mother : which == 1 ? &fp.approx_md_mother
                                                                                 : &fp.approx_li_mother;
        RationalApprox* ra = which == 0 ? &fp.approx_fi : which == 1 ? &fp.approx_md : &fp.approx_li;
        rescale_rational_approximation(mother, ra, minmax[f]);
        return !failed();
    };
    // inverter package (setup_inverter_package_dp / _sp, inverter_package.c):
    // nshifts = alloc_info.maxApproxOrder, the length of k_p_shiftferm
    const int max_order = alloc_info.maxApproxOrder;
    for (int f = 0; f < nflav; f++) {
        const ferm_param& fp = fermions_parameters[f];
        if (fp.approx_fi_mother.approx_order > max_order || fp.approx_md_mother.approx_order > max_order ||
            fp.approx_li_mother.approx_order > max_order) {
            fail("%s: flavour %d has an approximation of order above alloc_info.maxApproxOrder = %d", nm, f,
                 max_order);
            return acc;
        }
    }
    auto package = [&](inverter_package& ip) {
        memset(&ip, 0, sizeof ip);
        ip.u = gconf_as_fermionmatrix;
        ip.ferm_shift_temp = k_p_shiftferm;
        ip.nshifts = max_order;
        ip.loc_r = kloc_r;
        ip.loc_h = kloc_h;
        ip.loc_s = kloc_s;
        ip.loc_p = kloc_p;
        if (inverter_tricks.singlePInvAccelMultiInv) {
            const size_t vf = L.vec3f_bytes();
            void* cf = g_conf_f.get(8 * L.su3f_bytes());
            void* sf = g_shift_f.get((size_t)(max_order > 0 ? max_order : 1) * vf);
            char* lf = (char*)g_loc_f.get(5 * vf);
            if (!cf || !sf || !lf) return false;
            ostp_convert_double_to_float_gauge(gconf_as_fermionmatrix, (su3_soa_f*)cf);
            ip.u_f = (const su3_soa_f*)cf;
            ip.ferm_shift_temp_f = (vec3_soa_f*)sf;
            ip.loc_r_f = (vec3_soa_f*)lf;
            ip.loc_h_f = (vec3_soa_f*)(lf + vf);
            ip.loc_s_f = (vec3_soa_f*)(lf + 2 * vf);
            ip.loc_p_f = (vec3_soa_f*)(lf + 3 * vf);
            ip.out_f = (vec3_soa_f*)(lf + 4 * vf);
        }
        return !failed();
    };
    // y_ps = r(M^+M) x_ps for every pseudofermion of every flavour
    auto apply_rational = [&](int which, vec3_soa* xs, vec3_soa* ys, double res) {
        inverter_package ip;
        if (!package(ip)) return false;
        // every source's rational solve together where they share the operator
        // (inverter_multishift_wrapper's FP64 path, same phases), then each y = r(M) x
        {
            std::vector<ferm_param*> fps;
            std::vector<const RationalApprox*> ras;
            std::vector<const void*> ins;
            std::vector<void*> outs;
            std::vector<int> idx;
            for (int f = 0; f < nflav; f++) {
                ferm_param& fp = fermions_parameters[f];
                for (int ips = 0; ips < fp.number_of_ps; ips++) {
                    fps.push_back(&fp);
                    ras.push_back(which == 0 ? &fp.approx_fi : &fp.approx_li);
                    idx.push_back(fp.index_of_the_first_ps + ips);
                }
            }
            const int n = (int)fps.size();
            const size_t per = (size_t)(max_order > 0 ? max_order : 1) * vb;
            const unsigned long mark = error_count();
            char* xs_all = n >= 2 ? (char*)g_xs_many.get((size_t)n * per) : nullptr;
            if (n >= 2 && !xs_all) {
                (void)hipGetLastError();
                forgive_errors(mark);
            }
            if (xs_all) {
                const void* du_fm = resolve(ip.u, 8 * L.su3_bytes(), nm);
                if (!du_fm) return false;
                for (int k = 0; k < n; k++) {
                    const void* dx = resolve(at(xs, idx[k], vb), vb, nm);
                    if (!dx) return false;
                    ins.push_back(dx);
                    outs.push_back(xs_all + (size_t)k * per);
                }
                if (multishift_fp64_many(du_fm, fps.data(), ras.data(), ins.data(), outs.data(), n, res, max_cg,
                                         CONVERGENCE_CRITICAL)) {
                    for (int k = 0; k < n; k++) {
                        void* dy = resolve(at(ys, idx[k], vb), vb, nm);
                        if (!dy || !recombine(outs[k], ins[k], dy, ras[k])) return false;
                    }
                    // the last source's shifted solutions stay in ferm_shiftmulti_acc, as after the loop below
                    const int order = ras[n - 1]->approx_order;
                    void* dsm = resolve(ferm_shiftmulti_acc, per, nm);
                    if (!dsm || (order > 0 && !hip_ok(hipMemcpyAsync(dsm, outs[n - 1], (size_t)order * vb,
                                                                     hipMemcpyDeviceToDevice, ctx().stream),
                                                      nm)))
                        return false;
                    return !failed();
                }
                if (failed()) return false;
            }
        }
        for (int f = 0; f < nflav; f++) {
            ferm_param& fp = fermions_parameters[f];
            RationalApprox* ra = which == 0 ? &fp.approx_fi : &fp.approx_li;
            for (int ips = 0; ips < fp.number_of_ps; ips++) {
                const int k = fp.index_of_the_first_ps + ips;
                vec3_soa* x = (vec3_soa*)at(xs, k, vb);
                vec3_soa* y = (vec3_soa*)at(ys, k, vb);
                inverter_multishift_wrapper(ip, &fp, ra, ferm_shiftmulti_acc, x, res, max_cg, CONVERGENCE_CRITICAL);
                recombine_shifted_vec3_to_vec3(ferm_shiftmulti_acc, x, y, ra);
                if (failed()) return false;
            }
        }
        return true;
    };
    auto actions = [&](double* sg, double* sp, vec3_soa* a, vec3_soa* b, double* sf) {
        if (!gauge_action(du, sg)) return false;
        const void* dm = resolve(momenta, (size_t)8 * 64 * L.sizeh, nm);
        if (!dm || !momenta_action(dm, 0, 8, sp)) return false;
        *sf = 0.0;
        for (int f = 0; f < nflav; f++)
            for (int ips = 0; ips < fermions_parameters[f].number_of_ps; ips++) {
                const int k = fermions_parameters[f].index_of_the_first_ps + ips;
                *sf += real_scal_prod_global((const vec3_soa*)at(a, k, vb), (const vec3_soa*)at(b, k, vb));
            }
        return !failed();
    };

    // 3.-5. stout, bounds, first-inversion approximation, initial action
    if (!stout_links()) return acc;
    if (!bounds_all()) return acc;
    for (int f = 0; f < nflav; f++)
        if (!rescale(f, 0)) return acc;
    if (metro && !actions(&g_info.action_in, &g_info.action_mom_in, ferm_phi_acc, ferm_phi_acc, &g_info.action_ferm_in))
        return acc;
    // 6. heat bath chi = r_fi phi; 7. MD approximation
    multishift_invert_iterations = 0;
    if (!apply_rational(0, ferm_phi_acc, ferm_chi_acc, res_metro)) return acc;
    for (int f = 0; f < nflav; f++)
        if (!rescale(f, 1)) return acc;
    // 8. the trajectory
    const int iters_before = multishift_invert_iterations;
    {
        inverter_package ip;
        if (!package(ip)) return acc;
        multistep_2MN_SOLOOPENACC(ipdot_acc, tconf_acc, gstout_conf_acc_arr, auxbis_conf_acc, aux_conf_acc,
                                  fermions_parameters, nflav, ferm_chi_acc, ferm_shiftmulti_acc, ip, momenta,
                                  local_sums, res_md, max_cg);
        if (failed()) return acc;
    }
    g_info.md_multishift_iterations = multishift_invert_iterations - iters_before;
    // 9. stout again
    if (!stout_links()) return acc;
    if (metro) {
        // 10. bounds and last-inversion approximation; phi' = r_li chi
        if (!bounds_all()) return acc;
        for (int f = 0; f < nflav; f++)
            if (!rescale(f, 2)) return acc;
        if (!apply_rational(2, ferm_chi_acc, ferm_phi_acc, res_metro)) return acc;
        // 11. final action and the Metropolis test
        if (!actions(&g_info.action_fin, &g_info.action_mom_fin, ferm_chi_acc, ferm_phi_acc,
                     &g_info.action_ferm_fin))
            return acc;
        const double dS = -(-g_info.action_in + g_info.action_mom_in + g_info.action_ferm_in) +
                          (-g_info.action_fin + g_info.action_mom_fin + g_info.action_ferm_fin);
        g_info.delta_S = dS;
        int accepted;
        if (dS < 0) {
            accepted = 1;
            g_info.p1 = 1.0;
            g_info.p2 = -1.0;
        } else {
            g_info.p1 = exp(-dS);
            g_info.p2 = g_uniform ? g_uniform(g_uniform_ctx) : default_uniform();
            accepted = g_info.p2 < g_info.p1 ? 1 : 0;
        }
        g_info.accepted = accepted;
        if (accepted) {
            acc++;
            if (!hip_ok(hipMemcpyAsync(dbkp, du, gb, hipMemcpyDeviceToDevice, s), nm)) return acc;
        } else {
            if (!hip_ok(hipMemcpyAsync(du, dbkp, gb, hipMemcpyDeviceToDevice, s), nm)) return acc;
        }
        if (verbosity_lv > 0)
            printf("iterazione %i:  DELTA_ACTION = %.18lf. ---> [acc/iter] = [%i/%i]\n", g_info.iteration, dS, acc,
                   g_info.iteration);
    } else {
        g_info.accepted = 1;
        acc++;
    }
    hip_ok(hipStreamSynchronize(s), nm);
    return acc;
}

}  // extern "C"
