// Fermion measurements and the operators they apply:
//
//  dM_dmu_eo/oe{0..3}, d2M_dmu2_eo/oe{0..3}  Meas/baryon_number_utilities.c:14-935
//  acc_Deo_wf / acc_Doe_wf                    OpenAcc/field_times_fermion_matrix.c:77-236
//  idphase_dbz                                Meas/magnetic_susceptibility_utilities.c:13-56
//  ostp_fermion_measures                      Meas/ferm_meas.c:230-460 (the per-flavour body
//                                             of fermion_measures, noise vectors batched)
//
// All operators are one kernel, k_hop: for the output parity, over a mask of
// directions, 0.5 (U_mu(x) w e^{i theta} in(x+mu) + s U_mu(x-mu)^+ conj(w
// e^{i theta}) in(x-mu)) with s = +1 for dM/dmu and s = -1 for d2M/dmu2 and
// the field-weighted Dslash (w = field_re + i field_im at the link's site,
// 1 without a field).  These run a few times per configuration, so the
// kernel is a plain one thread per output half-site.
#include "ostp_blas.h"
#include "ostp_su3.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace ostp {

struct HopGeom {
    int nd[4];
    int sizeh;
};

template <int EO>
__global__ __launch_bounds__(kBlock) void k_hop(const double2* __restrict__ u, size_t su3_stride,
                                                double2* __restrict__ out, const double2* __restrict__ in,
                                                const double* __restrict__ ph, const double* __restrict__ wre,
                                                const double* __restrict__ wim, int mask, double s, HopGeom g) {
    const int idxh = blockIdx.x * kBlock + (int)threadIdx.x;
    if (idxh >= g.sizeh) return;
    const int sh = g.sizeh;
    const int nd0h = g.nd[0] >> 1;
    int d[4];
    int t = idxh / nd0h;
    const int hd0 = idxh - t * nd0h;
    int t2 = t / g.nd[1];
    d[1] = t - t2 * g.nd[1];
    d[3] = t2 / g.nd[2];
    d[2] = t2 - d[3] * g.nd[2];
    d[0] = 2 * hd0 + ((d[1] + d[2] + d[3] + EO) & 1);
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    auto snum = [&](const int e[4]) { return (e[0] + g.nd[0] * (e[1] + g.nd[1] * (e[2] + g.nd[2] * e[3]))) >> 1; };
    for (int mu = 0; mu < 4; mu++) {
        if (!(mask & (1 << mu))) continue;
        int e[4] = {d[0], d[1], d[2], d[3]};
        // forward: U_mu(x) (array 2mu + EO) times in(x + mu)
        e[mu] = d[mu] + 1 == g.nd[mu] ? 0 : d[mu] + 1;
        const int ip = snum(e);
        e[mu] = d[mu] == 0 ? g.nd[mu] - 1 : d[mu] - 1;
        const int im = snum(e);
        for (int back = 0; back < 2; back++) {
            const int l = 2 * mu + (back ? 1 - EO : EO);
            const int il = back ? im : idxh;  // the link's half-site
            const int iv = back ? im : ip;    // the spinor's half-site
            double sn, cs;
            sincos(ph[(size_t)l * sh + il], &sn, &cs);
            double2 f = make_double2(cs, sn);
            if (wre) f = cmul(f, make_double2(wre[(size_t)l * sh + il], wim[(size_t)l * sh + il]));
            if (back) f = cconj(f);
            const double2* U = (const double2*)((const char*)u + (size_t)l * su3_stride);
            Su3 m;
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) m.c[r][c] = U[(size_t)(3 * r + c) * sh + il];
            rebuild3row(m);
            double2 v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = cmul(in[(size_t)c * sh + iv], f);
            const double sg = back ? s : 1.0;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                double2 w = make_double2(0, 0);
#pragma unroll
                for (int c = 0; c < 3; c++) w = back ? cfmac(m.c[c][r], v[c], w) : cfma(m.c[r][c], v[c], w);
                acc[r] = make_double2(acc[r].x + sg * w.x, acc[r].y + sg * w.y);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[(size_t)c * sh + idxh] = make_double2(0.5 * acc[c].x, 0.5 * acc[c].y);
}

// im = 2 pi (a - b) ; re = 0   (phase_diff_in_place, mult_u1_phases, set_double_soa_to_zero)
__global__ __launch_bounds__(kBlock) void k_dphase(double* __restrict__ re, double* __restrict__ im,
                                                   const double* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    im[i] = (im[i] - b[i]) * (2 * 3.14159265358979323846);
    re[i] = 0.0;
}

static bool launch_hop(int eo, int mask, double s, const void* u, void* out, const void* in, const void* ph,
                       const void* wre, const void* wim) {
    const Lattice& L = ctx().lat;
    HopGeom g;
    for (int i = 0; i < 4; i++) g.nd[i] = L.nd[i];
    g.sizeh = (int)L.sizeh;
    auto k = eo ? k_hop<1> : k_hop<0>;
    hipLaunchKernelGGL(k, dim3(blocks_for(L.sizeh)), dim3(kBlock), 0, ctx().stream, (const double2*)u,
                       L.su3_bytes(), (double2*)out, (const double2*)in, (const double*)ph, (const double*)wre,
                       (const double*)wim, mask, s, g);
    return hip_ok(hipGetLastError(), "k_hop");
}

// operator on pointers the caller resolved; waits for the result
static bool hop_entry(const char* nm, int eo, int mask, double s, const su3_soa* u, vec3_soa* out,
                      const vec3_soa* in, const double_soa* ph, const double_soa* wre, const double_soa* wim) {
    if (!require_init(nm)) return false;
    const Lattice& L = ctx().lat;
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    void* dout = resolve(out, L.vec3_bytes(), nm);
    const void* din = resolve(in, L.vec3_bytes(), nm);
    const void* dph = resolve(ph, 8 * L.dsoa_bytes(), nm);
    const void* dre = wre ? resolve(wre, 8 * L.dsoa_bytes(), nm) : nullptr;
    const void* dim = wim ? resolve(wim, 8 * L.dsoa_bytes(), nm) : nullptr;
    if (!du || !dout || !din || !dph || (wre && !dre) || (wim && !dim)) return false;
    if ((dre == nullptr) != (dim == nullptr)) return fail("%s: field_re and field_im go together", nm);
    return launch_hop(eo, mask, s, du, dout, din, dph, dre, dim) &&
           hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

}  // namespace ostp

using namespace ostp;

extern "C" {

#define OSTP_DERIV(name, EO, MU, S)                                                                       \
    void name(const su3_soa* u, vec3_soa* out, const vec3_soa* in, const double_soa* backfield) {         \
        hop_entry(#name, EO, 1 << (MU), S, u, out, in, backfield, nullptr, nullptr);                    \
    }
OSTP_DERIV(dM_dmu_eo0, 0, 0, 1.0)
OSTP_DERIV(dM_dmu_eo1, 0, 1, 1.0)
OSTP_DERIV(dM_dmu_eo2, 0, 2, 1.0)
OSTP_DERIV(dM_dmu_eo3, 0, 3, 1.0)
OSTP_DERIV(dM_dmu_oe0, 1, 0, 1.0)
OSTP_DERIV(dM_dmu_oe1, 1, 1, 1.0)
OSTP_DERIV(dM_dmu_oe2, 1, 2, 1.0)
OSTP_DERIV(dM_dmu_oe3, 1, 3, 1.0)
OSTP_DERIV(d2M_dmu2_eo0, 0, 0, -1.0)
OSTP_DERIV(d2M_dmu2_eo1, 0, 1, -1.0)
OSTP_DERIV(d2M_dmu2_eo2, 0, 2, -1.0)
OSTP_DERIV(d2M_dmu2_eo3, 0, 3, -1.0)
OSTP_DERIV(d2M_dmu2_oe0, 1, 0, -1.0)
OSTP_DERIV(d2M_dmu2_oe1, 1, 1, -1.0)
OSTP_DERIV(d2M_dmu2_oe2, 1, 2, -1.0)
OSTP_DERIV(d2M_dmu2_oe3, 1, 3, -1.0)
#undef OSTP_DERIV

// baryon_number_utilities.h:13-31: indexed by the machine direction of time
// (geom_par.tmap at the call sites)
void (*dM_dmu_eo[4])(const su3_soa*, vec3_soa*, const vec3_soa*, const double_soa*) = {dM_dmu_eo0, dM_dmu_eo1,
                                                                                      dM_dmu_eo2, dM_dmu_eo3};
void (*dM_dmu_oe[4])(const su3_soa*, vec3_soa*, const vec3_soa*, const double_soa*) = {dM_dmu_oe0, dM_dmu_oe1,
                                                                                      dM_dmu_oe2, dM_dmu_oe3};
void (*d2M_dmu2_eo[4])(const su3_soa*, vec3_soa*, const vec3_soa*, const double_soa*) = {
    d2M_dmu2_eo0, d2M_dmu2_eo1, d2M_dmu2_eo2, d2M_dmu2_eo3};
void (*d2M_dmu2_oe[4])(const su3_soa*, vec3_soa*, const vec3_soa*, const double_soa*) = {
    d2M_dmu2_oe0, d2M_dmu2_oe1, d2M_dmu2_oe2, d2M_dmu2_oe3};

void acc_Deo_wf(const su3_soa* u, vec3_soa* out, const vec3_soa* in, const double_soa* phases,
                const double_soa* field_re, const double_soa* field_im) {
    hop_entry("acc_Deo_wf", 0, 15, -1.0, u, out, in, phases, field_re, field_im);
}
void acc_Doe_wf(const su3_soa* u, vec3_soa* out, const vec3_soa* in, const double_soa* phases,
                const double_soa* field_re, const double_soa* field_im) {
    hop_entry("acc_Doe_wf", 1, 15, -1.0, u, out, in, phases, field_re, field_im);
}

// idphase_dbz: i d(phase)/d(b_z) of the fermion's U(1) phases, from the
// unbounded, un-normalised phases at b_z = 1 and b_z = 0 (mu = 0): re = 0,
// im = 2 pi (phi(b_z = 1) - phi(0)); staggered and boundary signs cancel
void idphase_dbz(double_soa* idphi_dbz_re, double_soa* idphi_dbz_im, ferm_param* fpar) {
    const char* nm = "idphase_dbz";
    if (!require_init(nm) || !fpar) return;
    const Lattice& L = ctx().lat;
    const size_t pb = 8 * L.dsoa_bytes();
    double* re = (double*)resolve(idphi_dbz_re, pb, nm);
    double* im = (double*)resolve(idphi_dbz_im, pb, nm);
    if (!re || !im) return;
    if (!u1_phases_unbounded(im, bf_param{0, 0, 0, 0, 0, 1}, 0.0, fpar->ferm_charge) ||
        !u1_phases_unbounded(re, bf_param{0, 0, 0, 0, 0, 0}, 0.0, fpar->ferm_charge))
        return;
    const long n = 8 * L.sizeh;
    hipLaunchKernelGGL(k_dphase, dim3(blocks_for(n)), dim3(kBlock), 0, ctx().stream, re, im, (const double*)re, n);
    if (hip_ok(hipGetLastError(), nm)) hip_ok(hipStreamSynchronize(ctx().stream), nm);
}

// The measurements of fermion_measures for one flavour: nvec Z2 noise
// vectors (streams 2k and 2k+1 of seed on even and odd sites); all
// inversions of a kind go through the batched solver together.
int ostp_fermion_measures(const su3_soa* u, const su3_soa_f* u_f, ferm_param* pars, int nvec, int n_chiral_susc,
                          int n_qn_susc, uint64_t seed, double res, int max_cg, int half_inner, double* out,
                          int* cg_return) {
    const char* nm = "ostp_fermion_measures";
    if (!require_init(nm)) return 0;
    const unsigned long e0 = error_count();
    if (nvec < 1 || !pars || !out || n_chiral_susc < 0 || n_qn_susc < 0) {
        fail("%s: invalid arguments", nm);
        return 0;
    }
    const Lattice& L = ctx().lat;
    const size_t vb = L.vec3_bytes();
    const int tdir = L.map[3];
    const double m = pars->ferm_mass;
    const double factor = pars->degeneracy * 0.25 / ((double)L.nd[0] * L.nd[1] * L.nd[2] * L.nd[3]);
    const bool mag = pars->mag_re && pars->mag_im;
    enum { kRe, kRo, kXe, kXo, kPhi, kBe, kBo, kYe, kYo, kNv };
    char* buf = nullptr;
    if (!hip_ok(hipMalloc(&buf, (size_t)kNv * vb * nvec), "measure vectors")) return 0;
    auto V = [&](int k, int j) { return (vec3_soa*)(buf + ((size_t)k * kNv + j) * vb); };
    int cg_total = 0, unconverged = 0;
    // (y_e, y_o) = M^-1 (x_e, x_o) for the vectors in ks (ferm_meas.c:50-71)
    auto invert = [&](const std::vector<int>& ks, int ie, int io, int oe, int oo) {
        if (ks.empty()) return true;
        std::vector<ostp_batch_system> sys(ks.size());
        for (size_t q = 0; q < ks.size(); q++) {
            const int k = ks[q];
            acc_Deo(u, V(k, kPhi), V(k, io), pars->phases);
            combine_in1_x_fact1_minus_in2_back_into_in2(V(k, ie), m, V(k, kPhi));
            set_vec3_soa_to_zero(V(k, oe));
            ostp_batch_system& b = sys[q];
            for (int d = 0; d < 4; d++) b.nd[d] = L.nd[d];
            b.par_temp = L.par_temp;
            b.u = u;
            b.u_f = u_f;
            b.phases = pars->phases;
            b.phases_f = pars->phases_f;
            b.in = V(k, kPhi);
            b.out = V(k, oe);
            b.mass = m;
        }
        int cg = 0;
        const int conv = batched_solve_internal((int)ks.size(), sys.data(), res, max_cg, 0.1, 16, &cg, nullptr,
                                                nullptr, half_inner != 0);
        cg_total += cg;
        unconverged += (int)ks.size() - conv;
        for (int k : ks) {
            acc_Doe(u, V(k, kPhi), V(k, oe), pars->phases);
            combine_in1_minus_in2_allxfact(V(k, io), V(k, kPhi), 1.0 / m, V(k, oo));
        }
        return error_count() == e0;
    };
    // <rnd, (a_e, a_o)> times sign and factor into out[12 k + 2 q]
    auto dot = [&](int k, int q, int ae, int ao, double sign) {
        double re_e = 0, im_e = 0, re_o = 0, im_o = 0;
        if (!reduce_to_host(2, V(k, kRe), V(k, ae), 0, &re_e, &im_e) ||
            !reduce_to_host(2, V(k, kRo), V(k, ao), 0, &re_o, &im_o))
            return false;
        out[12 * k + 2 * q] = sign * (re_e + re_o) * factor;
        out[12 * k + 2 * q + 1] = sign * (im_e + im_o) * factor;
        return true;
    };
    for (int i = 0; i < 12 * nvec; i++) out[i] = NAN;
    bool ok = true;
    std::vector<int> all, chi, qn;
    for (int k = 0; k < nvec; k++) {
        ostp_generate_vec3_soa_z2noise(V(k, kRe), seed, 2u * (uint32_t)k);
        ostp_generate_vec3_soa_z2noise(V(k, kRo), seed, 2u * (uint32_t)k + 1u);
        all.push_back(k);
        if (k < n_chiral_susc) chi.push_back(k);
        if (k < n_qn_susc) qn.push_back(k);
    }
    // first inversion and the single-inversion measures
    ok = invert(all, kRe, kRo, kXe, kXo);
    for (int k = 0; ok && k < nvec; k++) {
        ok = dot(k, 0, kXe, kXo, 1.0);  // chiral condensate
        dM_dmu_eo[tdir](u, V(k, kBe), V(k, kXo), pars->phases);
        dM_dmu_oe[tdir](u, V(k, kBo), V(k, kXe), pars->phases);
        ok = ok && dot(k, 1, kBe, kBo, 1.0);  // quark number
        if (ok && mag) {
            acc_Deo_wf(u, V(k, kYe), V(k, kXo), pars->phases, pars->mag_re, pars->mag_im);
            acc_Doe_wf(u, V(k, kYo), V(k, kXe), pars->phases, pars->mag_re, pars->mag_im);
            ok = dot(k, 2, kYe, kYo, -1.0);  // magnetisation
        }
    }
    // chiral susceptibility: -<rnd, M^-2 rnd>
    ok = ok && invert(chi, kXe, kXo, kYe, kYo);
    for (int k : chi)
        if (ok) ok = dot(k, 3, kYe, kYo, -1.0);
    // quark number susceptibilities: <rnd, d2M/dmu2 M^-1 rnd>, -<rnd, (dM/dmu M^-1)^2 rnd>
    for (int k : qn) {
        if (!ok) break;
        d2M_dmu2_eo[tdir](u, V(k, kYe), V(k, kXo), pars->phases);
        d2M_dmu2_oe[tdir](u, V(k, kYo), V(k, kXe), pars->phases);
        ok = dot(k, 4, kYe, kYo, 1.0);
    }
    ok = ok && invert(qn, kBe, kBo, kYe, kYo);
    for (int k : qn) {
        if (!ok) break;
        dM_dmu_eo[tdir](u, V(k, kBe), V(k, kYo), pars->phases);
        dM_dmu_oe[tdir](u, V(k, kBo), V(k, kYe), pars->phases);
        ok = dot(k, 5, kBe, kBo, -1.0);
    }
    (void)hipFree(buf);
    if (cg_return) *cg_return = cg_total;
    return ok ? (unconverged < nvec ? nvec - unconverged : 0) : 0;
}

}  // extern "C"
