// Fermion BLAS on SoA vectors (OpenAcc/fermionic_utilities.c:126-313,
// OpenAcc/float_double_conv.c) and the fused solver update kernels.
//
// Each kernel streams whole vec3 planes with one half-site per lane
// (coalesced 128-bit loads); reductions use the deterministic grid reduce of
// ostp_reduce.h so FP64 solver iteration counts are reproducible run to run.
#include "ostp_blas.h"

namespace ostp {

// ------------------------------------------------------------ elementwise ops
// out = in1*factor + in2 (combine_in1xfactor_plus_in2)
template <typename C>
__global__ __launch_bounds__(kBlock) void k_axpy(const C* __restrict__ in1, double factor, const C* __restrict__ in2,
                                                 C* __restrict__ out, long n3, SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const C a = in1[i], b = in2[i];
    out[i] = mkc<C>((double)a.x * factor + (double)b.x, (double)a.y * factor + (double)b.y);
}

template <typename C>
__global__ __launch_bounds__(kBlock) void k_sub(const C* __restrict__ in1, const C* __restrict__ in2,
                                                C* __restrict__ out, long n3) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const C a = in1[i], b = in2[i];
    out[i] = mkc<C>((double)a.x - (double)b.x, (double)a.y - (double)b.y);
}

// in2 = in1*m2 - in2 (combine_in1xferm_mass_minus_in2)
// out = fact * (in1 - in2)   (fermionic_utilities.c:403-419)
__global__ __launch_bounds__(kBlock) void k_sub_scale(const double2* __restrict__ in1, const double2* __restrict__ in2,
                                                      double fact, double2* __restrict__ out, long n3) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double2 a = in1[i], b = in2[i];
    out[i] = make_double2(fact * (a.x - b.x), fact * (a.y - b.y));
}

__global__ __launch_bounds__(kBlock) void k_mass_minus(const double2* __restrict__ in1, double m2,
                                                       double2* __restrict__ in2, long n3) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n3) return;
    const double2 a = in1[i], b = in2[i];
    in2[i] = make_double2(a.x * m2 - b.x, a.y * m2 - b.y);
}

template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_convert(const A* __restrict__ in, B* __restrict__ out, long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const A a = in[i];
    out[i] = mkc<B>((double)a.x, (double)a.y);
}

__global__ __launch_bounds__(kBlock) void k_convert_real(const double* __restrict__ in, float* __restrict__ out,
                                                         long n) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// --------------------------------------------------------------- reductions
// sum over sizeh half-sites of f(a(i), b(i)) over 3 colours
template <typename C, int KIND, int F>
__global__ __launch_bounds__(kBlock) void k_reduce(const C* __restrict__ a, const C* __restrict__ b, int sizeh,
                                                   SolverCtl ctl) {
    if (!pred_pass(ctl)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v0 = 0.0, v1 = 0.0;
    if (i < sizeh) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const C x = a[(long)c * sizeh + i];
            if (KIND == 0) {  // |a|^2
                v0 += (double)x.x * x.x + (double)x.y * x.y;
            } else {
                const C y = b[(long)c * sizeh + i];
                v0 += (double)x.x * y.x + (double)x.y * y.y;    // Re <a,b>
                if (KIND == 2) v1 += (double)x.x * y.y - (double)x.y * y.x;  // Im <a,b>
            }
        }
    }
    block_partial<(KIND == 2 ? 2 : 1)>(v0, v1, ctl);
}

// ---------------------------------------------------------------- launchers
static inline hipStream_t S() { return ctx().stream; }
static inline int nb(long n) { return blocks_for(n); }

template <typename C>
bool launch_axpy(const void* in1, double factor, const void* in2, void* out, const SolverCtl& ctl) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL((k_axpy<C>), dim3(nb(n3)), dim3(kBlock), 0, S(), (const C*)in1, factor, (const C*)in2,
                       (C*)out, n3, ctl);
    return hip_ok(hipGetLastError(), "k_axpy");
}
template bool launch_axpy<double2>(const void*, double, const void*, void*, const SolverCtl&);
template bool launch_axpy<float2>(const void*, double, const void*, void*, const SolverCtl&);

bool launch_sub(const void* in1, const void* in2, void* out) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL((k_sub<double2>), dim3(nb(n3)), dim3(kBlock), 0, S(), (const double2*)in1,
                       (const double2*)in2, (double2*)out, n3);
    return hip_ok(hipGetLastError(), "k_sub");
}

bool launch_mass_minus(const void* in1, double m2, void* in2) {
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL(k_mass_minus, dim3(nb(n3)), dim3(kBlock), 0, S(), (const double2*)in1, m2, (double2*)in2,
                       n3);
    return hip_ok(hipGetLastError(), "k_mass_minus");
}

bool launch_d2f(const void* d, void* f, long n) {
    hipLaunchKernelGGL((k_convert<double2, float2>), dim3(nb(n)), dim3(kBlock), 0, S(), (const double2*)d,
                       (float2*)f, n);
    return hip_ok(hipGetLastError(), "k_convert");
}
bool launch_f2d(const void* f, void* d, long n) {
    hipLaunchKernelGGL((k_convert<float2, double2>), dim3(nb(n)), dim3(kBlock), 0, S(), (const float2*)f,
                       (double2*)d, n);
    return hip_ok(hipGetLastError(), "k_convert");
}

template <typename C, int F>
bool launch_norm2(const void* a, const SolverCtl& ctl) {
    const int sizeh = (int)ctx().lat.sizeh;
    hipLaunchKernelGGL((k_reduce<C, 0, F>), dim3(nb(sizeh)), dim3(kBlock), 0, S(), (const C*)a, (const C*)nullptr,
                       sizeh, ctl);
    return hip_ok(hipGetLastError(), "k_reduce") && launch_finalize(F, ctl, nb(sizeh));
}
template bool launch_norm2<double2, kFinStore>(const void*, const SolverCtl&);
template bool launch_norm2<double2, kFinPower>(const void*, const SolverCtl&);
template bool launch_norm2<float2, kFinStore>(const void*, const SolverCtl&);

bool launch_dot(const void* a, const void* b, int complex_dot, const SolverCtl& ctl) {
    const int sizeh = (int)ctx().lat.sizeh;
    if (complex_dot)
        hipLaunchKernelGGL((k_reduce<double2, 2, kFinStore2>), dim3(nb(sizeh)), dim3(kBlock), 0, S(),
                           (const double2*)a, (const double2*)b, sizeh, ctl);
    else
        hipLaunchKernelGGL((k_reduce<double2, 1, kFinStore>), dim3(nb(sizeh)), dim3(kBlock), 0, S(),
                           (const double2*)a, (const double2*)b, sizeh, ctl);
    return hip_ok(hipGetLastError(), "k_reduce") &&
           launch_finalize(complex_dot ? kFinStore2 : kFinStore, ctl, nb(sizeh));
}

bool launch_finalize(int fin, const SolverCtl& c, int nparts) {
#define OSTP_F(F, NV) hipLaunchKernelGGL((k_finalize<F, NV>), dim3(1), dim3(kBlock), 0, S(), c, nparts)
    switch (fin) {
        case kFinStore: OSTP_F(kFinStore, 1); break;
        case kFinStore2: OSTP_F(kFinStore2, 2); break;
        case kFinCgAlpha: OSTP_F(kFinCgAlpha, 1); break;
        case kFinCgLambda: OSTP_F(kFinCgLambda, 1); break;
        case kFinCgRestart: OSTP_F(kFinCgRestart, 1); break;
        case kFinMsAlpha: OSTP_F(kFinMsAlpha, 1); break;
        case kFinMsLambda: OSTP_F(kFinMsLambda, 1); break;
        case kFinMxmAlpha: OSTP_F(kFinMxmAlpha, 1); break;
        case kFinPower: OSTP_F(kFinPower, 1); break;
        case kFinMxmLambda: OSTP_F(kFinMxmLambda, 1); break;
        case kFinMxAlpha: OSTP_F(kFinMxAlpha, 1); break;
        case kFinMxLambda: OSTP_F(kFinMxLambda, 1); break;
        case kFinFcgAlpha: OSTP_F(kFinFcgAlpha, 1); break;
        case kFinFcgRes: OSTP_F(kFinFcgRes, 1); break;
        case kFinFcgBeta: OSTP_F(kFinFcgBeta, 2); break;
        case kFinFcgInit: OSTP_F(kFinFcgInit, 2); break;
        default: return fail("unknown finalizer %d", fin);
    }
#undef OSTP_F
    return hip_ok(hipGetLastError(), "k_finalize");
}

SolverCtl plain_ctl() {
    SolverCtl c;
    c.st = ctx().state;
    c.partials = ctx().partials;
    c.pred = kAlways;
    return c;
}

// stand-alone reduction returning to the host
bool reduce_to_host(int kind, const void* a, const void* b, int single, double* re, double* im) {
    SolverCtl c = plain_ctl();
    bool ok;
    if (kind == 0) ok = single ? launch_norm2<float2, kFinStore>(a, c) : launch_norm2<double2, kFinStore>(a, c);
    else ok = launch_dot(a, b, kind == 2, c);
    if (!ok) return false;
    double buf[2];
    if (!hip_ok(hipMemcpyAsync(buf, &ctx().state->sum, 2 * sizeof(double), hipMemcpyDeviceToHost, S()), "readback"))
        return false;
    if (!hip_ok(hipStreamSynchronize(S()), "synchronize")) return false;
    *re = buf[0];
    if (im) *im = buf[1];
    return true;
}

}  // namespace ostp

using namespace ostp;

// ------------------------------------------------------------ C entry points
extern "C" {

#define VEC(p, name) resolve(p, ctx().lat.vec3_bytes(), name)
#define VECF(p, name) resolve(p, ctx().lat.vec3f_bytes(), name)

double l2norm2_global(const vec3_soa* in_vect1) {
    if (!require_init("l2norm2_global")) return 0.0;
    const void* a = VEC(in_vect1, "l2norm2_global");
    double r = 0.0;
    if (a) reduce_to_host(0, a, nullptr, 0, &r, nullptr);
    return r;
}

double real_scal_prod_global(const vec3_soa* in_vect1, const vec3_soa* in_vect2) {
    if (!require_init("real_scal_prod_global")) return 0.0;
    const void* a = VEC(in_vect1, "real_scal_prod_global");
    const void* b = VEC(in_vect2, "real_scal_prod_global");
    double r = 0.0;
    if (a && b) reduce_to_host(1, a, b, 0, &r, nullptr);
    return r;
}

void ostp_scal_prod_global(const vec3_soa* in_vect1, const vec3_soa* in_vect2, double* re, double* im) {
    *re = 0.0;
    *im = 0.0;
    if (!require_init("scal_prod_global")) return;
    const void* a = VEC(in_vect1, "scal_prod_global");
    const void* b = VEC(in_vect2, "scal_prod_global");
    if (a && b) reduce_to_host(2, a, b, 0, re, im);
}

double _Complex scal_prod_global(const vec3_soa* in_vect1, const vec3_soa* in_vect2) {
    double re = 0.0, im = 0.0;
    ostp_scal_prod_global(in_vect1, in_vect2, &re, &im);
    double _Complex z;
    __real__ z = re;
    __imag__ z = im;
    return z;
}

static void finish(const char* what) { entry_done(what); }

void combine_in1xfactor_plus_in2(const vec3_soa* in_vect1, const double factor, const vec3_soa* in_vect2,
                                 vec3_soa* out) {
    if (!require_init("combine_in1xfactor_plus_in2")) return;
    const void* a = VEC(in_vect1, "combine_in1xfactor_plus_in2");
    const void* b = VEC(in_vect2, "combine_in1xfactor_plus_in2");
    void* o = VEC(out, "combine_in1xfactor_plus_in2");
    if (a && b && o && launch_axpy<double2>(a, factor, b, o, SolverCtl{})) finish("combine_in1xfactor_plus_in2");
}

void combine_add_factor_x_in2_to_in1(vec3_soa* in1, const vec3_soa* in2, double factor) {
    if (!require_init("combine_add_factor_x_in2_to_in1")) return;
    void* a = VEC(in1, "combine_add_factor_x_in2_to_in1");
    const void* b = VEC(in2, "combine_add_factor_x_in2_to_in1");
    if (a && b && launch_axpy<double2>(b, factor, a, a, SolverCtl{})) finish("combine_add_factor_x_in2_to_in1");
}

void combine_in1xferm_mass_minus_in2(const vec3_soa* in_vect1, double ferm_mass2, vec3_soa* in_vect2) {
    if (!require_init("combine_in1xferm_mass_minus_in2")) return;
    const void* a = VEC(in_vect1, "combine_in1xferm_mass_minus_in2");
    void* b = VEC(in_vect2, "combine_in1xferm_mass_minus_in2");
    if (a && b && launch_mass_minus(a, ferm_mass2, b)) finish("combine_in1xferm_mass_minus_in2");
}

void combine_in1_minus_in2(const vec3_soa* in_vect1, const vec3_soa* in_vect2, vec3_soa* out) {
    if (!require_init("combine_in1_minus_in2")) return;
    const void* a = VEC(in_vect1, "combine_in1_minus_in2");
    const void* b = VEC(in_vect2, "combine_in1_minus_in2");
    void* o = VEC(out, "combine_in1_minus_in2");
    if (a && b && o && launch_sub(a, b, o)) finish("combine_in1_minus_in2");
}

// in2 = fact1*in1 - in2: the same elementwise formula as
// combine_in1xferm_mass_minus_in2 (a real factor times a complex number
// commutes exactly), so the same kernel serves both.
void combine_in1_x_fact1_minus_in2_back_into_in2(const vec3_soa* in1, double fact1, vec3_soa* in2) {
    const char* nm = "combine_in1_x_fact1_minus_in2_back_into_in2";
    if (!require_init(nm)) return;
    const void* a = VEC(in1, nm);
    void* b = VEC(in2, nm);
    if (a && b && launch_mass_minus(a, fact1, b)) finish(nm);
}

void combine_in1_minus_in2_allxfact(const vec3_soa* in1, const vec3_soa* in2, double fact, vec3_soa* out) {
    const char* nm = "combine_in1_minus_in2_allxfact";
    if (!require_init(nm)) return;
    const void* a = VEC(in1, nm);
    const void* b = VEC(in2, nm);
    void* o = VEC(out, nm);
    if (!a || !b || !o) return;
    const long n3 = 3 * ctx().lat.sizeh;
    hipLaunchKernelGGL(k_sub_scale, dim3(nb(n3)), dim3(kBlock), 0, S(), (const double2*)a, (const double2*)b, fact,
                       (double2*)o, n3);
    if (hip_ok(hipGetLastError(), nm)) finish(nm);
}

void assign_in_to_out(const vec3_soa* in_vect1, vec3_soa* out) {
    if (!require_init("assign_in_to_out")) return;
    const void* a = VEC(in_vect1, "assign_in_to_out");
    void* o = VEC(out, "assign_in_to_out");
    if (a && o &&
        hip_ok(hipMemcpyAsync(o, a, ctx().lat.vec3_bytes(), hipMemcpyDeviceToDevice, ctx().stream), "copy"))
        finish("assign_in_to_out");
}

void set_vec3_soa_to_zero(vec3_soa* fermion) {
    if (!require_init("set_vec3_soa_to_zero")) return;
    void* o = VEC(fermion, "set_vec3_soa_to_zero");
    if (o && hip_ok(hipMemsetAsync(o, 0, ctx().lat.vec3_bytes(), ctx().stream), "memset"))
        finish("set_vec3_soa_to_zero");
}

void convert_double_to_float_vec3_soa(const vec3_soa* d_var, vec3_soa_f* f_var) {
    if (!require_init("convert_double_to_float_vec3_soa")) return;
    const void* d = VEC(d_var, "convert_double_to_float_vec3_soa");
    void* f = VECF(f_var, "convert_double_to_float_vec3_soa");
    if (d && f && launch_d2f(d, f, 3 * ctx().lat.sizeh)) finish("convert_double_to_float_vec3_soa");
}

void convert_float_to_double_vec3_soa(const vec3_soa_f* f_var, vec3_soa* d_var) {
    if (!require_init("convert_float_to_double_vec3_soa")) return;
    const void* f = VECF(f_var, "convert_float_to_double_vec3_soa");
    void* d = VEC(d_var, "convert_float_to_double_vec3_soa");
    if (d && f && launch_f2d(f, d, 3 * ctx().lat.sizeh)) finish("convert_float_to_double_vec3_soa");
}

void ostp_convert_double_to_float_gauge(const su3_soa* d_var, su3_soa_f* f_var) {
    if (!require_init("ostp_convert_double_to_float_gauge")) return;
    const Lattice& L = ctx().lat;
    const char* d = (const char*)resolve(d_var, 8 * L.su3_bytes(), "ostp_convert_double_to_float_gauge");
    char* f = (char*)resolve(f_var, 8 * L.su3f_bytes(), "ostp_convert_double_to_float_gauge");
    if (!d || !f) return;
    for (int l = 0; l < 8; l++) {
        if (!launch_d2f(d + l * L.su3_bytes(), f + l * L.su3f_bytes(), 9 * L.sizeh)) return;
        if (L.par_temp) {  // K plane: double -> float
            const double* K = (const double*)(d + l * L.su3_bytes() + 144 * L.sizeh);
            float* Kf = (float*)(f + l * L.su3f_bytes() + 72 * L.sizeh);
            hipLaunchKernelGGL(k_convert_real, dim3(nb(L.sizeh)), dim3(kBlock), 0, S(), K, Kf, L.sizeh);
            if (!hip_ok(hipGetLastError(), "k_convert_real")) return;
        }
    }
    finish("ostp_convert_double_to_float_gauge");
}

}  // extern "C"
