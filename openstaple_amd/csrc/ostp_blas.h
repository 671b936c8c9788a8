// Internal launch interface of the BLAS, Dslash and solver kernels.
#pragma once

#include "ostp_reduce.h"

namespace ostp {

template <typename C> __host__ __device__ __forceinline__ C mkc(double a, double b);
template <> __host__ __device__ __forceinline__ double2 mkc<double2>(double a, double b) { return make_double2(a, b); }
template <> __host__ __device__ __forceinline__ float2 mkc<float2>(double a, double b) {
    return make_float2((float)a, (float)b);
}

// Dslash (dslash.hip).  mode: 0 plain, 1 out = m2*x - D in, 2 = 1 + fused
// Re<x,out> reduction handed to finalizer `fin`.
template <typename R>
bool launch_dslash(int eo, int mode, int fin, const void* u, const void* ph, const void* in, void* out,
                   const void* x, double m2, const SolverCtl& ctl);
// D_oe of n FP64 even fields stored one after the other (xs -> hs), the
// fields sharing each link load (solvers.hip)
bool launch_doe_multi(const void* u, const void* ph, const void* xs, void* hs, int n);
template <typename R>
bool launch_mdagm(const void* u, const void* ph, const void* in, void* temp, void* out, double m2, int mode,
                  int fin, const SolverCtl& ctl);

// BLAS (blas.hip)
template <typename C>
bool launch_axpy(const void* in1, double factor, const void* in2, void* out, const SolverCtl& ctl);
bool launch_sub(const void* in1, const void* in2, void* out);
bool launch_mass_minus(const void* in1, double m2, void* in2);
bool launch_d2f(const void* d, void* f, long n);
bool launch_f2d(const void* f, void* d, long n);
template <typename C, int F>
bool launch_norm2(const void* a, const SolverCtl& ctl);
bool launch_dot(const void* a, const void* b, int complex_dot, const SolverCtl& ctl);
bool reduce_to_host(int kind, const void* a, const void* b, int single, double* re, double* im);
SolverCtl plain_ctl();

// Gauge force and action (stout.hip).  launch_gauge_force: staples (when stap
// is not null) and ipdot of the gauge action; with mom, the momentum step
// mom_sum_mult(mom, ipdot, fact) fused into the same kernel.
bool launch_gauge_force(const void* u, void* stap, void* ipdot, void* mom, double fact);
// kind 0: sum of Re tr of the plaquettes, 1: of the 2x1 and 1x2 rectangles
bool gauge_action_sum(const void* u, int kind, double* out);
double stout_c_zero();
bool stout_tlsm();
// momenta action (md.hip): sum of Tr(P^2)/2 over link arrays [first, first + nlinks)
bool momenta_action(const void* mom, int first, int nlinks, double* out);

// U(1) phases without the reduction and the 2 pi (fields.hip)
bool u1_phases_unbounded(double* phases, bf_param b, double im_chem_pot, double charge);

// Batched even-odd solves (batch.hip); the C entry points ostp_batched_mixed_solve
// (half = false) and ostp_batched_half_solve (half = true) call it as well.
int batched_solve_internal(int nsys, const ostp_batch_system* sys, double res, int max_cg, double mixed_delta,
                           int iters_per_graph, int* cg_return, int* ru_return, double* true_res, bool half);

}  // namespace ostp
