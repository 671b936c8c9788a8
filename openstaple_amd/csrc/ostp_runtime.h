// Internal runtime interface shared by the library's translation units.
#pragma once

#include "../../include/openstaple_amd_types.h"
#include "ostp_common.h"

namespace ostp {

constexpr int kMaxShifts = MAX_APPROX_ORDER;
constexpr int kBlock = 256;

// Scalar state of a solver, resident in device memory so that iterations can be
// queued without a host round trip per reduction.  Written only by the
// finalize kernel that closes each reduction (thread 0).
enum SolverStatus : int { kRunning = 0, kInnerExit = 1, kDone = 2 };

struct SolverState {
    // CG scalars (inverter_full.c:28, inverter_mixedp.c:54-56)
    double delta, alpha, lambda, omega, omega_save, gammag, source_norm;
    double res, safety, last_max_res, mixed_delta, shift_m2;
    double sum;               // result slot of stand-alone reductions
    double sum_im;
    double scale, rho;        // FP16 inner solver: inner vectors = scale x true; rescale at the last update
    double nentries;          // FP16 inner solver: complex entries of one vector (3 sizeh)
    double delta_old;         // FP16 inner solver: |r|^2 before the last update
    int cg, cg_restarted, restart_every, max_cg;
    int status;               // SolverStatus
    int ru;                   // mixed precision: reliable update in this iteration (FP16: due)
    int hold;                 // FP16 inner solver: inner iterations stop until the reliable update
    int xr_fresh;             // FP16 inner solver: 1 = |r|^2 partials of an x/r update not yet
                              // consumed, 2 = consumed by an iteration that stopped, 0 = none
    int ru_count;
    int resid_ok;             // FP16 inner solver: sum holds |b - M x|^2 (FP64) of the final x
    // multi-shift CG-M (inverter_multishift_full.c:48-53)
    int order, maxiter, maxiter_next, flag_cur;
    int flags[2][kMaxShifts];
    double zeta_i[kMaxShifts], zeta_ii[kMaxShifts], zeta_iii[kMaxShifts];
    double omegas[kMaxShifts], gammas[kMaxShifts], shifts[kMaxShifts];
};

struct Context {
    bool initialized = false;
    Lattice lat{};
    hipStream_t stream = nullptr;
    int device = 0;
    int abort_on_error = 1;
    int sync_at_return = 1;   // entry points without a result wait for their kernels
    // Dslash launches alternate the direction in which each XCD walks its
    // slab, so a launch starts on the links and phases the previous one
    // streamed last (still in the 256 MB memory-side cache)
    int dslash_serpentine = 1;
    unsigned dslash_launches = 0;
    // device scratch: reduction partials + solver state
    double* partials = nullptr;     // 2 * max_blocks
    SolverState* state = nullptr;   // device
    SolverState* state_host = nullptr;  // pinned mirror
    int max_blocks = 0;
    // extents for the host-only entry points (field files) set by
    // ostp_set_lattice_host; ostp_init clears them
    Lattice host_lat{};
    bool host_lat_set = false;
};

Context& ctx();

// Error policy: record `msg`; abort the process if configured (default).
// Returns false so callers can `return fail(...)`.
bool fail(const char* fmt, ...);
bool hip_ok(hipError_t e, const char* what);
// Failures recorded so far.  Code that calls other entry points compares two
// readings to learn whether those calls failed, whatever message an earlier
// call left recorded; a fallback that recovers from a failure it caused
// (an allocation it can do without) takes it back with forgive_errors(mark),
// which also clears the message.
unsigned long error_count();
void forgive_errors(unsigned long mark);
bool require_init(const char* what);

// Translate a host address registered with ostp_acc_* (or a device address)
// into the device address of `bytes` bytes; nullptr plus recorded error if the
// address is unknown or the range does not fit its allocation.
void* resolve(const void* p, size_t bytes, const char* what);
template <typename T>
inline T* resolve_as(const void* p, size_t bytes, const char* what) {
    return reinterpret_cast<T*>(resolve(p, bytes, what));
}

// Launch geometry over sizeh half-sites.
inline int blocks_for(long n, int nt = kBlock) { return (int)((n + nt - 1) / nt); }

// Gauge action selected by ostp_set_stout (true: tree-level Symanzik).
bool gauge_action_tlsm();

// End of a void entry point: wait for the stream unless asynchronous mode is on.
bool entry_done(const char* what);

// Set while fermion_force_soloopenacc runs (force.hip): the rational solves
// inside it are the MD force's (for a precision mode that applies there only).
bool in_fermion_force();
struct ForceScope {
    ForceScope();
    ~ForceScope();
};

// Address of the allocation-free solver scratch on `device` (solvers.hip);
// called by ostp_init with that device current.
bool noalloc_setup(int device);

// Read the device solver state into the pinned mirror (synchronises the stream).
bool pull_state();
bool push_state();

}  // namespace ostp

// Map the hardware workgroup id to a logical tile so that each of the 8 XCDs
// (which receive workgroups round-robin) works on a contiguous range of the
// lattice, keeping neighbour re-reads inside one XCD's L2.
__device__ __forceinline__ int xcd_block_id() {
    const int G = gridDim.x, i = blockIdx.x;
    const int q = G >> 3, r = G & 7, x = i & 7, k = i >> 3;
    return x * q + (x < r ? x : r) + k;
}

// The same with each XCD's range walked from its end when `rev` is set
// (workgroups are dispatched in hardware-id order, so a reversed launch
// starts where the previous, forward launch finished).
__device__ __forceinline__ int xcd_block_id_dir(int rev) {
    const int G = gridDim.x, i = blockIdx.x;
    const int q = G >> 3, r = G & 7, x = i & 7, k = i >> 3;
    const int len = q + (x < r ? 1 : 0);
    return x * q + (x < r ? x : r) + (rev ? len - 1 - k : k);
}
