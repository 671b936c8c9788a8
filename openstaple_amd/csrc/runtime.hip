// Runtime of the library: lattice context, device scratch, the host->device
// registry that replaces OpenACC's present table, and the error policy.
#include "ostp_runtime.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

// ------------------------------------------------------- reference globals
// Defaults of Include/setting_file_parser.c:462,490-492; a host program that
// defines these symbols itself overrides them (ELF interposition).
extern "C" {
int verbosity_lv = 0;
inv_tricks inverter_tricks = {0, 0, 0.1, 10000};
int multishift_invert_iterations = 0;
}

static_assert(sizeof(RationalApprox) == 456, "RationalApprox layout (rationalapprox.h:15-26)");
static_assert(offsetof(ferm_param, phases) == 72, "ferm_param layout (fermion_parameters.h:9-41)");
static_assert(offsetof(ferm_param, phases_f) == 104, "ferm_param layout (fermion_parameters.h:9-41)");
static_assert(sizeof(inverter_package) == 112, "inverter_package layout (inverter_package.h:12-29)");
static_assert(sizeof(inv_tricks) == 24, "inv_tricks layout (inverter_tricks.h:4-11)");

namespace ostp {

static Context g_ctx;
static std::string g_last_error;

struct Mapping {
    size_t bytes;
    void* dev;
    int refs;
};
static std::map<uintptr_t, Mapping> g_map;  // keyed by host start address
static std::mutex g_mu;

Context& ctx() { return g_ctx; }

static unsigned long g_error_count = 0;
unsigned long error_count() { return g_error_count; }
void forgive_errors(unsigned long mark) {
    g_error_count = mark;
    g_last_error.clear();
}

bool fail(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    g_error_count++;
    if (g_ctx.abort_on_error) {
        fprintf(stderr, "openstaple_amd: FATAL: %s\n", buf);
        fflush(stderr);
        abort();
    }
    return false;
}

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    return fail("%s: %s", what, hipGetErrorString(e));
}

bool require_init(const char* what) {
    if (g_ctx.initialized) return true;
    return fail("%s called before ostp_init()", what);
}

void* resolve(const void* p, size_t bytes, const char* what) {
    if (p == nullptr) {
        fail("%s: null pointer", what);
        return nullptr;
    }
    const uintptr_t a = (uintptr_t)p;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_map.upper_bound(a);
        if (it != g_map.begin()) {
            --it;
            const uintptr_t base = it->first;
            if (a >= base && a < base + it->second.bytes) {
                if (a + bytes > base + it->second.bytes) {
                    fail("%s: host range %p+%zu overruns its present region (%zu bytes)", what, p, bytes,
                         it->second.bytes);
                    return nullptr;
                }
                return (char*)it->second.dev + (a - base);
            }
        }
    }
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e == hipSuccess && (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged)) {
        void* base = nullptr;
        size_t size = 0;
        if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) == hipSuccess && base != nullptr) {
            if (a + bytes > (uintptr_t)base + size) {
                fail("%s: device range %p+%zu overruns its allocation (%zu bytes from %p)", what, p, bytes, size,
                     base);
                return nullptr;
            }
        }
        return const_cast<void*>(p);
    }
    (void)hipGetLastError();
    fail("%s: %p is neither present on the device (ostp_acc_create/copyin) nor device memory", what, p);
    return nullptr;
}

bool pull_state() {
    if (!hip_ok(hipMemcpyAsync(g_ctx.state_host, g_ctx.state, sizeof(SolverState), hipMemcpyDeviceToHost,
                               g_ctx.stream),
                "state readback"))
        return false;
    return hip_ok(hipStreamSynchronize(g_ctx.stream), "stream synchronize");
}

bool entry_done(const char* what) {
    if (!g_ctx.sync_at_return) return true;
    return hip_ok(hipStreamSynchronize(g_ctx.stream), what);
}

bool push_state() {
    if (!hip_ok(hipMemcpyAsync(g_ctx.state, g_ctx.state_host, sizeof(SolverState), hipMemcpyHostToDevice,
                               g_ctx.stream),
                "state upload"))
        return false;
    return hip_ok(hipStreamSynchronize(g_ctx.stream), "stream synchronize");
}

static void release_scratch() {
    if (g_ctx.partials) (void)hipFree(g_ctx.partials);
    if (g_ctx.state) (void)hipFree(g_ctx.state);
    if (g_ctx.state_host) (void)hipHostFree(g_ctx.state_host);
    g_ctx.partials = nullptr;
    g_ctx.state = nullptr;
    g_ctx.state_host = nullptr;
}

}  // namespace ostp

using namespace ostp;

namespace ostp {
static int g_force_depth = 0;
bool in_fermion_force() { return g_force_depth > 0; }
ForceScope::ForceScope() { ++g_force_depth; }
ForceScope::~ForceScope() { --g_force_depth; }
}  // namespace ostp

// HIP events on the library stream around a span of asynchronous calls
// (ostp_timer_start / _stop / _elapsed_us)
static hipEvent_t g_timer[2] = {nullptr, nullptr};

extern "C" {

int ostp_init(int nd0, int nd1, int nd2, int nd3, int par_temp) {
    if (nd0 <= 0 || nd1 <= 0 || nd2 <= 0 || nd3 <= 0 || (nd0 & 1))
        return fail("ostp_init: invalid lattice %dx%dx%dx%d (extents > 0, nd0 even)", nd0, nd1, nd2, nd3) ? 0 : 1;
    long vol = (long)nd0 * nd1 * nd2 * nd3;
    if (vol / 2 > (1L << 30)) return fail("ostp_init: lattice too large") ? 0 : 1;
    Context& c = g_ctx;
    if (!c.stream) {
        if (!hip_ok(hipGetDevice(&c.device), "hipGetDevice")) return 1;
        if (!hip_ok(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate")) return 1;
    }
    Lattice L{};
    L.nd[0] = nd0; L.nd[1] = nd1; L.nd[2] = nd2; L.nd[3] = nd3;
    L.sizeh = vol / 2;
    L.par_temp = par_temp ? 1 : 0;
    for (int i = 0; i < 4; i++) L.map[i] = c.initialized ? c.lat.map[i] : i;
    int blocks = blocks_for(L.sizeh, 64) + 64;
    if (!c.partials || blocks > c.max_blocks) {
        release_scratch();
        if (!hip_ok(hipMalloc(&c.partials, sizeof(double) * 4 * (size_t)blocks), "hipMalloc partials")) return 1;
        if (!hip_ok(hipMalloc(&c.state, sizeof(SolverState)), "hipMalloc state")) return 1;
        if (!hip_ok(hipHostMalloc(&c.state_host, sizeof(SolverState), hipHostMallocDefault), "hipHostMalloc"))
            return 1;
        c.max_blocks = blocks;
    }
    if (!noalloc_setup(c.device)) return 1;
    c.lat = L;
    c.initialized = true;
    c.host_lat_set = false;
    return 0;
}

void ostp_finalize(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_map) (void)hipFree(kv.second.dev);
    g_map.clear();
    release_scratch();
    for (auto& e : g_timer) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (g_ctx.stream) (void)hipStreamDestroy(g_ctx.stream);
    g_ctx.stream = nullptr;
    g_ctx.initialized = false;
}

int ostp_set_direction_map(int xmap, int ymap, int zmap, int tmap) {
    if (!require_init("ostp_set_direction_map")) return 1;
    int m[4] = {xmap, ymap, zmap, tmap};
    for (int i = 0; i < 4; i++) {
        if (m[i] < 0 || m[i] > 3) return fail("direction map entries must be 0..3") ? 0 : 1;
        for (int j = i + 1; j < 4; j++)
            if (m[i] == m[j]) return fail("found two equal direction mappings (geometry.c:160-171)") ? 0 : 1;
    }
    for (int i = 0; i < 4; i++) g_ctx.lat.map[i] = m[i];
    return 0;
}

int ostp_check_sizes(size_t sv, size_t su, size_t sd) {
    if (!require_init("ostp_check_sizes")) return 1;
    const Lattice& L = g_ctx.lat;
    if (sv != L.vec3_bytes() || su != L.su3_bytes() || sd != L.dsoa_bytes())
        return fail("struct sizes (%zu,%zu,%zu) do not match the lattice (%zu,%zu,%zu)", sv, su, sd,
                    L.vec3_bytes(), L.su3_bytes(), L.dsoa_bytes())
                   ? 0
                   : 1;
    return 0;
}

long ostp_sizeh(void) { return g_ctx.initialized ? g_ctx.lat.sizeh : -1; }

int ostp_acc_create(void* host, size_t bytes) {
    if (!host || bytes == 0) return fail("ostp_acc_create: empty range") ? 0 : 1;
    std::lock_guard<std::mutex> lk(g_mu);
    uintptr_t a = (uintptr_t)host;
    auto it = g_map.upper_bound(a);
    if (it != g_map.begin()) {
        auto pv = std::prev(it);
        if (a >= pv->first && a < pv->first + pv->second.bytes) {
            if (a + bytes > pv->first + pv->second.bytes)
                return fail("ostp_acc_create: range partially present") ? 0 : 1;
            pv->second.refs++;
            return 0;
        }
    }
    if (it != g_map.end() && it->first < a + bytes)
        return fail("ostp_acc_create: range overlaps a present region") ? 0 : 1;
    void* dev = nullptr;
    hipError_t e = hipMalloc(&dev, bytes);
    if (e != hipSuccess) return hip_ok(e, "ostp_acc_create hipMalloc") ? 0 : 1;
    g_map[a] = Mapping{bytes, dev, 1};
    return 0;
}

// Host<->device transfers and fills run on the library stream and wait for it:
// the stream is non-blocking, so a null-stream hipMemcpy/hipMemset would not be
// ordered against the kernels queued on it (a large fill could still be
// running while a kernel writes the same buffer).
static hipStream_t lib_stream() { return g_ctx.stream; }
static int stream_copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, const char* what) {
    hipStream_t s = lib_stream();
    if (!hip_ok(hipMemcpyAsync(dst, src, bytes, kind, s), what)) return 1;
    return hip_ok(s ? hipStreamSynchronize(s) : hipDeviceSynchronize(), what) ? 0 : 1;
}

int ostp_acc_update_device(const void* host, size_t bytes) {
    void* dev = resolve(host, bytes, "ostp_acc_update_device");
    if (!dev) return 1;
    return stream_copy(dev, host, bytes, hipMemcpyHostToDevice, "update device");
}

int ostp_acc_copyin(void* host, size_t bytes) {
    if (ostp_acc_create(host, bytes)) return 1;
    return ostp_acc_update_device(host, bytes);
}

int ostp_acc_update_self(void* host, size_t bytes) {
    void* dev = resolve(host, bytes, "ostp_acc_update_self");
    if (!dev) return 1;
    return stream_copy(host, dev, bytes, hipMemcpyDeviceToHost, "update self");
}

int ostp_acc_delete(void* host) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_map.find((uintptr_t)host);
    if (it == g_map.end()) return 1;
    if (--it->second.refs > 0) return 0;
    if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(it->second.dev);
    g_map.erase(it);
    return 0;
}

void* ostp_acc_deviceptr(const void* host) {
    std::lock_guard<std::mutex> lk(g_mu);
    uintptr_t a = (uintptr_t)host;
    auto it = g_map.upper_bound(a);
    if (it == g_map.begin()) return nullptr;
    --it;
    if (a >= it->first && a < it->first + it->second.bytes) return (char*)it->second.dev + (a - it->first);
    return nullptr;
}

int ostp_acc_is_present(const void* host, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_mu);
    uintptr_t a = (uintptr_t)host;
    auto it = g_map.upper_bound(a);
    if (it == g_map.begin()) return 0;
    --it;
    return (a >= it->first && a + bytes <= it->first + it->second.bytes) ? 1 : 0;
}

void* ostp_malloc(size_t bytes) {
    void* p = nullptr;
    if (!hip_ok(hipMalloc(&p, bytes), "ostp_malloc")) return nullptr;
    return p;
}

void ostp_free(void* dev) {
    if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
    (void)hipFree(dev);
}

int ostp_memcpy_htod(void* dev, const void* host, size_t bytes) {
    return stream_copy(dev, host, bytes, hipMemcpyHostToDevice, "ostp_memcpy_htod");
}

int ostp_memcpy_dtoh(void* host, const void* dev, size_t bytes) {
    return stream_copy(host, dev, bytes, hipMemcpyDeviceToHost, "ostp_memcpy_dtoh");
}

int ostp_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    return stream_copy(dst, src, bytes, hipMemcpyDeviceToDevice, "ostp_memcpy_dtod");
}

int ostp_memset(void* dev, int value, size_t bytes) {
    hipStream_t s = lib_stream();
    if (!hip_ok(hipMemsetAsync(dev, value, bytes, s), "ostp_memset")) return 1;
    return hip_ok(s ? hipStreamSynchronize(s) : hipDeviceSynchronize(), "ostp_memset") ? 0 : 1;
}

int ostp_synchronize(void) {
    hipError_t e = g_ctx.stream ? hipStreamSynchronize(g_ctx.stream) : hipDeviceSynchronize();
    return hip_ok(e, "ostp_synchronize") ? 0 : 1;
}

void ostp_set_abort_on_error(int a) { g_ctx.abort_on_error = a; }

int ostp_set_sync_at_return(int on) {
    const int prev = g_ctx.sync_at_return;
    g_ctx.sync_at_return = on ? 1 : 0;
    return prev;
}
int ostp_timer_start(void) {
    if (!require_init("ostp_timer_start")) return 1;
    for (auto& e : g_timer)
        if (!e && !hip_ok(hipEventCreate(&e), "hipEventCreate")) return 1;
    return hip_ok(hipEventRecord(g_timer[0], g_ctx.stream), "hipEventRecord") ? 0 : 1;
}
int ostp_timer_stop(void) {
    if (!g_timer[1]) return fail("ostp_timer_stop without ostp_timer_start") ? 1 : 1;
    return hip_ok(hipEventRecord(g_timer[1], g_ctx.stream), "hipEventRecord") ? 0 : 1;
}
double ostp_timer_elapsed_us(void) {
    float ms = 0.f;
    if (!g_timer[0] || !g_timer[1]) return -1.0;
    if (!hip_ok(hipEventSynchronize(g_timer[1]), "hipEventSynchronize") ||
        !hip_ok(hipEventElapsedTime(&ms, g_timer[0], g_timer[1]), "hipEventElapsedTime"))
        return -1.0;
    return 1000.0 * (double)ms;
}

int ostp_set_dslash_serpentine(int on) {
    const int prev = g_ctx.dslash_serpentine;
    g_ctx.dslash_serpentine = on ? 1 : 0;
    g_ctx.dslash_launches = 0;
    return prev;
}
const char* ostp_last_error(void) { return g_last_error.c_str(); }
void ostp_clear_error(void) { g_last_error.clear(); }

void ostp_set_inverter_tricks(int single_accel, int use_mixed, double delta, int restart_every) {
    inverter_tricks.singlePInvAccelMultiInv = single_accel;
    inverter_tricks.useMixedPrecision = use_mixed;
    inverter_tricks.mixedPrecisionDelta = delta;
    inverter_tricks.restartingEvery = restart_every;
}

}  // extern "C"
