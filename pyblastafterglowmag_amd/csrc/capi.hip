// Lifecycle entry points of the C ABI (include/pbagpu.h).
#include "pba_device.h"

extern "C" int pbagpu_device_count(int *count) {
    if (!count) return PBAGPU_E_ARG;
    return hipGetDeviceCount(count) == hipSuccess ? PBAGPU_OK : PBAGPU_E_HIP;
}

extern "C" int pbagpu_set_device(int device) {
    return hipSetDevice(device) == hipSuccess ? PBAGPU_OK : PBAGPU_E_HIP;
}

extern "C" int pbagpu_version(void) { return 1; }

// ------------------------------------------------------------------ kernel timing
#include <mutex>

namespace pbagpu {
static bool g_timing = false;
static std::mutex g_timing_mu;   // stages may run on several host threads (one stream each)
static double g_ms[PBAGPU_N_STAGES];
static long long g_launches[PBAGPU_N_STAGES];
static double g_units[PBAGPU_N_STAGES];

bool timing_enabled() { return g_timing; }

void timing_add(int stage, double ms) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return;
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_ms[stage] += ms;
    g_launches[stage] += 1;
}
void timing_add_units(int stage, double units) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return;
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_units[stage] += units;
}
}  // namespace pbagpu

extern "C" int pbagpu_timing_units(int stage, double *units, int reset) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return PBAGPU_E_ARG;
    std::lock_guard<std::mutex> lk(pbagpu::g_timing_mu);
    if (units) *units = pbagpu::g_units[stage];
    if (reset) pbagpu::g_units[stage] = 0.0;
    return PBAGPU_OK;
}

extern "C" int pbagpu_timing_enable(int on) {
    pbagpu::g_timing = on != 0;
    return PBAGPU_OK;
}

extern "C" int pbagpu_timing_read(int stage, double *ms_total, long long *launches, int reset) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return PBAGPU_E_ARG;
    std::lock_guard<std::mutex> lk(pbagpu::g_timing_mu);
    if (ms_total) *ms_total = pbagpu::g_ms[stage];
    if (launches) *launches = pbagpu::g_launches[stage];
    if (reset) { pbagpu::g_ms[stage] = 0.0; pbagpu::g_launches[stage] = 0; }
    return PBAGPU_OK;
}

// ------------------------------------------------------------- errors and workspace
#include "runtime.h"
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>

namespace pbagpu {
static thread_local char g_err[512] = "";

void set_error(const char *what, hipError_t e) {
    std::snprintf(g_err, sizeof(g_err), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

void set_error_msg(const char *msg) { std::snprintf(g_err, sizeof(g_err), "%s", msg); }

namespace {
struct Buf { void *p = nullptr; size_t cap = 0; };
struct StreamBufs { Buf b[WS_N]; };
// (device, stream) -> that stream's buffers; guarded by g_ws_mu.  Buffers are never moved, so
// a call holds its pointers without the lock while other streams allocate.
std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, StreamBufs *> g_ws;
}  // namespace

void *workspace(int slot, size_t bytes, hipStream_t stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_error("hipGetDevice", e); return nullptr; }
    if (slot < 0 || slot >= WS_N) { set_error_msg("workspace: bad slot"); return nullptr; }
    StreamBufs *sb;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto &ref = g_ws[{dev, stream}];
        if (!ref) ref = new StreamBufs();
        sb = ref;
    }
    Buf &b = sb->b[slot];
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return b.p;
    if (b.p) {
        // earlier work on this stream may still read the old buffer
        e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipFree(b.p);
        b.p = nullptr; b.cap = 0;
        if (e != hipSuccess) { set_error("hipFree(workspace)", e); return nullptr; }
    }
    size_t want = bytes + bytes / 2;
    e = hipMalloc(&b.p, want);
    if (e != hipSuccess) { b.p = nullptr; set_error("hipMalloc(workspace)", e); return nullptr; }
    b.cap = want;
    return b.p;
}

int release_stream(hipStream_t stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_error("hipGetDevice", e); return PBAGPU_E_HIP; }
    StreamBufs *sb = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto it = g_ws.find({dev, stream});
        if (it == g_ws.end()) return PBAGPU_OK;
        sb = it->second;
        g_ws.erase(it);
    }
    int rc = PBAGPU_OK;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) { set_error("hipStreamSynchronize", e); rc = PBAGPU_E_HIP; }
    for (int k = 0; k < WS_N; k++)
        if (sb->b[k].p) {
            e = hipFree(sb->b[k].p);
            if (e != hipSuccess) { set_error("hipFree(workspace)", e); rc = PBAGPU_E_HIP; }
        }
    delete sb;
    return rc;
}
}  // namespace pbagpu

extern "C" const char *pbagpu_last_error(void) { return pbagpu::g_err; }

extern "C" int pbagpu_release_stream(void *stream) { return pbagpu::release_stream((hipStream_t)stream); }

// ------------------------------------------------------------ CU-partitioned streams
namespace pbagpu {
int stream_cu_count(hipStream_t stream) {
    int dev = 0, n_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    uint32_t mask[32] = {0};
    if (hipExtStreamGetCUMask(stream, 32, mask) != hipSuccess) return n_cu;
    int c = 0;
    for (int i = 0; i < 32 && i * 32 < n_cu; i++) c += __builtin_popcount(mask[i]);
    return (c > 0 && c <= n_cu) ? c : n_cu;
}
}  // namespace pbagpu

extern "C" int pbagpu_stream_create(int cu_first, int cu_count, void **stream) {
    if (!stream) return PBAGPU_E_ARG;
    *stream = nullptr;
    int dev = 0, n_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) { pbagpu::set_error("hipDeviceGetAttribute", e); return PBAGPU_E_HIP; }
    hipStream_t s = nullptr;
    if (cu_count <= 0) {
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) { pbagpu::set_error("hipStreamCreateWithFlags", e); return PBAGPU_E_HIP; }
        *stream = (void *)s;
        return PBAGPU_OK;
    }
    if (cu_first < 0 || cu_first + cu_count > n_cu || n_cu > 32 * 32) {
        pbagpu::set_error_msg("pbagpu_stream_create: CU range outside the device");
        return PBAGPU_E_ARG;
    }
    uint32_t mask[32] = {0};
    for (int c = cu_first; c < cu_first + cu_count; c++) mask[c / 32] |= 1u << (c % 32);
    e = hipExtStreamCreateWithCUMask(&s, (uint32_t)((n_cu + 31) / 32), mask);
    if (e != hipSuccess) { pbagpu::set_error("hipExtStreamCreateWithCUMask", e); return PBAGPU_E_HIP; }
    *stream = (void *)s;
    return PBAGPU_OK;
}

extern "C" int pbagpu_stream_destroy(void *stream) {
    if (!stream) return PBAGPU_E_ARG;
    int rc = pbagpu::release_stream((hipStream_t)stream);
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    if (e != hipSuccess) { pbagpu::set_error("hipStreamDestroy", e); return PBAGPU_E_HIP; }
    return rc;
}

// ------------------------------------------------------------------- options
#include <atomic>

namespace pbagpu {
static std::atomic<long long> g_options[PBAGPU_N_OPTIONS];
long long option(int which) { return (which >= 0 && which < PBAGPU_N_OPTIONS) ? g_options[which].load() : 0; }
}  // namespace pbagpu

extern "C" int pbagpu_set_option(int which, long long value) {
    if (which < 0 || which >= PBAGPU_N_OPTIONS) {
        pbagpu::set_error_msg("pbagpu_set_option: unknown option");
        return PBAGPU_E_ARG;
    }
    if (which == PBAGPU_OPT_ELE_GRAPH && value != 0 && value != 1) {
        pbagpu::set_error_msg("pbagpu_set_option: PBAGPU_OPT_ELE_GRAPH takes 0 or 1");
        return PBAGPU_E_ARG;
    }
    if (which == PBAGPU_OPT_DYN_RUNS_PER_WAVE && (value < 0 || value > 64)) {
        pbagpu::set_error_msg("pbagpu_set_option: PBAGPU_OPT_DYN_RUNS_PER_WAVE takes 0 .. 64");
        return PBAGPU_E_ARG;
    }
    pbagpu::g_options[which].store(value);
    return PBAGPU_OK;
}

extern "C" int pbagpu_get_option(int which, long long *value) {
    if (which < 0 || which >= PBAGPU_N_OPTIONS || !value) {
        pbagpu::set_error_msg("pbagpu_get_option: unknown option or null value");
        return PBAGPU_E_ARG;
    }
    *value = pbagpu::g_options[which].load();
    return PBAGPU_OK;
}
