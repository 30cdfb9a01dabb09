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
namespace pbagpu {
static bool g_timing = false;
static double g_ms[PBAGPU_N_STAGES];
static long long g_launches[PBAGPU_N_STAGES];

bool timing_enabled() { return g_timing; }

void timing_add(int stage, double ms) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return;
    g_ms[stage] += ms;
    g_launches[stage] += 1;
}
}  // namespace pbagpu

extern "C" int pbagpu_timing_enable(int on) {
    pbagpu::g_timing = on != 0;
    return PBAGPU_OK;
}

extern "C" int pbagpu_timing_read(int stage, double *ms_total, long long *launches, int reset) {
    if (stage < 0 || stage >= PBAGPU_N_STAGES) return PBAGPU_E_ARG;
    if (ms_total) *ms_total = pbagpu::g_ms[stage];
    if (launches) *launches = pbagpu::g_launches[stage];
    if (reset) { pbagpu::g_ms[stage] = 0.0; pbagpu::g_launches[stage] = 0; }
    return PBAGPU_OK;
}
