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
