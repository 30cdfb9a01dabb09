// Host-side runtime shared by the stage entry points: error reporting and the library's
// device workspace (growable buffers per device, stream and slot, so that repeated calls do not
// allocate).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "../../include/pbagpu.h"

namespace pbagpu {

// Remember the failing call and its HIP error for pbagpu_last_error().
void set_error(const char *what, hipError_t e);
void set_error_msg(const char *msg);

enum WsSlot {
    WS_DYN_PARAMS = 0, WS_DYN_INIT, WS_DYN_RES, WS_DYN_NT, WS_DYN_K,
    WS_ELE_GRIDS, WS_ELE_KTAB, WS_ELE_POPS, WS_ELE_WORK, WS_ELE_LIST, WS_ELE_KBF, WS_ELE_MIXED,
    WS_EATS_NU, WS_EATS_LAYERS, WS_EATS_STORE, WS_EATS_COUNTER, WS_EATS_SAMPLES,
    WS_PW_NU, WS_PW_LAYERS, WS_PW_SAMPLES,
    WS_SKY_SEG, WS_SKY_OFF, WS_SKY_PART,
    WS_VEC_NU, WS_VEC_LAYERS, WS_VEC_SAMPLES,
    WS_MLP, WS_MLP_ELE, WS_SUR, WS_SSCX,
    WS_N
};

// Device buffer of at least `bytes` for `slot`, owned by (current device, `stream`).  Every
// stream has its own set of buffers, so calls on different streams (one host thread each) never
// share or free each other's memory.  Growing a buffer first waits for `stream`, then frees the
// old buffer, so a call must not hold a pointer from an earlier call on the same stream.
// Returns nullptr on failure (error recorded).
void *workspace(int slot, size_t bytes, hipStream_t stream);
// Free every workspace buffer of (current device, stream), after waiting for the stream.
int release_stream(hipStream_t stream);
// Compute units a stream may use (its CU mask; all of the device for a plain stream).
int stream_cu_count(hipStream_t stream);
// Value of a library option (pbagpu_set_option).
long long option(int which);

}  // namespace pbagpu

// Evaluate a HIP call; on failure record it and set rc to PBAGPU_E_HIP.
#define PBA_TRY(rc, call)                                                  \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess) { pbagpu::set_error(#call, e_); (rc) = PBAGPU_E_HIP; } \
    } while (0)
