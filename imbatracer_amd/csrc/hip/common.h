// Shared HIP helpers: error checks, per-thread default stream, device properties.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "imba_hip.h"
#include "../host/status.h"

#define IMBA_HIP_TRY(expr)                                                                        \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return ::imba::fail(_e == hipErrorOutOfMemory ? IMBA_ERR_OUT_OF_MEMORY : IMBA_ERR_HIP, \
                                "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

namespace imba {

struct DeviceProps {
    int device = -1;
    int cu_count = 0;
    int lds_per_block = 0;
    int l2_bytes = 0;
    char name[256] = {0};
};

// Properties of the current device (queried once per device).
const DeviceProps* device_props();

// The stream used when a caller passes NULL: one non-blocking stream per host thread and
// device, default priority.  Calls made with it synchronize before returning.
hipStream_t thread_stream();

// Resolves the caller's stream argument.  *sync is set when the call must synchronize.
inline hipStream_t resolve_stream(void* s, bool* sync) {
    if (s) { *sync = false; return (hipStream_t)s; }
    *sync = true;
    return thread_stream();
}

// Library-owned device scratch that grows on demand (one per device, guarded by a mutex;
// callers on one device must not overlap in time on the scratch they hold).
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Grows s to at least `bytes` (device-wide synchronisation first when it must be replaced).
// The replaced buffer is retired, not freed, while graphs exist (see retire_buffer).
int grow_buffer(Scratch& s, size_t bytes);

// Frees a library-owned device buffer, or keeps it until the last live graph is released when
// graphs exist (a captured launch may still name it).
int retire_buffer(void* p);

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace imba
