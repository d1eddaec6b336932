// Host stand-in for <hip/hip_runtime_api.h>: the slice of the HIP runtime that
// the serving scheduler (goofy_amd/csrc/serve.cpp) calls, backed by a fake
// device (tests/native/serve_fake_backend.cpp).  "Device memory" is host
// memory; each stream is a worker thread that runs its operations in order;
// events complete when their stream reaches them.  Used only to build the
// scheduler for CPU tests under ThreadSanitizer / AddressSanitizer.
#pragma once

#include <stddef.h>
#include <stdint.h>

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorInvalidDevice = 101,
    hipErrorNotReady = 600,
    hipErrorLaunchFailure = 719,
} hipError_t;

typedef struct FakeStream* hipStream_t;
typedef struct FakeEvent* hipEvent_t;

enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2, hipHostMallocDefault = 0 };

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
} hipMemcpyKind;

hipError_t hipMalloc(void** ptr, size_t bytes);
hipError_t hipFree(void* ptr);
hipError_t hipHostMalloc(void** ptr, size_t bytes, unsigned int flags);
hipError_t hipHostFree(void* ptr);
hipError_t hipSetDevice(int device);
hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned int flags);
hipError_t hipStreamDestroy(hipStream_t stream);
hipError_t hipStreamSynchronize(hipStream_t stream);
hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned int flags);
hipError_t hipEventDestroy(hipEvent_t event);
hipError_t hipEventRecord(hipEvent_t event, hipStream_t stream);
hipError_t hipEventQuery(hipEvent_t event);
hipError_t hipEventSynchronize(hipEvent_t event);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t stream);
const char* hipGetErrorString(hipError_t error);

// ---- fake-device controls (tests only) ----------------------------------------
struct FakeHipControl {
    int kernel_delay_us;     // each kernel runs this long on its stream (0 = no delay)
    int fail_malloc_after;   // hipMalloc fails once this many more calls succeeded (-1 = never)
    int fail_launch_after;   // a kernel launch call fails once this many more succeeded (-1 = never)
    int fault_kernel_after;  // a launched kernel faults once this many more ran (-1 = never); events
                             // recorded after a fault on that stream report hipErrorLaunchFailure
};
// The controls are read under the fake device's own lock; set them while no
// stream runs work, or accept that running work sees the old or new values.
void fake_hip_set(const FakeHipControl& c);
FakeHipControl fake_hip_get();
// Kernel launches (batch, inline batch and plain) the fake device has executed.
uint64_t fake_hip_kernels_run();
