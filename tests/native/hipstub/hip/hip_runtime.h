// A host-only stand-in for the HIP runtime, for building the engine's host
// code (plans, batches in flight, hit slots, the pinned-buffer pool and
// cache, the insertion thread, the request loop) with g++ under AddressSanitizer
// or ThreadSanitizer.  Streams are real: each runs its operations in order on a
// thread of its own, so copies and "kernels" happen asynchronously, after the
// call that enqueued them returns, as on the device.  hipFree and hipHostFree
// wait for every stream first (the device synchronisation they imply).
// Test infrastructure only.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

struct float2 {
    float x, y;
};
struct double2 {
    double x, y;
};
struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorUnknown = 999
} hipError_t;

typedef struct StubStream *hipStream_t;
typedef struct StubEvent *hipEvent_t;
typedef struct StubGraph *hipGraph_t;
typedef struct StubGraphExec *hipGraphExec_t;
typedef struct StubGraphNode *hipGraphNode_t;
struct hipKernelNodeParams {
    dim3 blockDim;
    void **extra;
    void *func;
    dim3 gridDim;
    void **kernelParams;
    unsigned sharedMemBytes;
};

enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
};
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipEventDisableSystemFence 0x20000000u
#define hipHostMallocDefault 0u

const char *hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceSynchronize();
hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2D(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                       hipMemcpyKind k);
hipError_t hipMemset(void *p, int v, size_t n);
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t st);
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned flags);
// CU masks are accepted and ignored by the stand-in (its streams are threads)
hipError_t hipExtStreamCreateWithCUMask(hipStream_t *st, uint32_t mask_words, const uint32_t *mask);
typedef enum hipDeviceAttribute_t {
    hipDeviceAttributeMultiprocessorCount = 1,
    hipDeviceAttributeNumberOfXccs = 2
} hipDeviceAttribute_t;
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int device);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t ev, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *ev);
hipError_t hipEventCreateWithFlags(hipEvent_t *ev, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t ev);
hipError_t hipEventRecord(hipEvent_t ev, hipStream_t st);
hipError_t hipEventQuery(hipEvent_t ev);
hipError_t hipEventSynchronize(hipEvent_t ev);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g);
hipError_t hipGraphCreate(hipGraph_t *g, unsigned flags);
// Kernel nodes of the stand-in: kernelParams[0] points to the address of the
// recorded operation (see the launchers in hipstub.cpp).
hipError_t hipGraphAddKernelNode(hipGraphNode_t *n, hipGraph_t g, const hipGraphNode_t *deps, size_t ndeps,
                                 const hipKernelNodeParams *p);
hipError_t hipLaunchKernel(const void *f, dim3 g, dim3 b, void **args, size_t smem, hipStream_t st);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t st);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);

template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags)
{
    return hipHostMalloc((void **)p, n, flags);
}

// Stand-in bookkeeping the stress driver checks: bytes held by hipMalloc.
long long hipstub_device_bytes();
