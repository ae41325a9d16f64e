// Host build of this project's FP64 kernels and runtime for the sanitizer test
// (tests/test_cpu_sanitized.py): the HIP names the runtime and the FP64 kernels
// use, with "device" memory on the host heap (so that AddressSanitizer sees
// every kernel access) and each kernel launch run one work-item after another.
// Only for code without cross-lane operations; the reductions that use them
// are replaced in host_kernels.cpp.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1 };
enum { hipEventDisableTiming = 2 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };

struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct float4 { float x, y, z, w; };
inline int2 make_int2(int a, int b) { return {a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return {a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __syncthreads()
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
extern dim3 blockIdx, threadIdx, blockDim, gridDim;

template <class T> inline void unsafeAtomicAdd(T* p, T v) { *p += v; }
// the device reductions that shuffle are compiled but replaced (host_kernels.cpp)
template <class T> [[noreturn]] inline T __shfl_down(T, unsigned, int) { __builtin_trap(); }
inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
#define __builtin_amdgcn_sqrtf(x) std::sqrt((float)(x))
#define __builtin_amdgcn_rcpf(x) (1.0f / (float)(x))
#define __builtin_amdgcn_rcp(x) (1.0 / (double)(x))

// every work-item of the grid in turn (x dimension only, as the launches here use)
template <class K, class... A> inline void host_launch(K k, dim3 g, dim3 b, A... a)
{
    gridDim = g;
    blockDim = b;
    for (unsigned bx = 0; bx < g.x; ++bx)
        for (unsigned tx = 0; tx < b.x; ++tx) {
            blockIdx = dim3(bx);
            threadIdx = dim3(tx);
            k(a...);
        }
}
#define hipLaunchKernelGGL(k, g, b, shm, st, ...) host_launch(k, g, b, __VA_ARGS__)

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned f);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* n, char* l, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
template <typename T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc((void**)p, bytes); }
// a __device__ variable is host memory here: its address is the symbol's own
inline hipError_t hipGetSymbolAddress(void** p, const void* sym)
{
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
