// Host build of the FP64 kernels (csrc/dmpm_kernels.hip) for the sanitizer test:
// with hip/hip_runtime.h from this folder every launch runs its work-items one
// after another on the host, and device buffers are heap blocks, so that
// AddressSanitizer checks every kernel access and UndefinedBehaviorSanitizer
// every integer operation of the kernels' index arithmetic.  The three
// reductions that use cross-lane operations are restated here as host loops
// (same per-node terms and dL/dm as k_loss).  The FP32 block kernels are not
// part of this build: their launchers abort.
#include "hip/hip_runtime.h"

#include <cstdio>
#include <cstdlib>

dim3 blockIdx, threadIdx, blockDim, gridDim;

#define launch_loss launch_loss_device
#define launch_sum_rows launch_sum_rows_device
#define launch_sumsq launch_sumsq_device
#include "dmpm_kernels.hip"
#include "dmpm_phases.hip"
#undef launch_loss
#undef launch_sum_rows
#undef launch_sumsq

namespace dmpm {

template <typename R>
void launch_loss(const StepParams& sp, const R* gm, const R* target, double* dLdm, double* part, double* loss, hipStream_t)
{
    double sum = 0.0;
    for (long long i = 0; i < sp.G; ++i) {
        const double cm = gm[i], tm = target[i];
        const double c = std::log(cm + 1.0 + 1e-4) - std::log(tm + 1.0 + 1e-4);
        double l = 0.5 * c * c, g = c / (cm + 1.0 + 1e-4);
        if (cm < 1e-3) {
            const double d = 1e-3 - cm;
            l += 10.0 * d * d;
            g -= 20.0 * d;
        }
        sum += l;
        dLdm[i] = g;
    }
    part[0] = sum;
    *loss = sum;
}

void launch_sum_rows(const double* part, int rows, int len, double* out, hipStream_t)
{
    for (int r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int i = 0; i < len; ++i) s += part[(long long)r * len + i];
        out[r] = s;
    }
}

template <typename R>
void launch_sumsq(const R* a, long long count, double* part, double* out, hipStream_t)
{
    double s = 0.0;
    for (long long i = 0; i < count; ++i) s += double(a[i]) * double(a[i]);
    part[0] = s;
    *out = s;
}

template void launch_loss<double>(const StepParams&, const double*, const double*, double*, double*, double*, hipStream_t);
template void launch_loss<float>(const StepParams&, const float*, const float*, double*, double*, double*, hipStream_t);
template void launch_sumsq<double>(const double*, long long, double*, double*, hipStream_t);
template void launch_sumsq<float>(const float*, long long, double*, double*, hipStream_t);

[[noreturn]] static void fp32_only(const char* what)
{
    fprintf(stderr, "host kernel build: %s is an FP32 block kernel, not part of this build\n", what);
    abort();
}
void launch_clear_tiles(const Batch&, float*, unsigned*, hipStream_t) { fp32_only("clear_tiles"); }
void launch_set_flags(unsigned*, long long, hipStream_t) { fp32_only("set_flags"); }
void launch_p2g_blk(const Batch&, const StepIO<float>&, const TileClear&, hipStream_t) { fp32_only("p2g_blk"); }
void launch_step_blk(const Batch&, const StepIO<float>&, const StepIO<float>&, const TileClear&, hipStream_t) { fp32_only("step_blk"); }
void launch_p2g_mass_blk(const Batch&, const float*, const float*, float*, hipStream_t) { fp32_only("p2g_mass_blk"); }
void launch_g2p_blk(const Batch&, const StepIO<float>&, const TileClear&, hipStream_t) { fp32_only("g2p_blk"); }
void launch_g2p_adj_blk(const Batch&, const AdjIO<float>&, hipStream_t) { fp32_only("g2p_adj_blk"); }
void launch_p2g_adj_blk(const Batch&, const AdjIO<float>&, const TileClear&, int, int, const int*, hipStream_t) { fp32_only("p2g_adj_blk"); }
void launch_loss_b(const Batch&, const float*, const float*, double*, double*, double*, long long, hipStream_t)
{
    fp32_only("loss_b");
}
void launch_loss_particles_b(const Batch&, const float*, const float*, const double*, float*, hipStream_t)
{
    fp32_only("loss_particles_b");
}
size_t sort_temp_bytes(int, long long)
{
    fp32_only("sort_temp_bytes");
    return 0;
}
hipError_t launch_sort_order(const SortArgs&, hipStream_t)
{
    fp32_only("sort_order");
    return hipErrorInvalidValue;
}
template <typename T> void launch_gather(const T*, T*, long long, int, int, const int*, hipStream_t) { fp32_only("gather"); }
template void launch_gather<float>(const float*, float*, long long, int, int, const int*, hipStream_t);
template void launch_gather<int>(const int*, int*, long long, int, int, const int*, hipStream_t);
template <typename T> void launch_scatter(const T*, T*, long long, int, int, const int*, hipStream_t) { fp32_only("scatter"); }
template void launch_scatter<float>(const float*, float*, long long, int, int, const int*, hipStream_t);
void launch_invert(const int*, int*, int, hipStream_t) { fp32_only("invert"); }
void launch_iota(int*, int, hipStream_t) { fp32_only("iota"); }

}  // namespace dmpm

// ------------------------------------------------ runtime API on the host --
// the error of the last failed call, returned and cleared by hipGetLastError as
// the HIP runtime does (a failed hipMalloc leaves it set)
static hipError_t g_last = hipSuccess;

// DMPM_HOST_MALLOC_LIMIT (bytes, read at every call): larger requests fail as
// out of memory, so that tests can take the library's allocation-failure paths
hipError_t hipMalloc(void** p, size_t bytes)
{
    const char* lim = getenv("DMPM_HOST_MALLOC_LIMIT");
    if (lim && bytes > strtoull(lim, nullptr, 10)) {
        *p = nullptr;
        g_last = hipErrorOutOfMemory;
        return hipErrorOutOfMemory;
    }
    *p = malloc(bytes ? bytes : 1);
    if (!*p) {
        g_last = hipErrorOutOfMemory;
        return hipErrorOutOfMemory;
    }
    memset(*p, 0x7f, bytes);  // not zero: a read before the first write shows
    return hipSuccess;
}
hipError_t hipFree(void* p)
{
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = (void*)0x1;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = malloc(1);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// launches run when they are made: every wait is already satisfied
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.f;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    free(e);
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError()
{
    const hipError_t e = g_last;
    g_last = hipSuccess;
    return e;
}
const char* hipGetErrorString(hipError_t e) { return e == hipErrorOutOfMemory ? "out of memory (host build)" : "host build"; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorInvalidValue; }
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, hipGraphNode_t*, char*, size_t) { return hipErrorInvalidValue; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
