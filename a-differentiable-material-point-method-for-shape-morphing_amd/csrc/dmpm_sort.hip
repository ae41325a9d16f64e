// Device order of the FP32 path's particles, computed on the GPU.
//
// Within each instance the particles are ordered by the T^3-cell block (tile)
// of their cell (blocks x-major, T = 2^tshift: 4 or 8), then by their rank
// inside their cell ("slot", stable in the current order), then by the cell's
// x-major index inside the block.  With T = 4 a 64-lane wave holds one
// particle of each of a block's 64 cells, so its LDS atomics hit 64 distinct
// nodes.  Two stable radix sorts give that order:
//   1. key (block, cell-in-block), value = current slot: particles grouped by
//      cell, each cell's particles in their current order;
//   2. the slot of each particle = its position minus the first position of
//      its cell (a max-scan of the cell heads); key (block, slot, cell),
//      value carried along: the final order.
// Blocks are numbered across instances (instance i's blocks from boff[i]), so
// instances keep their slot ranges.  The result is perm[d] = v1[d], the current
// slot of the particle that goes to slot d, and blk[d], its global block.
#include "dmpm_kernels.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/functional.hpp>

namespace dmpm {

namespace {

constexpr int kSortBlock = 256;

inline unsigned sort_blocks(long long n) { return (unsigned)((n + kSortBlock - 1) / kSortBlock); }

// instance owning slot d (poff ascending, B+1 entries)
__device__ __forceinline__ int instance_of_slot(const int* poff, int B, int d)
{
    int lo = 0, hi = B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (poff[mid] <= d) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// key (global block << 3 tshift | cell in block), value = slot
__global__ __launch_bounds__(kSortBlock) void k_key_cell(SortArgs a)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= a.N) return;
    const int in = instance_of_slot(a.poff, a.B, d);
    const StepParams& sp = a.sps[in];
    int cell[3], nb[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        // the cell of the position (relative to the instance's origin), clamped
        // into the grid; NaN goes to cell 0
        double u = floor((double(a.x[(long long)(FX + ax) * a.S + d]) - sp.min_point[ax]) * sp.inv_dx);
        if (!(u >= 0.0)) u = 0.0;
        if (u > double(sp.dim[ax] - 1)) u = double(sp.dim[ax] - 1);
        cell[ax] = int(u);
        nb[ax] = (sp.dim[ax] + (1 << a.tshift) - 1) >> a.tshift;
    }
    const int t = a.tshift, m = (1 << t) - 1;
    const long long block = a.boff[in] + ((long long)(cell[0] >> t) * nb[1] + (cell[1] >> t)) * nb[2] + (cell[2] >> t);
    const unsigned local = (unsigned)((cell[0] & m) << (2 * t) | (cell[1] & m) << t | (cell[2] & m));
    a.k0[d] = ((unsigned long long)block << (3 * t)) | local;
    a.v0[d] = d;
}

// head[d] = d where a new cell starts in the sorted keys, else 0
__global__ __launch_bounds__(kSortBlock) void k_heads(const unsigned long long* key, int N, int* head)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    head[d] = (d == 0 || key[d] != key[d - 1]) ? d : 0;
}

// key (global block << (16 + c) | slot << c | cell in block), c = 3 tshift bits
// of cell, slots beyond 2^16 - 1 clamped
__global__ __launch_bounds__(kSortBlock) void k_key_slot(const unsigned long long* k1, const int* first, const int* v1, int N, int c,
                                                        unsigned long long* k2, int* v2)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    const unsigned long long k = k1[d];
    int slot = d - first[d];
    if (slot > 65535) slot = 65535;
    k2[d] = ((k >> c) << (16 + c)) | ((unsigned long long)slot << c) | (k & ((1ull << c) - 1));
    v2[d] = v1[d];
}

__global__ __launch_bounds__(kSortBlock) void k_block_of(const unsigned long long* key, int N, int shift, int* blk)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    blk[d] = (int)(key[d] >> shift);
}

// dst[f S + d] = src[f S + perm[d]] for nf fields
template <typename T>
__global__ __launch_bounds__(kSortBlock) void k_gather(const T* src, T* dst, long long S, int N, int nf, const int* perm)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    const long long s = perm[d];
    for (int f = 0; f < nf; ++f) dst[(long long)f * S + d] = src[(long long)f * S + s];
}

// dst[f S + idx[d]] = src[f S + d] for nf fields
template <typename T>
__global__ __launch_bounds__(kSortBlock) void k_scatter(const T* src, T* dst, long long S, int N, int nf, const int* idx)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    const long long t = idx[d];
    for (int f = 0; f < nf; ++f) dst[(long long)f * S + t] = src[(long long)f * S + d];
}

// inv[a[d]] = d (a: a permutation of [0, N))
__global__ __launch_bounds__(kSortBlock) void k_invert(const int* a, int* inv, int N)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < N) inv[a[d]] = d;
}

__global__ __launch_bounds__(kSortBlock) void k_iota(int* a, int N)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < N) a[d] = d;
}

int bits_for(long long v)
{
    int b = 1;
    while (b < 40 && (1LL << b) <= v) ++b;
    return b;
}

}  // namespace

size_t sort_temp_bytes(int N, long long nblocks)
{
    size_t a = 0, b = 0;
    unsigned long long* kp = nullptr;
    int* ip = nullptr;
    // the widest keys of either tile (8^3: 9 bits of cell)
    const unsigned hi1 = 9 + bits_for(nblocks), hi2 = 25 + bits_for(nblocks);
    if (rocprim::radix_sort_pairs(nullptr, a, kp, kp, ip, ip, (size_t)N, 0u, hi1 > hi2 ? hi1 : hi2) != hipSuccess) return 0;
    if (rocprim::inclusive_scan(nullptr, b, ip, ip, (size_t)N, rocprim::maximum<int>()) != hipSuccess) return 0;
    return a > b ? a : b;
}

hipError_t launch_sort_order(const SortArgs& a, hipStream_t s)
{
    if (a.N <= 0) return hipSuccess;
    const unsigned nb = sort_blocks(a.N);
    const int c = 3 * a.tshift;
    const unsigned hi1 = c + bits_for(a.nblocks), hi2 = 16 + c + bits_for(a.nblocks);
    size_t bytes = a.temp_bytes;
    hipLaunchKernelGGL(k_key_cell, dim3(nb), dim3(kSortBlock), 0, s, a);
    hipError_t e = rocprim::radix_sort_pairs(a.temp, bytes, a.k0, a.k1, a.v0, a.v1, (size_t)a.N, 0u, hi1, s);
    if (e != hipSuccess) return e;
    // first position of each cell into v0 (free after the first sort); k_key_slot
    // then reads first[d] and writes v0[d] in the same thread
    hipLaunchKernelGGL(k_heads, dim3(nb), dim3(kSortBlock), 0, s, a.k1, a.N, a.head);
    bytes = a.temp_bytes;
    e = rocprim::inclusive_scan(a.temp, bytes, a.head, a.v0, (size_t)a.N, rocprim::maximum<int>(), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_key_slot, dim3(nb), dim3(kSortBlock), 0, s, a.k1, a.v0, a.v1, a.N, c, a.k0, a.v0);
    bytes = a.temp_bytes;
    e = rocprim::radix_sort_pairs(a.temp, bytes, a.k0, a.k1, a.v0, a.v1, (size_t)a.N, 0u, hi2, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_block_of, dim3(nb), dim3(kSortBlock), 0, s, a.k1, a.N, 16 + c, a.blk);
    return hipGetLastError();
}

template <typename T>
void launch_gather(const T* src, T* dst, long long S, int N, int nf, const int* perm, hipStream_t s)
{
    if (N <= 0 || nf <= 0) return;
    hipLaunchKernelGGL(k_gather<T>, dim3(sort_blocks(N)), dim3(kSortBlock), 0, s, src, dst, S, N, nf, perm);
}
template void launch_gather<float>(const float*, float*, long long, int, int, const int*, hipStream_t);
template void launch_gather<int>(const int*, int*, long long, int, int, const int*, hipStream_t);

template <typename T>
void launch_scatter(const T* src, T* dst, long long S, int N, int nf, const int* idx, hipStream_t s)
{
    if (N <= 0 || nf <= 0) return;
    hipLaunchKernelGGL(k_scatter<T>, dim3(sort_blocks(N)), dim3(kSortBlock), 0, s, src, dst, S, N, nf, idx);
}
template void launch_scatter<float>(const float*, float*, long long, int, int, const int*, hipStream_t);

void launch_invert(const int* a, int* inv, int N, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_invert, dim3(sort_blocks(N)), dim3(kSortBlock), 0, s, a, inv, N);
}

void launch_iota(int* a, int N, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(sort_blocks(N)), dim3(kSortBlock), 0, s, a, N);
}

}  // namespace dmpm
