// imc_util.hpp -- host-side helpers shared by the IMC C-ABI translation units:
// error reporting, device buffers, deterministic reductions and scans.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include <cstdio>
#include <string>

namespace imc {

// Status codes returned across the C-ABI (0 = ok).
enum Status : int {
  kOk = 0,
  kErrHip = -1,
  kErrArg = -2,
  kErrCapacity = -3,
  kErrState = -4,
};

void set_error(const char* fmt, ...);
const char* get_error();

#define IMC_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      ::imc::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
      return ::imc::kErrHip;                                                           \
    }                                                                                  \
  } while (0)

#define IMC_REQUIRE(cond, code, ...)  \
  do {                                \
    if (!(cond)) {                    \
      ::imc::set_error(__VA_ARGS__);  \
      return (code);                  \
    }                                 \
  } while (0)

// Device allocation that remembers its size; grows by reallocation.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  int64_t cap = 0;
  // A buffer that has to grow again gets 25% headroom, so sizes that creep
  // up from step to step (particle counts) do not reallocate every step.
  int64_t next_cap(int64_t n) const { return cap > 0 ? std::max<int64_t>(n, cap + cap / 4) : n; }
  int reserve(int64_t n) {
    if (n <= cap) return kOk;
    n = next_cap(n);
    T* q = nullptr;
    IMC_HIP(hipMalloc(&q, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    if (p) IMC_HIP(hipFree(p));
    p = q;
    cap = n;
    return kOk;
  }
  // grow keeping the first `keep` elements
  int grow(int64_t n, int64_t keep, hipStream_t s) {
    if (n <= cap) return kOk;
    n = next_cap(n);
    T* q = nullptr;
    IMC_HIP(hipMalloc(&q, sizeof(T) * (size_t)n));
    if (p && keep > 0) IMC_HIP(hipMemcpyAsync(q, p, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice, s));
    IMC_HIP(hipStreamSynchronize(s));
    if (p) IMC_HIP(hipFree(p));
    p = q;
    cap = n;
    return kOk;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Page-locked host staging for per-cell transfers.  hipMemcpyAsync from
// pageable memory runs through the runtime's own staging with host-side
// waits (about 0.1 ms per 320 KB array measured); copying into a
// page-locked buffer first costs one host memcpy and lets the DMA run
// asynchronously.  A caller must synchronise its stream before the staged
// region is reused.
struct PinnedBuf {
  char* p = nullptr;
  char* dev = nullptr;  // the same memory as seen by kernels
  size_t cap = 0;
  int reserve(size_t n) {
    if (n <= cap) return kOk;
    if (p) IMC_HIP(hipHostFree(p));
    p = dev = nullptr;
    IMC_HIP(hipHostMalloc((void**)&p, n, hipHostMallocMapped));
    IMC_HIP(hipHostGetDevicePointer((void**)&dev, p, 0));
    cap = n;
    return kOk;
  }
  double* slot(int64_t n, int k) const { return (double*)p + (size_t)k * (size_t)n; }
  double* dslot(int64_t n, int k) const { return (double*)dev + (size_t)k * (size_t)n; }
  void release() {
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    cap = 0;
  }
};

// Up to four copies in one kernel launch (4-byte granularity), between
// device buffers and the kernel-visible address of page-locked staging: a
// kernel reads or writes host memory directly over PCIe, which for the
// per-cell arrays of a step costs less than queueing DMA copies.
struct CopyJob {
  const void* src;
  void* dst;
  int64_t bytes;
};
int copy_jobs(const CopyJob* jobs, int njobs, hipStream_t s);

// Kernel-visible address of [p, p + bytes) when it lies inside one block of
// imc_host_alloc, else null.
void* host_device_alias(const void* p, size_t bytes);

// Up to eight device ranges cleared in one kernel launch (8-byte
// granularity).  A kernel rather than hipMemsetAsync: one launch for all of a
// step's tally buffers, and safe under graph capture.
struct ZeroJob {
  void* p;
  int64_t bytes;
};
int zero_jobs(const ZeroJob* jobs, int njobs, hipStream_t s);

// Deterministic "canonical" sum: 1024-element chunks summed left to right,
// repeated on the chunk sums until one value is left.  The result lands in
// *out (device).  `scratch` must hold ceil(n/1024) + ceil(n/2^20) + 2 doubles.
int canonical_sum(const double* x, int64_t n, double* out, double* scratch, hipStream_t s);
int64_t canonical_sum_scratch(int64_t n);

// Fast deterministic sum for long particle arrays: a fixed grid of 1024
// workgroups accumulates grid-strided partial sums and reduces them in a
// fixed tree; the result depends only on n, not on scheduling.  Used where
// the oracle does not need the canonical order (weight floor, census totals).
// `scratch` must hold 1024 doubles.
int tree_sum(const double* x, int64_t n, double* out, double* scratch, hipStream_t s);
// The same with the length read from device memory when the kernels run.
int tree_sum_dev(const double* x, const int64_t* n_dev, double* out, double* scratch, hipStream_t s);

// Exclusive scan of int64 counts; writes n+1 offsets (last = total).
// `scratch` must hold scan_scratch(n) int64.
int exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, int64_t* scratch, hipStream_t s);
int64_t scan_scratch(int64_t n);
int exclusive_scan_i32_to_i64(const int32_t* in, int64_t n, int64_t* out, int64_t* scratch, hipStream_t s);

int launch_blocks(int64_t n, int block);

// Counting sort of per-chunk keys in [0, kChunkKeys) into a permutation of
// chunk ids, key 0 first; order within a key is arbitrary.  The keys are
// written by a kernel that also leaves per-workgroup key histograms in
// `hist` (kChunkGrid x kChunkKeys ints); chunk c belongs to workgroup
// (c / kChunkBlock) % grid of a grid-strided loop over chunks with
// kChunkBlock threads per workgroup.  `grid` <= kChunkGrid.
constexpr int kChunkKeys = 64;
constexpr int kChunkGrid = 128;
constexpr int kChunkBlock = 256;
int chunk_grid(int64_t nchunk);
int chunk_order(const uint8_t* key, int64_t nchunk, int grid, int32_t* hist, int32_t* order, hipStream_t s);
// The same for ceil(n / 2^shift) chunks with n read from device memory when
// the kernels run (the key kernel ran with that n as well).
int chunk_order_dev(const uint8_t* key, const int64_t* n_dev, int shift, int grid, int32_t* hist, int32_t* order,
                    hipStream_t s);

// Block-local key histogram for the key kernels: zero it, count, publish.
__device__ __forceinline__ void chunk_hist_zero(int* h) {
  if (threadIdx.x < kChunkKeys) h[threadIdx.x] = 0;
  __syncthreads();
}
__device__ __forceinline__ void chunk_hist_publish(const int* h, int32_t* hist) {
  __syncthreads();
  if (threadIdx.x < kChunkKeys) hist[blockIdx.x * kChunkKeys + threadIdx.x] = h[threadIdx.x];
}

}  // namespace imc
