// imc_util.hip -- error reporting, deterministic reductions and scans.
#include <cstdarg>
#include <cstdio>

#include <algorithm>
#include <map>
#include <mutex>

#include "../../include/imc_hip.h"
#include "imc_util.hpp"

namespace imc {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_err; }

// Blocks of imc_host_alloc: host base -> (bytes, kernel-visible base).
static std::mutex g_host_mu;
static std::map<uintptr_t, std::pair<size_t, char*>> g_host_blocks;

void* host_device_alias(const void* p, size_t bytes) {
  if (!p) return nullptr;
  const uintptr_t a = (uintptr_t)p;
  std::lock_guard<std::mutex> lock(g_host_mu);
  auto it = g_host_blocks.upper_bound(a);
  if (it == g_host_blocks.begin()) return nullptr;
  --it;
  if (a + bytes > it->first + it->second.first) return nullptr;
  return it->second.second + (a - it->first);
}

int chunk_grid(int64_t nchunk) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kChunkGrid, (nchunk + kChunkBlock - 1) / kChunkBlock));
}

// Offsets: off[b][key] = (chunks with a smaller key) + (chunks with this key
// in workgroups before b).  One workgroup; histograms staged in LDS.
__global__ __launch_bounds__(1024) void k_chunk_scan(int32_t* __restrict__ hist, int grid) {
  __shared__ int h[kChunkGrid * kChunkKeys];
  __shared__ int tot[kChunkKeys];
  const int n = grid * kChunkKeys;
  for (int a = threadIdx.x; a < n; a += blockDim.x) h[a] = hist[a];
  __syncthreads();
  if (threadIdx.x < kChunkKeys) {
    int run = 0;
    for (int b = 0; b < grid; ++b) {
      const int v = h[b * kChunkKeys + threadIdx.x];
      h[b * kChunkKeys + threadIdx.x] = run;
      run += v;
    }
    tot[threadIdx.x] = run;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int k = 0; k < kChunkKeys; ++k) {
      const int v = tot[k];
      tot[k] = run;
      run += v;
    }
  }
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += blockDim.x) hist[a] = h[a] + tot[a % kChunkKeys];
}

__global__ __launch_bounds__(kChunkBlock) void k_chunk_scatter(const uint8_t* __restrict__ key, int64_t nchunk,
                                                               const int32_t* __restrict__ off,
                                                               int32_t* __restrict__ order,
                                                               const int64_t* __restrict__ n_dev, int shift) {
  __shared__ int cur[kChunkKeys];
  if (n_dev) nchunk = (*n_dev + (1ll << shift) - 1) >> shift;
  if (threadIdx.x < kChunkKeys) cur[threadIdx.x] = off[blockIdx.x * kChunkKeys + threadIdx.x];
  __syncthreads();
  for (int64_t c = blockIdx.x * (int64_t)kChunkBlock + threadIdx.x; c < nchunk;
       c += (int64_t)gridDim.x * kChunkBlock)
    order[atomicAdd(&cur[key[c]], 1)] = (int32_t)c;
}

int chunk_order(const uint8_t* key, int64_t nchunk, int grid, int32_t* hist, int32_t* order, hipStream_t s) {
  IMC_REQUIRE(nchunk >= 0 && nchunk <= INT32_MAX, kErrArg, "chunk_order: %lld chunks", (long long)nchunk);
  IMC_REQUIRE(grid >= 1 && grid <= kChunkGrid, kErrArg, "chunk_order: grid %d", grid);
  if (nchunk == 0) return kOk;
  k_chunk_scan<<<1, 1024, 0, s>>>(hist, grid);
  IMC_HIP(hipGetLastError());
  k_chunk_scatter<<<grid, kChunkBlock, 0, s>>>(key, nchunk, hist, order, nullptr, 0);
  IMC_HIP(hipGetLastError());
  return kOk;
}

int chunk_order_dev(const uint8_t* key, const int64_t* n_dev, int shift, int grid, int32_t* hist, int32_t* order,
                    hipStream_t s) {
  IMC_REQUIRE(grid >= 1 && grid <= kChunkGrid, kErrArg, "chunk_order_dev: grid %d", grid);
  k_chunk_scan<<<1, 1024, 0, s>>>(hist, grid);
  IMC_HIP(hipGetLastError());
  k_chunk_scatter<<<grid, kChunkBlock, 0, s>>>(key, 0, hist, order, n_dev, shift);
  IMC_HIP(hipGetLastError());
  return kOk;
}

int launch_blocks(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > 0x7fffffff) b = 0x7fffffff;
  return (int)b;
}

// ---------------------------------------------------------------------------
// kernel copies
// ---------------------------------------------------------------------------
struct CopyJobs4 {
  CopyJob j[4];
  int n;
};

__global__ void k_copy_jobs(CopyJobs4 J) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int q = 0; q < J.n; ++q) {
    const uint32_t* src = (const uint32_t*)J.j[q].src;
    uint32_t* dst = (uint32_t*)J.j[q].dst;
    const int64_t words = J.j[q].bytes >> 2;
    for (int64_t i = i0; i < words; i += stride) dst[i] = src[i];
  }
}

int copy_jobs(const CopyJob* jobs, int njobs, hipStream_t s) {
  CopyJobs4 J{};
  int64_t most = 0;
  J.n = 0;
  for (int q = 0; q < njobs; ++q) {
    if (!jobs[q].src || !jobs[q].dst || jobs[q].bytes <= 0) continue;
    IMC_REQUIRE((jobs[q].bytes & 3) == 0, kErrArg, "copy_jobs: size not a multiple of 4");
    J.j[J.n++] = jobs[q];
    most = std::max<int64_t>(most, jobs[q].bytes >> 2);
  }
  if (J.n == 0) return kOk;
  int blocks = (int)std::min<int64_t>((most + 255) / 256, 1024);
  k_copy_jobs<<<blocks, 256, 0, s>>>(J);
  IMC_HIP(hipGetLastError());
  return kOk;
}

struct ZeroJobs8 {
  ZeroJob j[8];
  int n;
};

__global__ void k_zero_jobs(ZeroJobs8 J) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int q = 0; q < J.n; ++q) {
    unsigned long long* p = (unsigned long long*)J.j[q].p;
    const int64_t words = J.j[q].bytes >> 3;
    for (int64_t i = i0; i < words; i += stride) p[i] = 0ull;
  }
}

int zero_jobs(const ZeroJob* jobs, int njobs, hipStream_t s) {
  IMC_REQUIRE(njobs <= 8, kErrArg, "zero_jobs: more than eight ranges");
  ZeroJobs8 J{};
  int64_t most = 0;
  for (int q = 0; q < njobs; ++q) {
    if (!jobs[q].p || jobs[q].bytes <= 0) continue;
    IMC_REQUIRE((jobs[q].bytes & 7) == 0 && ((uintptr_t)jobs[q].p & 7) == 0, kErrArg,
                "zero_jobs: range not 8-byte aligned");
    J.j[J.n++] = jobs[q];
    most = std::max<int64_t>(most, jobs[q].bytes >> 3);
  }
  if (J.n == 0) return kOk;
  int blocks = (int)std::min<int64_t>((most + 255) / 256, 2048);
  k_zero_jobs<<<blocks, 256, 0, s>>>(J);
  IMC_HIP(hipGetLastError());
  return kOk;
}

// ---------------------------------------------------------------------------
// canonical sum
// ---------------------------------------------------------------------------
// One workgroup per 1024-element chunk: the chunk is staged through LDS
// with coalesced loads, then one lane adds it left to right (the canonical
// order; a lane walking global memory pays a load latency per element).
__global__ __launch_bounds__(256) void k_chunk_sum(const double* __restrict__ x, int64_t n,
                                                   double* __restrict__ out) {
  __shared__ double v[1024];
  const int64_t lo = (int64_t)blockIdx.x * 1024;
  const int m = (int)((n - lo) < 1024 ? (n - lo) : 1024);
  for (int i = threadIdx.x; i < m; i += 256) v[i] = x[lo + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll 16
    for (int i = 0; i < m; ++i) s += v[i];
    out[blockIdx.x] = s;
  }
}

int64_t canonical_sum_scratch(int64_t n) {
  int64_t tot = 2;
  while (n > 1) {
    n = (n + 1023) / 1024;
    tot += n;
  }
  return tot;
}

int canonical_sum(const double* x, int64_t n, double* out, double* scratch, hipStream_t s) {
  if (n <= 0) {
    IMC_HIP(hipMemsetAsync(out, 0, sizeof(double), s));
    return kOk;
  }
  const double* src = x;
  double* dst = scratch;
  int64_t m = n;
  while (true) {
    int64_t nc = (m + 1023) / 1024;
    double* target = (nc == 1) ? out : dst;
    k_chunk_sum<<<(unsigned)nc, 256, 0, s>>>(src, m, target);
    IMC_HIP(hipGetLastError());
    if (nc == 1) break;
    src = dst;
    dst += nc;
    m = nc;
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// tree sum
// ---------------------------------------------------------------------------
constexpr int kTreeBlocks = 1024;

__device__ __forceinline__ double block_sum_256(double v) {
  __shared__ double ws[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  return r;
}

__global__ __launch_bounds__(256) void k_tree_partial(const double* __restrict__ x, int64_t n,
                                                      const int64_t* __restrict__ n_dev,
                                                      double* __restrict__ part) {
  if (n_dev) n = *n_dev;
  double v = 0.0;
  const int64_t stride = (int64_t)kTreeBlocks * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  // four loads in flight, added in the same order as the one-at-a-time loop
  for (; i + 3 * stride < n; i += 4 * stride) {
    const double a = x[i], b = x[i + stride], c = x[i + 2 * stride], d = x[i + 3 * stride];
    v += a;
    v += b;
    v += c;
    v += d;
  }
  for (; i < n; i += stride) v += x[i];
  double r = block_sum_256(v);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void k_tree_final(const double* __restrict__ part, double* __restrict__ out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < kTreeBlocks; i += 256) v += part[i];
  double r = block_sum_256(v);
  if (threadIdx.x == 0) *out = r;
}

int tree_sum(const double* x, int64_t n, double* out, double* scratch, hipStream_t s) {
  k_tree_partial<<<kTreeBlocks, 256, 0, s>>>(x, n, nullptr, scratch);
  IMC_HIP(hipGetLastError());
  k_tree_final<<<1, 256, 0, s>>>(scratch, out);
  IMC_HIP(hipGetLastError());
  return kOk;
}

int tree_sum_dev(const double* x, const int64_t* n_dev, double* out, double* scratch, hipStream_t s) {
  k_tree_partial<<<kTreeBlocks, 256, 0, s>>>(x, 0, n_dev, scratch);
  IMC_HIP(hipGetLastError());
  k_tree_final<<<1, 256, 0, s>>>(scratch, out);
  IMC_HIP(hipGetLastError());
  return kOk;
}

// ---------------------------------------------------------------------------
// exclusive scan (int64), 256 threads x 8 items per block
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename TIn>
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const TIn* __restrict__ in, int64_t n,
                                                            int64_t* __restrict__ out, int64_t* __restrict__ sums) {
  __shared__ int64_t wsum[kScanThreads / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t run = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int64_t idx = base + q;
    int64_t x = idx < n ? (int64_t)in[idx] : 0;
    v[q] = run;
    run += x;
  }
  // inclusive wave scan of per-thread totals
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t inc = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int64_t wofs = 0;
  for (int w = 0; w < wid; ++w) wofs += wsum[w];
  int64_t excl = wofs + inc - run;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int64_t idx = base + q;
    if (idx < n) out[idx] = excl + v[q];
  }
  if (threadIdx.x == kScanThreads - 1 && sums) sums[blockIdx.x] = wofs + inc;
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ sums) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += sums[i / kScanTile];
}

__global__ void k_scan_total(const int64_t* __restrict__ excl, const int64_t* __restrict__ last_in_i64,
                             const int32_t* __restrict__ last_in_i32, int64_t n, int64_t* __restrict__ out_total) {
  int64_t last = n > 0 ? (last_in_i64 ? last_in_i64[n - 1] : (int64_t)last_in_i32[n - 1]) : 0;
  *out_total = (n > 0 ? excl[n - 1] : 0) + last;
}

int64_t scan_scratch(int64_t n) {
  int64_t tot = 1;
  while (n > kScanTile) {
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    tot += 2 * nb + 1;
    n = nb;
  }
  return tot;
}

template <typename TIn>
static int scan_impl(const TIn* in, int64_t n, int64_t* out, int64_t* scratch, hipStream_t s) {
  if (n <= 0) {
    IMC_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
    return kOk;
  }
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  int64_t* sums = nb > 1 ? scratch : nullptr;
  k_scan_tile<TIn><<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, out, sums);
  IMC_HIP(hipGetLastError());
  if (nb > 1) {
    int64_t* sums_scan = scratch + nb;
    int rc = scan_impl<int64_t>(sums, nb, sums_scan, scratch + 2 * nb + 1, s);
    if (rc) return rc;
    k_scan_add<<<launch_blocks(n, 256), 256, 0, s>>>(out, n, sums_scan);
    IMC_HIP(hipGetLastError());
  }
  return kOk;
}

int exclusive_scan_i64(const int64_t* in, int64_t n, int64_t* out, int64_t* scratch, hipStream_t s) {
  int rc = scan_impl<int64_t>(in, n, out, scratch, s);
  if (rc) return rc;
  k_scan_total<<<1, 1, 0, s>>>(out, in, nullptr, n, out + n);
  IMC_HIP(hipGetLastError());
  return kOk;
}

int exclusive_scan_i32_to_i64(const int32_t* in, int64_t n, int64_t* out, int64_t* scratch, hipStream_t s) {
  int rc = scan_impl<int32_t>(in, n, out, scratch, s);
  if (rc) return rc;
  k_scan_total<<<1, 1, 0, s>>>(out, nullptr, in, n, out + n);
  IMC_HIP(hipGetLastError());
  return kOk;
}

}  // namespace imc

extern "C" {

int imc_host_alloc(uint64_t bytes, void** out) {
  using namespace imc;
  IMC_REQUIRE(out && bytes > 0, kErrArg, "imc_host_alloc: bad arguments");
  *out = nullptr;
  void* p = nullptr;
  void* d = nullptr;
  IMC_HIP(hipHostMalloc(&p, (size_t)bytes, hipHostMallocMapped | hipHostMallocPortable));
  if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess || !d) {
    (void)hipHostFree(p);
    set_error("imc_host_alloc: no device address for the block");
    return kErrHip;
  }
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    g_host_blocks[(uintptr_t)p] = std::make_pair((size_t)bytes, (char*)d);
  }
  *out = p;
  return kOk;
}

int imc_host_free(void* p) {
  using namespace imc;
  IMC_REQUIRE(p, kErrArg, "imc_host_free: null pointer");
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    auto it = g_host_blocks.find((uintptr_t)p);
    IMC_REQUIRE(it != g_host_blocks.end(), kErrArg, "imc_host_free: not a block of imc_host_alloc");
    g_host_blocks.erase(it);
  }
  IMC_HIP(hipHostFree(p));
  return kOk;
}

}  // extern "C"
