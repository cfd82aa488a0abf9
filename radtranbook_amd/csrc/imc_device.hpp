// imc_device.hpp -- device-side building blocks shared by the IMC kernels:
// counter-based Philox4x32-10 streams, uniform conversions, constants and
// wave64 helpers for CDNA (gfx950).
//
// Random-stream layout (replaces the reference's hidden per-thread MT19937 of
// IMC/IMC2D.py:697-707,756-765 with a counter-based generator so any history
// can be replayed from its counter):
//   key = { lo32(seed), hi32(seed) }
//   ctr = { lo32(idx), (hi32(idx) & 0xFFFFFF) | tag << 24, step, block }
// A block yields 4 words = 2 slots (a,b); slot q is block q/2, half q%2.
//   FP64 uniform: ((a >> 5) * 2^26 + (b >> 6)) * 2^-53 in [0,1)
//   FP32 uniform: (a >> 8) * 2^-24                     in [0,1)
// Philox4x32-10 is the SC'11 generator of Salmon et al.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace imc {

constexpr double kC = 29.98;     // cm/ns (IMC/IMC2D.py:48)
constexpr double kA = 0.01372;   // GJ/(cm^3 keV^4) (IMC/IMC2D.py:49)
constexpr double kPi = 3.14159265358979323846;

enum Tag : uint32_t {
  kTagTransport = 1,
  kTagEmission = 2,
  kTagSource = 3,
  kTagComb = 4,
  kTagIC = 5,
  kTagBoundary0 = 8,  // + side
};

struct Key {
  uint32_t k0, k1;
};

__host__ __device__ inline Key make_key(uint64_t seed) { return Key{(uint32_t)seed, (uint32_t)(seed >> 32)}; }

struct Ctr {
  uint32_t c0, c1, c2;  // c3 is the block number
};

__host__ __device__ inline Ctr make_ctr(uint32_t tag, uint64_t idx, uint32_t step) {
  return Ctr{(uint32_t)idx, ((uint32_t)(idx >> 32) & 0x00FFFFFFu) | (tag << 24), step};
}

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox10(Ctr c, uint32_t block, Key k) {
  uint32_t x0 = c.c0, x1 = c.c1, x2 = c.c2, x3 = block;
  uint32_t k0 = k.k0, k1 = k.k1;
  // The key is uniform (a kernel argument).  Left visible, the compiler
  // hoists the 20 round keys out of a tracker's loop and, short of scalar
  // registers, parks them in vector-register lanes: two v_readlane and an
  // extra xor per round.  Made opaque here, the round keys are rebuilt with
  // scalar adds at every call and each xor takes its key as a scalar operand.
  asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // full 32x32 -> 64 products: one v_mad_u64_u32 each instead of a
    // mul_lo/mul_hi pair
    uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    // three-input xor in one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
    uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), x1, k0, 0x96);
    uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), x3, k1, 0x96);
    x0 = n0;
    x1 = (uint32_t)p1;
    x2 = n2;
    x3 = (uint32_t)p0;
  }
  return U4{x0, x1, x2, x3};
}

__device__ __forceinline__ double u01_d(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ float u01_f(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }

// Uniform from slot q of a stream (one Philox evaluation).
__device__ __forceinline__ double slot_d(Ctr c, Key k, uint32_t q) {
  U4 u = philox10(c, q >> 1, k);
  return (q & 1u) ? u01_d(u.z, u.w) : u01_d(u.x, u.y);
}

// Both slots of block b.
struct Pair {
  double u0, u1;
};
__device__ __forceinline__ Pair block_d(Ctr c, Key k, uint32_t b) {
  U4 u = philox10(c, b, k);
  return Pair{u01_d(u.x, u.y), u01_d(u.z, u.w)};
}

struct QuadF {
  float u0, u1, u2, u3;
};
// Four 24-bit uniforms from one block (FP32 tracking stream).
__device__ __forceinline__ QuadF block4_f(Ctr c, Key k, uint32_t b) {
  U4 u = philox10(c, b, k);
  return QuadF{u01_f(u.x), u01_f(u.y), u01_f(u.z), u01_f(u.w)};
}

// FP32 tracking arithmetic on the hardware transcendental units (about 1 ulp
// each): a/b as a times the reciprocal, and -ln(x) from log2.  (HIP's
// __fdividef is an IEEE division and __logf the full-range logf.)
__device__ __forceinline__ float fdiv_fast(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float neg_ln_fast(float x) { return -0.693147180559945f * __builtin_amdgcn_logf(x); }

// First m in [0, n-1] with u < cum[m] for a non-decreasing running sum cum,
// n - 1 when there is none (the index a left-to-right scan stops at).
__device__ __forceinline__ int first_above(const double* __restrict__ cum, int n, double u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (u < cum[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Element idx of an array whose base pointer is uniform across the wave,
// addressed as base + idx * sizeof(T) computed in the offset type Off.  With
// a 32-bit Off the load or store uses the scalar-base + 32-bit vector-offset
// form, with no 64-bit address arithmetic in the vector ALU; the caller
// guarantees idx * sizeof(T) < 2^32 then.
template <typename Off, typename T>
__device__ __forceinline__ T& elem(T* base, Off idx) {
  using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *reinterpret_cast<T*>(reinterpret_cast<Byte*>(base) + (Off)(idx * (Off)sizeof(T)));
}
// Largest element count whose byte offsets fit 32 bits for 8-byte elements.
constexpr int64_t kNarrowMax = (int64_t)(0xFFFFFFFFull / 8);

// wave64 helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  return (__lane_id() == 0) ? 0ull : ((~0ull) >> (64 - __lane_id()));
}

// Pickup order of the persistent trackers.  The queue hands out batches of
// `batch` consecutive particles (consecutive particles share cells, which is
// what keeps the LDS tally caches hitting).  Pools are laid out cell by
// cell, so a batch from an optically thick region is a batch of long
// histories; handed out late, it keeps its wave busy long after the rest of
// the device has drained.  With `order` set, queue batch b is particle chunk
// order[b] (chunk size 2^shift == batch), and the host sorts the chunks by
// expected history length, longest first.  Only the order of pickup changes:
// a particle keeps its index and so its random stream.
struct QueueMap {
  const int32_t* order;  // chunk permutation, or null for pool order
  int64_t len;           // queue positions: chunks << shift, or n
  int shift;             // log2 of the chunk size
};
inline QueueMap pool_order_queue(int64_t n) { return QueueMap{nullptr, n, 0}; }

// Wave-local batch of particles; one global atomic per batch.
struct WaveQueue {
  int64_t qbase = 0, qend = 0;
  bool exhausted = false;
};
// Called by every lane of the wave when `need` (the ballot of lanes without
// a history) is non-zero.  Hands the next particles of the wave's batch to
// the needy lanes in lane order, fetching a new batch of `batch` queue
// positions when the current one is used up.  Returns the particle index a
// needy lane takes, or -1.
__device__ __forceinline__ int64_t wave_take(WaveQueue& W, bool needy, uint64_t need, unsigned long long* next,
                                             int batch, const QueueMap& Q, int64_t n) {
  if (W.qbase >= W.qend && !W.exhausted) {
    unsigned long long b = 0;
    if (__lane_id() == 0) b = atomicAdd(next, (unsigned long long)batch);
    b = __shfl(b, 0, 64);
    if ((int64_t)b >= Q.len) {
      W.exhausted = true;
    } else {
      const int64_t base = Q.order ? ((int64_t)Q.order[b >> Q.shift] << Q.shift) : (int64_t)b;
      W.qbase = base;
      W.qend = (base + batch < n) ? base + batch : n;
    }
  }
  int64_t k = -1;
  const int64_t avail = W.qend - W.qbase;
  if (avail > 0) {
    const int64_t rank = __popcll(need & lanemask_lt());
    if (needy && rank < avail) k = W.qbase + rank;
    const int64_t take = __popcll(need);
    W.qbase += take < avail ? take : avail;
  }
  return k;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// Default number of private copies of per-cell histograms (tallies, census
// energy): workgroup b updates copy b % copies, so the global atomics that
// miss the LDS cache spread over copies (and XCDs).
constexpr int kTallyCopies = 8;
constexpr int kMaxTallyCopies = 64;

// Sum of v over each run of equal `key` in consecutive lanes of the wave.
// Every lane of the wave must call it.  The first lane of a run receives the
// run total and returns true.
template <typename T>
__device__ __forceinline__ bool wave_run_sum(int key, T& v) {
  const int lane = __lane_id();
  int prev = __shfl_up(key, 1, 64);
  bool head = (lane == 0) || (prev != key);
  uint64_t heads = __ballot(head);
  uint64_t above = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
  int run_end = above ? (int)__builtin_ctzll(above) : 64;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_down(v, o, 64);
    if (lane + o < run_end) v += t;
  }
  return head;
}

// Per-workgroup tally cache in LDS: an open-addressing hash of cell -> FP64
// sum.  A slot is claimed by the first cell that lands on it and keeps it
// for the kernel's lifetime; adds to cached cells are LDS atomics, adds that
// find no slot within kLdsProbes fall back to the caller's global atomic.
// Global atomics on gfx950 cap near 2.4e10/s device-wide, while the tally
// traffic of an IMC step concentrates on a few hundred hot cells, so the
// cache removes most global atomics.
constexpr int kLdsSlots = 1024;
constexpr int kLdsProbes = 4;

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

template <int SLOTS = kLdsSlots>
__device__ __forceinline__ void lds_tally_init(int* key, double* val) {
  static_assert((SLOTS & (SLOTS - 1)) == 0, "slot count must be a power of two");
  for (int s = threadIdx.x; s < SLOTS; s += blockDim.x) {
    key[s] = -1;
    val[s] = 0.0;
  }
}

// Returns false if the value still has to go to global memory.
template <int SLOTS = kLdsSlots>
__device__ __forceinline__ bool lds_tally_add(int* key, double* val, int cell, double v) {
  uint32_t s = ((uint32_t)cell * 2654435761u) >> (32 - ilog2(SLOTS));
#pragma unroll
  for (int p = 0; p < kLdsProbes; ++p) {
    int k = __hip_atomic_load(key + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == -1) {
      int expected = -1;
      __hip_atomic_compare_exchange_strong(key + s, &expected, cell, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
      k = (expected == -1) ? cell : expected;
    }
    if (k == cell) {
      __hip_atomic_fetch_add(val + s, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return true;
    }
    s = (s + 1) & (SLOTS - 1);
  }
  return false;
}

// After a __syncthreads(): add every cached cell into dst (global).
template <int SLOTS = kLdsSlots>
__device__ __forceinline__ void lds_tally_flush(const int* key, const double* val, double* dst) {
  for (int s = threadIdx.x; s < SLOTS; s += blockDim.x) {
    int k = key[s];
    if (k >= 0) {
      double v = val[s];
      if (v != 0.0) atomic_add_f64(dst + k, v);
    }
  }
}

// Per-wave tally cache in LDS: each wave owns W direct-mapped slots of
// (cell, FP64 sum, owner lane).  A cell that lands on an occupied slot
// evicts the holder's sum to global memory and takes the slot, so the cache
// follows the cells in current use whatever order the histories arrive in.
// Only the owning wave touches its slots.  Lanes that land on one slot take
// turns: each round every waiting lane writes its id to `own`, the lane whose
// id survives updates the slot, and the round ends at a wave-wide ballot, so
// one round's updates are complete before the next round's claims (the loop
// exit is uniform; a per-lane exit would let the compiler move the update
// past the other lanes' claims).  The slots are volatile LDS (address space
// 3) so each read-back is a real ds_read in program order.
typedef __attribute__((address_space(3))) volatile int lds_vint;
typedef __attribute__((address_space(3))) volatile double lds_vdouble;

template <int W>
struct WaveTally {
  static_assert((W & (W - 1)) == 0 && W >= 64, "slot count must be a power of two >= 64");
  static constexpr int kBytes = W * (4 + 8 + 4);
  lds_vdouble* val;
  lds_vint* key;
  lds_vint* own;

  // lds: this wave's kBytes of a __shared__ array (8-byte aligned)
  __device__ __forceinline__ explicit WaveTally(char* lds)
      : val((lds_vdouble*)(lds)), key((lds_vint*)(lds + W * 8)), own((lds_vint*)(lds + W * 12)) {
    for (int s = __lane_id(); s < W; s += 64) {
      key[s] = -1;
      val[s] = 0.0;
    }
  }

  // Called by any subset of the wave's lanes.
  __device__ __forceinline__ void add(int cell, double v, double* __restrict__ dst) {
    const uint32_t s = ((uint32_t)cell * 2654435761u) >> (32 - ilog2(W));
    const int lane = __lane_id();
    bool pending = true;
    do {
      if (pending) own[s] = lane;
      if (pending && own[s] == lane) {
        const int k = key[s];
        if (k == cell) {
          val[s] = val[s] + v;
        } else {
          if (k >= 0) atomic_add_f64(dst + k, val[s]);
          key[s] = cell;
          val[s] = v;
        }
        pending = false;
      }
    } while (__ballot(pending) != 0);
  }

  // Every lane of the wave: add the resident sums to dst.
  __device__ __forceinline__ void flush(double* __restrict__ dst) {
    for (int s = __lane_id(); s < W; s += 64) {
      const int k = key[s];
      if (k >= 0) {
        const double v = val[s];
        if (v != 0.0) atomic_add_f64(dst + k, v);
      }
    }
  }
};

// FP32 / FP64 pair type of a tracking precision.
template <typename R>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// Sum the private copies of a per-cell histogram.
template <typename T, typename U>
__global__ void k_reduce_copies(int64_t n, int copies, const T* __restrict__ part, U* __restrict__ out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  U a = 0;
  for (int k = 0; k < copies; ++k) a += (U)part[(int64_t)k * n + c];
  out[c] = a;
}

}  // namespace imc
