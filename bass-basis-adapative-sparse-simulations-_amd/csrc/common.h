// Shared device primitives for the BASS sparse-state kernels (gfx950, wave64).
//
// State layout in HBM (structure of arrays):
//   keys : uint64_t[n]   basis-state bit strings, qubit q = bit q
//   amps : double2[n]    complex128 amplitudes, (re, im) interleaved
//
// All floating-point work is FP64 and is compiled with -ffp-contract=off; the
// only fused operations are the explicit ones in cmul below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bass {

constexpr uint64_t EMPTY = 0xFFFFFFFFFFFFFFFFull;  // hash sentinel, never a key for N <= 63
constexpr int NT = 256;                             // threads per block (4 waves)
constexpr int WAVES = NT / 64;
constexpr int NTB = 1024;                           // single-block reduction / scan kernels
constexpr int WAVESB = NTB / 64;

// Error bits reported through the per-call status word.
enum : int {
  ERR_PROBE = 1,      // hash table probe sequence exhausted (table full)
  ERR_MISSING = 2,    // a key inserted by the build kernel was not found
  ERR_LOOKBACK = 4,   // decoupled look-back spin bound exceeded
  ERR_OVERFLOW = 8,   // output capacity exceeded
  ERR_SELECT = 16,    // top-k selection bookkeeping inconsistent
};

// Complex product x * y with one fused multiply-add per component:
//   re = fma(xr, yr, -(xi * yi)),  im = fma(xr, yi, xi * yr).
// This is the rounding of numpy's complex128 multiply loops (AVX2/AVX-512
// builds) and of LLVM's fast-math contraction of (ac - bd, ad + bc) in the
// reference's Numba kernels; the operand order matters for the imaginary part.
__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
  return make_double2(__builtin_fma(x.x, y.x, -(x.y * y.y)), __builtin_fma(x.x, y.y, x.y * y.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ bool cnonzero(double2 a) { return a.x != 0.0 || a.y != 0.0; }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Fibonacci hashing of a 64-bit key into a power-of-two table: the top
// log2(cap) bits of key * 2^64/phi.  shift = 64 - log2(cap).
__device__ __forceinline__ uint64_t hslot(uint64_t key, int shift) {
  return (key * 0x9E3779B97F4A7C15ull) >> shift;
}

__device__ __forceinline__ void set_err(int* err, int bit) { atomicOr(err, bit); }

// ---------------------------------------------------------------- scans ----
// Inclusive scan over the 64 lanes of a wave (all lanes active) with DPP
// moves: shifts by 1, 2, 4, 8 inside each 16-lane row, then row 0's and
// row 1's last lanes broadcast into the rows after them (row_bcast:15 into
// rows 1 and 3, row_bcast:31 into rows 2 and 3).  Lanes a move does not
// reach add 0.  No LDS traffic, unlike lane shuffles.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// Block-wide exclusive scan for NTHR threads.  lds must hold NTHR/64 ints.
// Returns the exclusive prefix; total receives the block sum.
template <int NTHR = NT>
__device__ __forceinline__ int block_excl_scan(int v, int& total, int* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NTHR / 64; i++) {
    int s = lds[i];
    pre += (i < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return pre + inc - v;
}

// Exclusive scan of R values per thread in item-major order (all threads'
// item 0, then item 1, ...: the order of R consecutive block_excl_scan calls)
// with one barrier pair.  lds must hold R * NTHR/64 ints.
template <int R, int NTHR = NT>
__device__ __forceinline__ void block_excl_scan_items(const int (&v)[R], int (&off)[R], int& total, int* lds) {
  constexpr int NW = NTHR / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    inc[r] = wave_incl_scan(v[r]);
    if (lane == 63) lds[r * NW + w] = inc[r];
  }
  __syncthreads();
  int base = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const int s = lds[r * NW + i];
      pre += (i < w) ? s : 0;
      tot += s;
    }
    off[r] = base + pre + inc[r] - v[r];
    base += tot;
  }
  __syncthreads();
  total = base;
}

// Deterministic block sum of doubles (fixed shuffle tree, then waves in order).
// lds must hold NTHR/64 doubles.  Result valid in every thread.
template <int NTHR = NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NTHR / 64; i++) s += lds[i];
  __syncthreads();
  return s;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ------------------------------------------------------ decoupled look-back --
// Status word per tile: bits 63..62 flag (0 = not ready, 1 = aggregate,
// 2 = inclusive prefix), bits 61..0 value.  Tiles take their id from an atomic
// ticket so every predecessor of a tile is already running when the tile
// starts; a predecessor publishes its aggregate before it waits on anyone, so
// each wait below ends.  A spin bound turns a broken invariant into an error
// bit instead of a hang.
//
// The status words carry only counts -- no tile reads data another tile wrote
// in the same launch -- so they are relaxed agent-scope atomics.  (A release
// store at agent scope would write back the whole XCD L2 on every publish.)
constexpr unsigned long long LB_MASK = (1ull << 62) - 1;

__device__ __forceinline__ unsigned long long lb_load(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by ONE thread of the tile.  Returns the exclusive prefix of `count`
// over all tiles before `tile`.
__device__ inline unsigned long long lookback_prefix(unsigned long long* status, unsigned int tile,
                                                     unsigned long long count, int* err) {
  if (tile == 0) {
    lb_store(&status[0], (2ull << 62) | count);
    return 0;
  }
  lb_store(&status[tile], (1ull << 62) | count);
  unsigned long long excl = 0;
  long t = (long)tile - 1;
  unsigned int spins = 0;
  while (t >= 0) {
    unsigned long long s = lb_load(&status[t]);
    unsigned int flag = (unsigned int)(s >> 62);
    if (flag == 0) {
      if (++spins > (1u << 26)) {
        set_err(err, ERR_LOOKBACK);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += s & LB_MASK;
    if (flag == 2) break;
    --t;
  }
  lb_store(&status[tile], (2ull << 62) | (excl + count));
  return excl;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Wave-parallel look-back: all 64 lanes of ONE wave call this together.  Each
// probe reads the status words of the 64 nearest unresolved predecessors at
// once; if every word up to the nearest inclusive prefix is published, their
// values are summed with one wave reduction, otherwise the wave waits and
// re-reads.  Returns the exclusive prefix of `count` in every lane.
__device__ inline unsigned long long lookback_prefix_wave(unsigned long long* status, unsigned int tile,
                                                          unsigned long long count, int* err) {
  const int lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0) lb_store(&status[0], (2ull << 62) | count);
    return 0;
  }
  if (lane == 0) lb_store(&status[tile], (1ull << 62) | count);
  unsigned long long excl = 0;
  long base = (long)tile - 1;
  unsigned int spins = 0;
  while (true) {
    const long t = base - lane;
    // before tile 0 reads as an inclusive prefix of 0
    const unsigned long long s = (t >= 0) ? lb_load(&status[t]) : (2ull << 62);
    const unsigned int flag = (unsigned int)(s >> 62);
    const unsigned long long incl = __ballot(flag == 2);
    const unsigned long long zero = __ballot(flag == 0);
    const int first = incl ? (__ffsll((long long)incl) - 1) : 64;
    const unsigned long long upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
    if (zero & upto) {
      if (++spins > (1u << 24)) {
        if (lane == 0) set_err(err, ERR_LOOKBACK);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += wave_sum_u64(lane <= first ? (s & LB_MASK) : 0ull);
    if (first < 64) break;
    base -= 64;
  }
  if (lane == 0) lb_store(&status[tile], (2ull << 62) | (excl + count));
  return excl;
}

// Input count of a counted operation: the device slot, clamped to [0, n_max]
// (the host's bound, which the grid and the buffers are sized for).  n_dev
// null: n_max is the count.
__device__ __forceinline__ int64_t dev_count(const int64_t* n_dev, int64_t n_max) {
  if (!n_dev) return n_max;
  const int64_t v = *n_dev;
  return v < 0 ? 0 : (v < n_max ? v : n_max);
}

// Tile id in dispatch order.
__device__ __forceinline__ unsigned int take_ticket(unsigned int* ticket, unsigned int* lds_slot) {
  if (threadIdx.x == 0) *lds_slot = atomicAdd(ticket, 1u);
  __syncthreads();
  unsigned int t = *lds_slot;
  __syncthreads();
  return t;
}

}  // namespace bass
