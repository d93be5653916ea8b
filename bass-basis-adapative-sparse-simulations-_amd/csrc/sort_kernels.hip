// Stable LSD radix sort of uint64 keys with complex128 payload, and the
// run-length merge that follows it (src/core/sparse_state.py:127-172).
//
// 8-bit digits, P = ceil(key_bits / 8) passes.  One kernel reads every key
// once and builds the global histograms of all P digits (8 B/entry).  Each
// pass is then a single kernel (Onesweep style): a tile of 2048 entries ranks
// its digits stably in LDS -- wave-level peer masks from 8 ballots, then a
// running per-digit count across the tile's rounds -- obtains its per-digit
// global offsets from a decoupled look-back over earlier tiles (one status
// word per tile and digit) and scatters key + payload (48 B/entry/pass).
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = NT * SORT_ITEMS;
constexpr int RADIX = 256;
constexpr unsigned int ST_MASK = (1u << 30) - 1u;

int sort_tiles(int64_t n) { return (int)((n + SORT_TILE - 1) / SORT_TILE); }

__global__ __launch_bounds__(NT) void k_sort_hist(const uint64_t* __restrict__ keys, int64_t n, int passes,
                                                   unsigned int* ghist) {
  __shared__ unsigned int s_h[8][RADIX];
  for (int j = threadIdx.x; j < 8 * RADIX; j += NT) (&s_h[0][0])[j] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    const uint64_t k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&s_h[p][(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < passes * RADIX; j += NT) {
    const unsigned int v = (&s_h[0][0])[j];
    if (v) atomicAdd(&ghist[j], v);
  }
}

__device__ __forceinline__ unsigned int st_load(unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_store(unsigned int* p, unsigned int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(NT) void k_sort_pass(const uint64_t* __restrict__ kin,
                                                   const double2* __restrict__ vin, int64_t n, int shift,
                                                   const unsigned int* __restrict__ ghist_pass,
                                                   uint64_t* __restrict__ kout, double2* __restrict__ vout,
                                                   unsigned int* status, unsigned int* ticket, int* err) {
  __shared__ unsigned int s_cnt[RADIX];
  __shared__ unsigned int s_wc[WAVES][RADIX];
  __shared__ unsigned int s_off[RADIX];
  __shared__ int s_scan[WAVES];
  __shared__ unsigned int s_tile;
  const unsigned int tile = take_ticket(ticket, &s_tile);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  s_cnt[threadIdx.x] = 0;
  for (int ww = 0; ww < WAVES; ++ww) s_wc[ww][threadIdx.x] = 0;
  __syncthreads();
  uint64_t kk[SORT_ITEMS];
  double2 vv[SORT_ITEMS];
  unsigned int rank[SORT_ITEMS];
  unsigned int dig[SORT_ITEMS];
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = (int64_t)tile * SORT_TILE + (int64_t)it * NT + threadIdx.x;
    const bool ok = i < n;
    kk[it] = ok ? kin[i] : 0ull;
    vv[it] = ok ? vin[i] : make_double2(0.0, 0.0);
    const unsigned int d = (unsigned int)(kk[it] >> shift) & 255u;
    dig[it] = d;
    // lanes of this wave holding the same digit
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const unsigned int in_wave = (unsigned int)__popcll(peers & lt_mask);
    if (ok && in_wave == 0) s_wc[w][d] = (unsigned int)__popcll(peers);
    __syncthreads();
    if (ok) {
      unsigned int r = s_cnt[d] + in_wave;
      for (int ww = 0; ww < w; ++ww) r += s_wc[ww][d];
      rank[it] = r;
    }
    __syncthreads();
    {
      const int d2 = threadIdx.x;  // NT == RADIX: one digit per thread
      unsigned int add = 0;
      for (int ww = 0; ww < WAVES; ++ww) {
        add += s_wc[ww][d2];
        s_wc[ww][d2] = 0;
      }
      s_cnt[d2] += add;
    }
    __syncthreads();
  }
  // per-digit look-back over earlier tiles; thread t owns digit t
  {
    const int d = threadIdx.x;
    const unsigned int cnt = s_cnt[d];
    unsigned int* row = status + (size_t)tile * RADIX;
    unsigned int excl = 0;
    if (tile == 0) {
      st_store(&row[d], (2u << 30) | cnt);
    } else {
      st_store(&row[d], (1u << 30) | cnt);
      long t = (long)tile - 1;
      unsigned int spins = 0;
      while (t >= 0) {
        const unsigned int sv = st_load(&status[(size_t)t * RADIX + d]);
        const unsigned int flag = sv >> 30;
        if (flag == 0) {
          if (++spins > (1u << 26)) {
            set_err(err, ERR_LOOKBACK);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += sv & ST_MASK;
        if (flag == 2) break;
        --t;
      }
      st_store(&row[d], (2u << 30) | (excl + cnt));
    }
    // global start of digit d: exclusive scan of the pass histogram
    int tot;
    const unsigned int gstart = (unsigned int)block_excl_scan((int)ghist_pass[d], tot, s_scan);
    s_off[d] = gstart + excl;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = (int64_t)tile * SORT_TILE + (int64_t)it * NT + threadIdx.x;
    if (i < n) {
      const unsigned int pos = s_off[dig[it]] + rank[it];
      kout[pos] = kk[it];
      vout[pos] = vv[it];
    }
  }
}

__global__ __launch_bounds__(NT) void k_copy_kv(const uint64_t* __restrict__ ks, const double2* __restrict__ vs,
                                                 int64_t n, uint64_t* __restrict__ kd, double2* __restrict__ vd) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

__global__ void k_zero_u32(unsigned int* p, int64_t m) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) p[i] = 0u;
}

int launch_radix_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, uint64_t* tk, double2* tv,
                      unsigned int* ghist, unsigned int* status, unsigned int* tickets, int* err,
                      hipStream_t s) {
  if (n <= 1) return 0;
  const int passes = (key_bits + 7) / 8;
  const int tiles = sort_tiles(n);
  hipLaunchKernelGGL(k_zero_u32, dim3(8), dim3(NT), 0, s, ghist, (int64_t)(8 * RADIX));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, tickets, (int64_t)16);
  int64_t hb = (n + NT - 1) / NT;
  if (hb > 1024) hb = 1024;
  hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)hb), dim3(NT), 0, s, keys, n, passes, ghist);
  const uint64_t* ki = keys;
  const double2* vi = vals;
  uint64_t* ko = tk;
  double2* vo = tv;
  for (int p = 0; p < passes; ++p) {
    int64_t zb = ((int64_t)tiles * RADIX + NT - 1) / NT;
    if (zb > 4096) zb = 4096;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)zb), dim3(NT), 0, s, status, (int64_t)tiles * RADIX);
    hipLaunchKernelGGL(k_sort_pass, dim3(tiles), dim3(NT), 0, s, ki, vi, n, 8 * p, ghist + p * RADIX, ko, vo,
                       status, tickets + p, err);
    const uint64_t* t0 = ki;
    const double2* t1 = vi;
    ki = ko;
    vi = vo;
    ko = (uint64_t*)t0;
    vo = (double2*)t1;
  }
  if (passes & 1) {
    int64_t cb = (n + NT - 1) / NT;
    if (cb > 8192) cb = 8192;
    hipLaunchKernelGGL(k_copy_kv, dim3((unsigned)cb), dim3(NT), 0, s, tk, tv, n, keys, vals);
  }
  return (int)hipGetLastError();
}

// Reduce-by-key of a sorted array: heads are entries whose key differs from
// the previous one; a head sums its run left to right (the reference's
// merge_sorted_kernel order).  Heads are compacted in order.
__global__ __launch_bounds__(NT) void k_merge(const uint64_t* __restrict__ x, const double2* __restrict__ a,
                                               int64_t n, uint64_t* __restrict__ xo, double2* __restrict__ ao,
                                               unsigned long long* lb, unsigned int* ticket, int64_t* n_out,
                                               int* err) {
  __shared__ int s_scan[WAVES];
  __shared__ unsigned int s_tile;
  __shared__ unsigned long long s_base;
  const unsigned int tile = take_ticket(ticket, &s_tile);
  bool head[SORT_ITEMS];
  int off[SORT_ITEMS];
  int rtot[SORT_ITEMS];
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = (int64_t)tile * SORT_TILE + (int64_t)it * NT + threadIdx.x;
    head[it] = i < n && (i == 0 || x[i] != x[i - 1]);
    int tot;
    off[it] = block_excl_scan(head[it] ? 1 : 0, tot, s_scan);
    rtot[it] = tot;
  }
  unsigned long long ttot = 0;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) ttot += (unsigned long long)rtot[it];
  if (threadIdx.x == 0) s_base = lookback_prefix(lb, tile, ttot, err);
  __syncthreads();
  unsigned long long rb = s_base;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    if (head[it]) {
      const int64_t i = (int64_t)tile * SORT_TILE + (int64_t)it * NT + threadIdx.x;
      const uint64_t key = x[i];
      double2 acc = a[i];
      for (int64_t j = i + 1; j < n && x[j] == key; ++j) acc = cadd(acc, a[j]);
      const unsigned long long pos = rb + (unsigned long long)off[it];
      xo[pos] = key;
      ao[pos] = acc;
    }
    rb += (unsigned long long)rtot[it];
  }
  if (tile == gridDim.x - 1 && threadIdx.x == 0) *n_out = (int64_t)(s_base + ttot);
}

int launch_merge(const uint64_t* x, const double2* a, int64_t n, uint64_t* xo, double2* ao,
                 unsigned long long* lb, unsigned int* ticket, int64_t* n_out, int* err, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_merge, dim3((unsigned)sort_tiles(n)), dim3(NT), 0, s, x, a, n, xo, ao, lb, ticket, n_out,
                     err);
  return (int)hipGetLastError();
}

}  // namespace bass
