// Stable LSD radix sort of uint64 keys with complex128 payload, and the
// run-length merge that follows it (src/core/sparse_state.py:127-172).
//
// 8-bit digits, P = ceil(key_bits / 8) passes, each a reduce-then-scan:
//   k_sort_count   per tile of 2048 entries, LDS histogram of the digit
//                  -> counts[digit][tile]                      (reads 8 B/entry)
//   k_sort_rowscan one block per digit: exclusive scan of its row over tiles,
//                  row total to totals[digit]
//   k_sort_scatter per tile: stable in-tile rank of every entry (wave peer
//                  masks from 8 ballots + running per-digit counts across the
//                  tile's rounds), position = digit start + row prefix + rank,
//                  scatter key + payload                  (reads 24 B, writes 24 B)
// No inter-block waits: every kernel finishes independently.
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = NT * SORT_ITEMS;
constexpr int RADIX = 256;

int sort_tiles(int64_t n) { return (int)((n + SORT_TILE - 1) / SORT_TILE); }

__global__ __launch_bounds__(NT) void k_sort_count(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                    unsigned int* __restrict__ counts, int tiles) {
  __shared__ unsigned int s_h[RADIX];
  s_h[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = (int64_t)blockIdx.x * SORT_TILE + (int64_t)it * NT + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(unsigned int)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * tiles + blockIdx.x] = s_h[threadIdx.x];
}

// One block per digit: in-place exclusive scan of counts[digit][0:tiles].
__global__ __launch_bounds__(NT) void k_sort_rowscan(unsigned int* __restrict__ counts, int tiles,
                                                      unsigned int* __restrict__ totals) {
  __shared__ int s_scan[WAVES];
  unsigned int* row = counts + (size_t)blockIdx.x * tiles;
  unsigned int carry = 0;
  for (int t0 = 0; t0 < tiles; t0 += NT) {
    const int t = t0 + threadIdx.x;
    const unsigned int v = t < tiles ? row[t] : 0u;
    int tot;
    const unsigned int ex = (unsigned int)block_excl_scan((int)v, tot, s_scan);
    if (t < tiles) row[t] = carry + ex;
    carry += (unsigned int)tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(NT) void k_sort_scatter(const uint64_t* __restrict__ kin,
                                                      const double2* __restrict__ vin, int64_t n, int shift,
                                                      const unsigned int* __restrict__ counts, int tiles,
                                                      const unsigned int* __restrict__ totals,
                                                      uint64_t* __restrict__ kout, double2* __restrict__ vout) {
  __shared__ unsigned int s_cnt[RADIX];
  __shared__ unsigned int s_wc[WAVES][RADIX];
  __shared__ unsigned int s_off[RADIX];
  __shared__ int s_scan[WAVES];
  const unsigned int tile = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    // digit start (scan of the totals) + this tile's row prefix
    int tot;
    const unsigned int gstart = (unsigned int)block_excl_scan((int)totals[threadIdx.x], tot, s_scan);
    s_off[threadIdx.x] = gstart + counts[(size_t)threadIdx.x * tiles + tile];
    s_cnt[threadIdx.x] = 0;
    for (int ww = 0; ww < WAVES; ++ww) s_wc[ww][threadIdx.x] = 0;
  }
  __syncthreads();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int it = 0; it < SORT_ITEMS; ++it) {
    const int64_t i = (int64_t)tile * SORT_TILE + (int64_t)it * NT + threadIdx.x;
    const bool ok = i < n;
    const uint64_t kk = ok ? kin[i] : 0ull;
    const double2 vv = ok ? vin[i] : make_double2(0.0, 0.0);
    const unsigned int d = (unsigned int)(kk >> shift) & 255u;
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
      unsigned int r = s_off[d] + s_cnt[d] + in_wave;
      for (int ww = 0; ww < w; ++ww) r += s_wc[ww][d];
      kout[r] = kk;
      vout[r] = vv;
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
}

__global__ __launch_bounds__(NT) void k_copy_kv(const uint64_t* __restrict__ ks, const double2* __restrict__ vs,
                                                 int64_t n, uint64_t* __restrict__ kd, double2* __restrict__ vd) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

// counts: RADIX * tiles words; totals: RADIX words.
int launch_radix_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, uint64_t* tk, double2* tv,
                      unsigned int* totals, unsigned int* counts, unsigned int* /*unused*/, int* /*err*/,
                      hipStream_t s) {
  if (n <= 1) return 0;
  const int passes = (key_bits + 7) / 8;
  const int tiles = sort_tiles(n);
  const uint64_t* ki = keys;
  const double2* vi = vals;
  uint64_t* ko = tk;
  double2* vo = tv;
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_sort_count, dim3(tiles), dim3(NT), 0, s, ki, n, 8 * p, counts, tiles);
    hipLaunchKernelGGL(k_sort_rowscan, dim3(RADIX), dim3(NT), 0, s, counts, tiles, totals);
    hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(NT), 0, s, ki, vi, n, 8 * p, counts, tiles, totals, ko,
                       vo);
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
  if (threadIdx.x < 64) {
    const unsigned long long e = lookback_prefix_wave(lb, tile, ttot, err);
    if (threadIdx.x == 0) s_base = e;
  }
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
