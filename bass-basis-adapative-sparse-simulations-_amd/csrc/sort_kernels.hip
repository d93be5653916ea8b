// Stable LSD radix sort of uint64 keys with complex128 payload, and the
// run-length merge that follows it (src/core/sparse_state.py:127-172).
//
// One-sweep design (one read of the keys for every digit histogram, then one
// kernel per digit pass; Adinets & Merrill, "Onesweep", 2022):
//   k_sort_hist   every pass's digit histogram from one read of the keys
//                 (LDS histograms, one global add per bin and block)
//   k_sort_sweep  per pass, tiles of 4096 entries taken in dispatch order:
//                 stable in-tile ranks (wave ballots on the digit bits, running
//                 per-wave digit counts in LDS), per-digit decoupled look-back
//                 over the preceding tiles, then the tile is staged in LDS in
//                 digit order so each digit run is written out contiguously
//                 (keys first, then amplitudes through the same LDS buffer).
// Digits are 8 bits wide (10 on request: fewer passes, but on MI355X each
// 10-bit pass takes about 1.85x an 8-bit one -- one 88 KB block per CU and
// digit runs of ~4 records per tile -- so 8 bits win even where they cost
// an extra pass plus the copy back after an odd pass count).  Positions and
// look-back counts are 32-bit: n < 2^30.
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = NT * SORT_ITEMS;  // the merge's tile

int sort_tiles(int64_t n) { return (int)((n + SORT_TILE - 1) / SORT_TILE); }

constexpr int SW_NT = 512;
constexpr int SW_WAVES = SW_NT / 64;
constexpr int SW_ITEMS = 8;
constexpr int SW_TILE = SW_NT * SW_ITEMS;
constexpr unsigned SW_AGG = 1u << 30;  // status flag: tile aggregate published
constexpr unsigned SW_INC = 2u << 30;  // status flag: inclusive prefix published
constexpr unsigned SW_CNT = (1u << 30) - 1u;

int sweep_tiles(int64_t n) { return (int)((n + SW_TILE - 1) / SW_TILE); }

struct SortPasses {
  int np;
  int shift[SORT_MAX_PASSES];
  int width[SORT_MAX_PASSES];
};

__device__ __forceinline__ unsigned sw_load(unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void sw_store(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// hist[p * 2^RB + digit] += count, for every pass p.
template <int RB>
__global__ __launch_bounds__(NT) void k_sort_hist(const uint64_t* __restrict__ keys, int64_t n, SortPasses sp,
                                                   unsigned* __restrict__ hist) {
  constexpr int R = 1 << RB;
  constexpr int U = 4;  // keys per thread loaded together
  __shared__ unsigned s_h[SORT_MAX_PASSES * R];
  for (int i = threadIdx.x; i < sp.np * R; i += NT) s_h[i] = 0u;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * NT * U;
  for (int64_t b = (int64_t)blockIdx.x * NT * U; b < n; b += stride) {
    uint64_t k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + (int64_t)u * NT + threadIdx.x;
      k[u] = keys[i < n ? i : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (b + (int64_t)u * NT + threadIdx.x < n) {
        for (int p = 0; p < sp.np; ++p)
          atomicAdd(&s_h[p * R + (unsigned)((k[u] >> sp.shift[p]) & ((1ull << sp.width[p]) - 1ull))], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < sp.np * R; i += NT)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// One digit pass.  status: this pass's look-back words (zero on entry, tiles x
// 2^RB); status_next: the next pass's, cleared here row by row (no tile of
// the next pass reads them before this launch ends).
template <int RB>
__global__ __launch_bounds__(SW_NT) void k_sort_sweep(const uint64_t* __restrict__ kin,
                                                       const double2* __restrict__ vin, int64_t n, int shift,
                                                       int width, const unsigned* __restrict__ hist,
                                                       unsigned* status, unsigned* __restrict__ status_next,
                                                       unsigned* ticket, uint64_t* __restrict__ kout,
                                                       double2* __restrict__ vout, int* err) {
  constexpr int R = 1 << RB;
  constexpr int DPT = R > SW_NT ? R / SW_NT : 1;  // digits handled per thread
  __shared__ double2 s_buf[SW_TILE];              // staging: keys, then amplitudes
  __shared__ unsigned short s_wc[SW_WAVES][R];    // per-wave digit counts, then per-wave prefixes
  __shared__ unsigned s_lstart[R];                // tile-local start of each digit
  __shared__ unsigned s_gofs[R];                  // global minus tile-local position, per digit
  __shared__ int s_scan[SW_WAVES];
  __shared__ unsigned s_tile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (int i = tid; i < SW_WAVES * R; i += SW_NT) (&s_wc[0][0])[i] = 0;
  __syncthreads();
  const unsigned tile = s_tile;
  const int64_t t0 = (int64_t)tile * SW_TILE;
  const int cnt = (int)((n - t0) < SW_TILE ? (n - t0) : SW_TILE);
  for (int d = tid; d < R; d += SW_NT) status_next[(size_t)tile * R + d] = 0u;

  // load: wave w owns entries [w*512, w*512+512) of the tile, item it of lane l
  // is entry w*512 + it*64 + l (so (it, l) order is input order)
  const int64_t wb = t0 + (int64_t)w * 64 * SW_ITEMS;
  const unsigned dmask = (1u << width) - 1u;
  uint64_t kk[SW_ITEMS];
  double vr[SW_ITEMS], vi[SW_ITEMS];  // separate doubles: a double2 array is kept in scratch
#pragma unroll
  for (int it = 0; it < SW_ITEMS; ++it) {
    const int64_t i = wb + it * 64 + lane;
    const int64_t c = i < n ? i : t0;
    kk[it] = kin[c];
    const double2 v = vin[c];
    vr[it] = v.x;
    vi[it] = v.y;
  }
  // stable rank of each item among the wave's items of the same digit
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned rank[SW_ITEMS];
#pragma unroll
  for (int it = 0; it < SW_ITEMS; ++it) {
    const bool ok = wb + it * 64 + lane < n;
    const unsigned d = (unsigned)(kk[it] >> shift) & dmask;
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const unsigned long long bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const unsigned before = s_wc[w][d];
    __builtin_amdgcn_wave_barrier();
    rank[it] = before + (unsigned)__popcll(peers & lt);
    if (ok && (peers & lt) == 0ull) s_wc[w][d] = (unsigned short)(before + (unsigned)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  // per digit: tile count, per-wave exclusive prefixes, tile-local starts
  unsigned tot[DPT];
  int mine = 0;
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = tid * DPT + j;
    unsigned run = 0;
    if (d < R) {
#pragma unroll
      for (int ww = 0; ww < SW_WAVES; ++ww) {
        const unsigned c = s_wc[ww][d];
        s_wc[ww][d] = (unsigned short)run;
        run += c;
      }
    }
    tot[j] = run;
    mine += (int)run;
  }
  // global digit starts: exclusive scan of this pass's histogram
  unsigned hv[DPT];
  int hsum = 0;
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = tid * DPT + j;
    hv[j] = d < R ? hist[d] : 0u;
    hsum += (int)hv[j];
  }
  int total;
  const int lex = block_excl_scan<SW_NT>(mine, total, s_scan);
  const int gex = block_excl_scan<SW_NT>(hsum, total, s_scan);
  unsigned lacc = (unsigned)lex, gacc = (unsigned)gex;
  unsigned lstart[DPT], gstart[DPT];
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    lstart[j] = lacc;
    gstart[j] = gacc;
    lacc += tot[j];
    gacc += hv[j];
    const int d = tid * DPT + j;
    if (d < R) s_lstart[d] = lstart[j];
  }
  // decoupled look-back, one digit per thread: publish every aggregate first
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = tid * DPT + j;
    if (d < R) sw_store(&status[(size_t)tile * R + d], (tile == 0 ? SW_INC : SW_AGG) | tot[j]);
  }
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = tid * DPT + j;
    if (d >= R) continue;
    unsigned excl = 0;
    if (tile > 0) {
      long t = (long)tile - 1;
      unsigned spins = 0;
      while (t >= 0) {
        const unsigned sv = sw_load(&status[(size_t)t * R + d]);
        const unsigned f = sv & ~SW_CNT;
        if (f == 0u) {
          if (++spins > (1u << 24)) {
            set_err(err, ERR_LOOKBACK);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += sv & SW_CNT;
        if (f == SW_INC) break;
        --t;
      }
      sw_store(&status[(size_t)tile * R + d], SW_INC | (excl + tot[j]));
    }
    s_gofs[d] = gstart[j] + excl - lstart[j];
  }
  __syncthreads();

  // stage the keys in digit order, write each digit run contiguously
  uint64_t* s_k = (uint64_t*)s_buf;
  unsigned lp[SW_ITEMS];
#pragma unroll
  for (int it = 0; it < SW_ITEMS; ++it) {
    const unsigned d = (unsigned)(kk[it] >> shift) & dmask;
    lp[it] = s_lstart[d] + s_wc[w][d] + rank[it];
    if (wb + it * 64 + lane < n) s_k[lp[it]] = kk[it];
  }
  __syncthreads();
  unsigned gp[SW_ITEMS];
#pragma unroll
  for (int m = 0; m < SW_ITEMS; ++m) {
    const int j = m * SW_NT + tid;
    if (j < cnt) {
      const uint64_t key = s_k[j];
      gp[m] = s_gofs[(unsigned)(key >> shift) & dmask] + (unsigned)j;
      kout[gp[m]] = key;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SW_ITEMS; ++it)
    if (wb + it * 64 + lane < n) s_buf[lp[it]] = make_double2(vr[it], vi[it]);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < SW_ITEMS; ++m) {
    const int j = m * SW_NT + tid;
    if (j < cnt) vout[gp[m]] = s_buf[j];
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

static int g_sort_digit_bits = 0;  // 0: automatic

int sort_digit_bits_set(int bits) {
  const int old = g_sort_digit_bits;
  g_sort_digit_bits = bits;
  return old;
}

// Digit width and pass count for keys of key_bits bits: 8-bit digits unless
// 10 were requested with sort_digit_bits_set.
int sort_plan(int key_bits, int* passes) {
  const int rb = g_sort_digit_bits == 10 ? 10 : 8;
  *passes = (key_bits + rb - 1) / rb;
  return rb;
}

size_t sort_status_bytes(int64_t n) { return (size_t)2 * (size_t)(sweep_tiles(n) + 1) * 1024 * 4; }

template <int RB>
static void run_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, int passes, uint64_t* tk, double2* tv,
                     unsigned* hist, unsigned* status, unsigned* ticket, int* err, hipStream_t s) {
  constexpr int R = 1 << RB;
  SortPasses sp{};
  sp.np = passes;
  int sh = 0;
  for (int p = 0; p < passes; ++p) {
    sp.width[p] = key_bits / passes + (p < key_bits % passes ? 1 : 0);
    sp.shift[p] = sh;
    sh += sp.width[p];
  }
  const int tiles = sweep_tiles(n);
  int64_t hb = (n + NT * 16 - 1) / (NT * 16);
  if (hb > 1024) hb = 1024;
  hipLaunchKernelGGL(k_sort_hist<RB>, dim3((unsigned)hb), dim3(NT), 0, s, keys, n, sp, hist);
  const uint64_t* ki = keys;
  const double2* vi = vals;
  uint64_t* ko = tk;
  double2* vo = tv;
  unsigned* st[2] = {status, status + (size_t)tiles * R};
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_sort_sweep<RB>, dim3((unsigned)tiles), dim3(SW_NT), 0, s, ki, vi, n, sp.shift[p],
                       sp.width[p], hist + (size_t)p * R, st[p & 1], st[(p + 1) & 1], ticket + p, ko, vo, err);
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
}

// hist: SORT_MAX_PASSES * 1024 words, ticket: SORT_MAX_PASSES words and
// status: sort_status_bytes(n), all zero on entry except status past its
// first tiles x 2^RB words (cleared pass by pass).
int launch_radix_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, uint64_t* tk, double2* tv,
                      unsigned int* hist, unsigned int* status, unsigned int* ticket, int* err, hipStream_t s) {
  if (n <= 1) return 0;
  int passes = 0;
  const int rb = sort_plan(key_bits, &passes);
  if (rb == 10)
    run_sort<10>(keys, vals, n, key_bits, passes, tk, tv, hist, status, ticket, err, s);
  else
    run_sort<8>(keys, vals, n, key_bits, passes, tk, tv, hist, status, ticket, err, s);
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
