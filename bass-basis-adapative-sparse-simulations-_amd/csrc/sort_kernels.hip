// Stable LSD radix sort of uint64 keys with complex128 payload, and the
// run-length merge that follows it (src/core/sparse_state.py:127-172).
//
// One-sweep design (one read of the keys for every digit histogram, then one
// kernel per digit pass; Adinets & Merrill, "Onesweep", 2022):
//   k_sort_hist   every pass's digit histogram from one read of the keys
//                 (striped LDS histograms, per-block partial counts summed
//                 by k_sort_hist_sum)
//   k_sort_sweep  per pass, tiles of 8192 entries taken in dispatch order:
//                 stable in-tile ranks (wave ballots on the digit bits, running
//                 per-wave digit counts in LDS), per-digit decoupled look-back
//                 over the preceding tiles, then the tile is staged in LDS in
//                 digit order so each digit run is written out contiguously
//                 (keys first, then amplitudes through the same LDS buffer).
// An odd pass count rotates through a third buffer so the result ends in place.
// Digits are 8 bits wide (10 on request: fewer passes, but on MI355X a 10-bit
// pass takes about 1.85x an 8-bit one -- more LDS per block and a quarter of
// the digit-run length -- so 8 bits win even where they cost an extra pass).
// Positions and look-back counts are 32-bit: n < 2^30.
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = NT * SORT_ITEMS;  // the merge's tile

int sort_tiles(int64_t n) { return (int)((n + SORT_TILE - 1) / SORT_TILE); }

// 8192-entry tiles: on MI355X a 56-bit sort of 2^27 records takes 10.4 ms with
// them against 11.4 ms with 4096-entry tiles (longer digit runs per tile, i.e.
// fewer and longer write streams) and 10.8-11.0 ms with 6144-entry tiles
constexpr int SW_NT = 1024;   // threads per sweep block
constexpr int SW_ITEMS = 8;   // entries per thread
constexpr int SW_LB = 8;      // look-back words read per round trip
constexpr int SW_TILE = SW_NT * SW_ITEMS;
constexpr int HS_NT = 512;    // histogram blocks
constexpr int HS_U = 8;       // keys per thread per round
constexpr int HS_KEYS = HS_NT * HS_U;
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

// Per-block digit histograms of every pass from one read of the keys:
// part[block][p * 2^RB + digit].  LDS counters are striped 4 ways (8-bit
// digits) so lanes with equal digits mostly hit different banks; each block
// stores its counts plainly (no device-scope atomics on shared bins).
template <int RB>
__global__ __launch_bounds__(HS_NT) void k_sort_hist(const uint64_t* __restrict__ keys, int64_t n, SortPasses sp,
                                                      unsigned* __restrict__ part) {
  constexpr int R = 1 << RB;
  constexpr int ST = RB == 8 ? 4 : 1;
  __shared__ unsigned s_h[SORT_MAX_PASSES * R * ST];
  const int np = sp.np;
  for (int i = threadIdx.x; i < np * R * ST; i += HS_NT) s_h[i] = 0u;
  __syncthreads();
  const unsigned stripe = threadIdx.x % ST;
  const int64_t stride = (int64_t)gridDim.x * HS_KEYS;
  for (int64_t b = (int64_t)blockIdx.x * HS_KEYS; b < n; b += stride) {
    uint64_t k[HS_U];
#pragma unroll
    for (int u = 0; u < HS_U; ++u) {
      const int64_t i = b + (int64_t)u * HS_NT + threadIdx.x;
      k[u] = keys[i < n ? i : 0];
    }
#pragma unroll
    for (int u = 0; u < HS_U; ++u) {
      if (b + (int64_t)u * HS_NT + threadIdx.x < n) {
        for (int p = 0; p < np; ++p) {
          const unsigned d = (unsigned)((k[u] >> sp.shift[p]) & ((1ull << sp.width[p]) - 1ull));
          atomicAdd(&s_h[(p * R + d) * ST + stripe], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < np * R; i += HS_NT) {
    unsigned c = 0;
#pragma unroll
    for (int t = 0; t < ST; ++t) c += s_h[i * ST + t];
    part[(size_t)blockIdx.x * np * R + i] = c;
  }
}

// hist[i] = sum over blocks of part[block][i], i < cols: 64 columns per
// block, the rows split over 16 thread groups (several loads in flight per
// thread), group sums combined in LDS.
constexpr int HSUM_COLS = 64, HSUM_GROUPS = 16;

__global__ __launch_bounds__(HSUM_COLS* HSUM_GROUPS) void k_sort_hist_sum(const unsigned* __restrict__ part,
                                                                          int blocks, int cols,
                                                                          unsigned* __restrict__ hist) {
  __shared__ unsigned s_p[HSUM_GROUPS][HSUM_COLS];
  const int c = threadIdx.x % HSUM_COLS, g = threadIdx.x / HSUM_COLS;
  const int i = blockIdx.x * HSUM_COLS + c;
  unsigned acc = 0;
  if (i < cols) {
#pragma unroll 4
    for (int b = g; b < blocks; b += HSUM_GROUPS) acc += part[(size_t)b * cols + i];
  }
  s_p[g][c] = acc;
  __syncthreads();
  if (g == 0 && i < cols) {
    unsigned t = 0;
#pragma unroll
    for (int k = 0; k < HSUM_GROUPS; ++k) t += s_p[k][c];
    hist[i] = t;
  }
}

// One digit pass.  status: this pass's look-back words (zero on entry, tiles x
// 2^RB); status_next: the next pass's, cleared here row by row (no tile of
// the next pass reads them before this launch ends).
template <int RB, int NTHR, int ITEMS>
__global__ __launch_bounds__(NTHR) void k_sort_sweep(const uint64_t* __restrict__ kin,
                                                       const double2* __restrict__ vin, int64_t n, int shift,
                                                       int width, const unsigned* __restrict__ hist,
                                                       unsigned* status, unsigned* __restrict__ status_next,
                                                       unsigned* ticket, uint64_t* __restrict__ kout,
                                                       double2* __restrict__ vout, int* err) {
  constexpr int R = 1 << RB;
  constexpr int WV = NTHR / 64;
  constexpr int TILE = NTHR * ITEMS;
  constexpr int DPT = R > NTHR ? R / NTHR : 1;  // digits handled per thread
  // staging buffer of at most 64 KB: all keys, then the amplitudes in HALVES rounds
  constexpr int HALVES = TILE * 16 > 65536 ? 2 : 1;
  constexpr int HT = TILE / HALVES;             // staged amplitudes per round
  static_assert(TILE * 8 <= HT * 16 && HT % NTHR == 0, "staging buffer shape");
  __shared__ double2 s_buf[HT];                 // staging: keys, then amplitudes
  __shared__ unsigned short s_wc[WV][R];        // per-wave digit counts, then per-wave prefixes
  __shared__ unsigned s_lstart[R];              // tile-local start of each digit
  __shared__ unsigned s_gofs[R];                // global minus tile-local position, per digit
  __shared__ int s_scan[WV];
  __shared__ unsigned s_tile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (int i = tid; i < WV * R; i += NTHR) (&s_wc[0][0])[i] = 0;
  __syncthreads();
  const unsigned tile = s_tile;
  const int64_t t0 = (int64_t)tile * TILE;
  const int cnt = (int)((n - t0) < TILE ? (n - t0) : TILE);
  for (int d = tid; d < R; d += NTHR) status_next[(size_t)tile * R + d] = 0u;

  // load: wave w owns entries [w*64*ITEMS, (w+1)*64*ITEMS) of the tile, item
  // it of lane l is entry w*64*ITEMS + it*64 + l (so (it, l) order is input order)
  const int64_t wb = t0 + (int64_t)w * 64 * ITEMS;
  const unsigned dmask = (1u << width) - 1u;
  uint64_t kk[ITEMS];
  double vr[ITEMS], vi[ITEMS];  // separate doubles: a double2 array is kept in scratch
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = wb + it * 64 + lane;
    const int64_t c = i < n ? i : t0;
    kk[it] = kin[c];
    const double2 v = vin[c];
    vr[it] = v.x;
    vi[it] = v.y;
  }
  // stable rank of each item among the wave's items of the same digit
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned rank[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
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
      for (int ww = 0; ww < WV; ++ww) {
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
  const int lex = block_excl_scan<NTHR>(mine, total, s_scan);
  const int gex = block_excl_scan<NTHR>(hsum, total, s_scan);
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
  __syncthreads();  // tile-local digit starts and per-wave prefixes
  // stage the keys in digit order before the look-back: the key registers are
  // free while the predecessor words are in flight
  uint64_t* s_k = (uint64_t*)s_buf;
  unsigned lp[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const unsigned d = (unsigned)(kk[it] >> shift) & dmask;
    lp[it] = s_lstart[d] + s_wc[w][d] + rank[it];
    if (wb + it * 64 + lane < n) s_k[lp[it]] = kk[it];
  }
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = tid * DPT + j;
    if (d >= R) continue;
    unsigned excl = 0;
    if (tile > 0) {
      // read the words of the SW_LB nearest unresolved predecessors at once
      // (one memory round trip), then consume them in distance order up to
      // the first inclusive prefix or the first unpublished word
      long t = (long)tile - 1;
      unsigned spins = 0;
      while (true) {
        unsigned v[SW_LB];
#pragma unroll
        for (int k = 0; k < SW_LB; ++k) v[k] = (t - k >= 0) ? sw_load(&status[(size_t)(t - k) * R + d]) : SW_INC;
        int stop = SW_LB;
        bool inc = false;
#pragma unroll
        for (int k = 0; k < SW_LB; ++k) {
          if (k < stop) {
            const unsigned f = v[k] & ~SW_CNT;
            if (f == 0u) {
              stop = k;
            } else {
              excl += v[k] & SW_CNT;
              if (f == SW_INC) {
                inc = true;
                stop = k;
              }
            }
          }
        }
        if (inc) break;
        t -= stop;  // all SW_LB consumed, or resume at the first unpublished word
        if (stop == 0) {
          if (++spins > (1u << 24)) {
            set_err(err, ERR_LOOKBACK);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      sw_store(&status[(size_t)tile * R + d], SW_INC | (excl + tot[j]));
    }
    s_gofs[d] = gstart[j] + excl - lstart[j];
  }
  __syncthreads();

  // write each digit run of keys contiguously
  unsigned gp[ITEMS];
#pragma unroll
  for (int m = 0; m < ITEMS; ++m) {
    const int j = m * NTHR + tid;
    if (j < cnt) {
      const uint64_t key = s_k[j];
      gp[m] = s_gofs[(unsigned)(key >> shift) & dmask] + (unsigned)j;
      kout[gp[m]] = key;
    }
  }
#pragma unroll
  for (int h = 0; h < HALVES; ++h) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const unsigned q = lp[it] - (unsigned)(h * HT);
      if (wb + it * 64 + lane < n && q < (unsigned)HT) s_buf[q] = make_double2(vr[it], vi[it]);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < ITEMS; ++m) {
      const int j = m * NTHR + tid;
      if ((m * NTHR) / HT == h && j < cnt) vout[gp[m]] = s_buf[j - h * HT];
    }
  }
}

static int g_sort_digit_bits = 0;  // 0: automatic

int sort_digit_bits_get() { return g_sort_digit_bits; }

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
                     uint64_t* tk2, double2* tv2, unsigned* hist, unsigned* status, unsigned* ticket, int* err,
                     hipStream_t s) {
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
  // histogram blocks: at most 512 and one per HS_KEYS keys, so the partial
  // counts (blocks x passes x 2^RB words: <= 7 n bytes, or 28 KB when
  // n < HS_KEYS) fit in the scratch record buffers, which no pass has written
  // yet (24 n bytes at tk, inside a scratch region with 128 KB to spare)
  int64_t hb = n / HS_KEYS;
  if (hb > 512) hb = 512;
  if (hb < 1) hb = 1;
  unsigned* part = (unsigned*)tk;
  const int cols = passes * R;
  hipLaunchKernelGGL(k_sort_hist<RB>, dim3((unsigned)hb), dim3(HS_NT), 0, s, keys, n, sp, part);
  hipLaunchKernelGGL(k_sort_hist_sum, dim3((unsigned)((cols + HSUM_COLS - 1) / HSUM_COLS)),
                     dim3(HSUM_COLS * HSUM_GROUPS), 0, s, part, (int)hb, cols, hist);
  // buffers of each pass: an even count ping-pongs between (keys, vals) and
  // T1, an odd count >= 3 runs keys -> T1 -> T2 -> T1 ... -> keys, so the
  // result always ends in (keys, vals) without a copy (one pass: one copy)
  uint64_t* kb[3] = {keys, tk, tk2};
  double2* vb[3] = {vals, tv, tv2};
  unsigned* st[2] = {status, status + (size_t)tiles * R};
  int src = 0;
  for (int p = 0; p < passes; ++p) {
    int dst;
    if (passes & 1)
      dst = (p == passes - 1 && passes > 1) ? 0 : (src == 1 ? 2 : 1);
    else
      dst = src == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_sort_sweep<RB, SW_NT, SW_ITEMS>), dim3((unsigned)tiles), dim3(SW_NT), 0, s, kb[src],
                       vb[src], n, sp.shift[p], sp.width[p], hist + (size_t)p * R, st[p & 1], st[(p + 1) & 1],
                       ticket + p, kb[dst], vb[dst], err);
    src = dst;
  }
  if (src != 0) {
    (void)hipMemcpyAsync(keys, kb[src], sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(vals, vb[src], sizeof(double2) * (size_t)n, hipMemcpyDeviceToDevice, s);
  }
}

// hist: SORT_MAX_PASSES * 1024 words, ticket: SORT_MAX_PASSES words and
// status: sort_status_bytes(n), all zero on entry except status past its
// first tiles x 2^RB words (cleared pass by pass).
int launch_radix_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, uint64_t* tk, double2* tv,
                      uint64_t* tk2, double2* tv2, unsigned int* hist, unsigned int* status, unsigned int* ticket,
                      int* err, hipStream_t s, int digit_bits) {
  if (n <= 1) return 0;
  int passes = 0;
  int rb = sort_plan(key_bits, &passes);
  if (digit_bits == 8 || digit_bits == 10) {
    rb = digit_bits;
    passes = (key_bits + rb - 1) / rb;
  }
  if (rb == 10)
    run_sort<10>(keys, vals, n, key_bits, passes, tk, tv, tk2, tv2, hist, status, ticket, err, s);
  else
    run_sort<8>(keys, vals, n, key_bits, passes, tk, tv, tk2, tv2, hist, status, ticket, err, s);
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
