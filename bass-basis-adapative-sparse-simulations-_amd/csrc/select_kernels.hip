// Exact top-k truncation by radix select on the IEEE bit pattern of |a|^2.
//
// Order: larger |a|^2 first; equal |a|^2 ordered by ascending logical key
// (the reference leaves ties to np.argpartition, whose choice is implementation
// defined; src/core/truncation.py:68-113, src/simulation/bass_simulator.py:899-916).
// Entries of a key-ordered state store permuted keys; a KeyMap maps them back
// for the tie-break.  |a|^2 >= 0, so its bit pattern orders like the value.
//
// A truncation problem (SelProblem) reads a sequence of tiles of at most
// SEG_ENTRIES entries: tile t holds cnt_t entries.  A dense array is the
// special case cnt_t = min(SEG_ENTRIES, n - t * SEG_ENTRIES); the segmented
// output of a gate supplies cnt_t per tile.  Per-tile scratch (SelTiles) is
// indexed by global tile, so several problems -- a batch of circuits -- run in
// one launch of every kernel: per-tile kernels find their problem from the
// tile index, per-problem kernels take one block each.  One problem by value
// (BATCH = false) or a device table of problems (BATCH = true) run the same
// code.  No kernel coordinates with other workgroups while it runs.
//
// Passes over the full input (24 B per entry):
//   digit 1 (sign + exponent) histogram: fused into the producing gate, or
//     k_hist_psum (16 B/entry);
//   k_extract1: read |a|^2 (16 B/entry); per tile, entries above the digit-1
//     bucket are counted and summed, entries inside it are kept as candidates
//     (in input order, per tile) together with the digit-2 histogram;
//   k_compact: read key + amplitude (24 B/entry), write the k kept entries in
//     input order at offsets scanned from the per-tile kept counts, divided by
//     sqrt(kept) when renormalising (24 B per kept entry).
// Everything else touches only candidates or per-tile scalars.
#include <string.h>

#include <mutex>
#include <unordered_map>

#include "common.h"
#include "extract_core.h"
#include "kernels.h"

namespace bass {


// Byte tables of P.km into P.ktab (the threads of one block share the
// entries; the map is staged in s_q, 64 bytes of LDS).  Uniform branch.
template <int NTHR>
__device__ __forceinline__ void build_ktab(const SelProblem& P, unsigned char* s_q) {
  if (!P.ktab || P.km.nbits == 0) return;
  const int nb = P.km.nbits;
  if (threadIdx.x < 64) s_q[threadIdx.x] = threadIdx.x < nb ? P.km.q[threadIdx.x] : 64;
  __syncthreads();
  for (int e = threadIdx.x; e < KTAB_WORDS; e += NTHR) {
    const int b = e >> 8, v = e & 255;
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int q = s_q[8 * b + i];
      if (((v >> i) & 1) && q < 64) r |= 1ull << q;
    }
    P.ktab[e] = r;
  }
}

// Find, in a 2048-bin histogram, the bucket holding the need-th largest element.
// Block of NTHR threads, HIST_BINS/NTHR bins per thread.  Results in *b and
// *above (count in higher buckets).  lds_scan: NTHR/64 ints.
template <int NTHR = NT>
__device__ void find_from_top(const unsigned int* hist, unsigned long long need, unsigned int* b,
                              unsigned long long* above, int* lds_scan) {
  constexpr int PER = HIST_BINS / NTHR;
  const int t = threadIdx.x;
  unsigned int c[PER];
  int v = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    c[j] = hist[t * PER + j];
    v += (int)c[j];
  }
  int tot;
  const int pre = block_excl_scan<NTHR>(v, tot, lds_scan);
  unsigned long long ab = (unsigned long long)(tot - pre - v);
  if (ab < need && need <= ab + (unsigned long long)v) {
    for (int j = PER - 1; j >= 0; --j) {
      if (ab + c[j] >= need) {
        *b = (unsigned int)(t * PER + j);
        *above = ab;
        break;
      }
      ab += c[j];
    }
  }
  __syncthreads();
}

// Same, counting from the smallest bucket (for the key tie-break).
template <int NTHR = NT>
__device__ void find_from_bottom(const unsigned int* hist, unsigned long long need, unsigned int* b,
                                 unsigned long long* below, int* lds_scan) {
  constexpr int PER = HIST_BINS / NTHR;
  const int t = threadIdx.x;
  unsigned int c[PER];
  int v = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    c[j] = hist[t * PER + j];
    v += (int)c[j];
  }
  int tot;
  unsigned long long bl = (unsigned long long)block_excl_scan<NTHR>(v, tot, lds_scan);
  if (bl < need && need <= bl + (unsigned long long)v) {
    for (int j = 0; j < PER; ++j) {
      if (bl + c[j] >= need) {
        *b = (unsigned int)(t * PER + j);
        *below = bl;
        break;
      }
      bl += c[j];
    }
  }
  __syncthreads();
}

// Fixed-order sum of a double array by one block (each thread adds its strided
// elements in order; loads are issued four at a time).
template <int NTHR = NT>
__device__ double block_sum_array(const double* v, int64_t m, double* lds) {
  double s = 0.0;
  int64_t j = threadIdx.x;
  for (; j + 3 * NTHR < m; j += 4 * NTHR) {
    const double a0 = v[j], a1 = v[j + NTHR], a2 = v[j + 2 * NTHR], a3 = v[j + 3 * NTHR];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; j < m; j += NTHR) s += v[j];
  return block_sum<NTHR>(s, lds);
}

constexpr int CMP_BUF = 1024;        // k_compact's staging of the written |a|^2 (doubles; two passes a tile)
constexpr int EX1_ABV = 1024;        // k_extract1's staging of above-bucket |a|^2 (doubles; two passes a tile)
constexpr int EX1_TILES = 4;         // tiles per k_extract1 block of a single problem
constexpr int EX1_TILES_MIN = 12;    // range for large problems (row counts <= 31 * SEL_TILE fit 16 bits)
constexpr int EX1_TILES_MAX = 31;

// ------------------------------------------------------ problem lookup ----
// Problem of a global tile: the last problem whose tile0 <= gt (problems are
// stored in tile order).  Uniform per block.
__device__ __forceinline__ int prob_of_tile(const SelProblem* probs, int nprob, int64_t gt) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile0 <= gt)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <bool BATCH>
__device__ __forceinline__ const SelProblem& tile_problem(const SelProblem* probs, int nprob, const SelProblem& one,
                                                          int64_t gt) {
  if constexpr (BATCH)
    return probs[prob_of_tile(probs, nprob, gt)];
  else
    return one;
}

template <bool BATCH>
__device__ __forceinline__ const SelProblem& block_problem(const SelProblem* probs, const SelProblem& one,
                                                           unsigned int b) {
  if constexpr (BATCH)
    return probs[b];
  else
    return one;
}

// entries of global tile gt of problem P, and their first input slot
__device__ __forceinline__ int tile_count(const SelProblem& P, int64_t gt) {
  if (P.seg_cnt) return P.seg_cnt[gt];
  const int64_t r = P.n_dense - (gt - P.tile0) * SEL_TILE;
  return r < SEL_TILE ? (int)r : SEL_TILE;
}
__device__ __forceinline__ int64_t tile_base(const SelProblem& P, int64_t gt) {
  return (P.seg_cnt ? gt : gt - P.tile0) * SEL_TILE;
}

// ------------------------------------------------------------ kernels -----
template <bool BATCH>
__global__ __launch_bounds__(NT) void k_hist_psum(const SelProblem* probs, int nprob, SelProblem one, SelTiles T) {
  __shared__ unsigned int s_hist[HIST_BINS];
  __shared__ double s_red[WAVES];
  const int64_t gt = blockIdx.x;
  const SelProblem& P = tile_problem<BATCH>(probs, nprob, one, gt);
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  __syncthreads();
  const int cnt = tile_count(P, gt);
  const int64_t base = tile_base(P, gt);
  double ps = 0.0;
  // all loads first (clamped to a valid slot), so they are in flight together
  double vr[SEL_ITEMS], vi[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int j = it * NT + threadIdx.x;
    const double2 v = P.a[base + (j < cnt ? j : 0)];
    vr[it] = v.x;
    vi[it] = v.y;
  }
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int j = it * NT + threadIdx.x;
    if (j < cnt) {
      const double p = cabs2(make_double2(vr[it], vi[it]));
      ps += p;
      atomicAdd(&s_hist[(unsigned int)(__double_as_longlong(p) >> 52) & (HIST_BINS - 1)], 1u);
    }
  }
  const double s = block_sum(ps, s_red);
  if (threadIdx.x == 0) T.tile_psum[gt] = s;
  __syncthreads();
  for (int b = threadIdx.x; b < HIST_BINS; b += NT)
    if (s_hist[b]) atomicAdd(&P.hist1[b], s_hist[b]);
}

// Digit 1: find the threshold bucket; also reduce the input's sum |a|^2.
__device__ __forceinline__ void find1_body(const SelProblem& P, const SelTiles& T) {
  __shared__ int s_scan[WAVESB];
  __shared__ double s_red[WAVESB];
  __shared__ unsigned int s_b;
  __shared__ unsigned long long s_above;
  __shared__ unsigned char s_q[64];
  // the select's zeroed header (digit-2 and wide histograms, state, tickets),
  // when the caller leaves it to this first kernel; the barrier below orders
  // it before this kernel's own writes to the state
  if (P.zero_hdr)
    for (int64_t i = threadIdx.x; i < P.zero_words; i += NTB) P.zero_hdr[i] = 0ull;
  build_ktab<NTB>(P, s_q);
  if (threadIdx.x == 0) {
    s_b = 0xFFFFFFFFu;
    s_above = 0;
  }
  // the bucket a speculative (and pruning) lift predicted, read before this
  // kernel replaces it
  const unsigned int pred = P.spec_word ? *P.spec_word : 0xFFFFFFFFu;
  const unsigned int margin = P.pruned > 0 ? (unsigned int)(P.pruned - 1) : 0u;
  __syncthreads();
  // input count = sum of the digit-1 histogram; a pruned lift's dropped
  // outputs are those below bucket pred - margin
  int cnt = 0, drop = 0;
  for (int b = threadIdx.x; b < HIST_BINS; b += NTB) {
    const int h = (int)P.hist1[b];
    cnt += h;
    drop += (unsigned int)b + margin < pred ? h : 0;
  }
  int count;
  (void)block_excl_scan<NTB>(cnt, count, s_scan);
  int dropped = 0;
  if (P.pruned && pred != 0xFFFFFFFFu) (void)block_excl_scan<NTB>(drop, dropped, s_scan);
  const bool keepall = (unsigned long long)count <= P.k;
  if (!keepall) find_from_top<NTB>(P.hist1, P.k, &s_b, &s_above, s_scan);
  // a chunked problem sums its input |a|^2 by chunk (k_sel_chunks, k_sel_final)
  const double total = P.chunked ? 0.0 : block_sum_array<NTB>(T.tile_psum + P.tile0, P.npsum, s_red);
  if (threadIdx.x == 0) {
    SelState* sel = P.sel;
    OpStats* st = P.st;
    if (!P.chunked) st->total = total;
    st->n_in = count;
    if (keepall) {
      // every entry passes kept_pred: bits > 0, or bits == 0 and key <= ~0
      sel->keepall = 1;
      sel->pstar = 0;
      sel->keystar = ~0ull;
      st->flags = 0;
    } else {
      if (s_b == 0xFFFFFFFFu) set_err(&st->err, ERR_SELECT);
      sel->b1 = s_b;
      sel->need = P.k - s_above;
      sel->prefix = (unsigned long long)s_b << 52;
      st->flags = 1;
    }
    if (P.spec_word) {
      // a speculative lift ran the digit-1 pass against *spec_word: usable
      // when that was this bucket (stats flags bit 2); the bucket is the next
      // lift's prediction
      if (P.spec_in && !keepall && pred == s_b) {
        sel->spec_hit = 1;
        st->flags |= 4;
      }
      // a pruned lift dropped outputs this select needs (it keeps everything,
      // or its bucket lies more than the margin below the prediction): the
      // relift stores them all, with the digit-1 pass of this bucket (flags
      // bit 3)
      if (P.pruned && dropped > 0 && (keepall || s_b + margin < pred)) {
        sel->relift = 1;
        if (!keepall) sel->spec_hit = 1;
        st->flags |= 8;
      }
      *P.spec_word = keepall ? 0xFFFFFFFFu : s_b;
    }
  }
}

// Per-problem kernels of a batch walk the problems with a persistent grid
// (BatchHdr::nrun); a single problem is its block's.  The barrier after each
// problem keeps the next one's LDS writes behind this one's reads.
#define BASS_PER_PROBLEM(BATCH, call_batch, call_one)               \
  do {                                                              \
    if constexpr (BATCH) {                                          \
      const int np_ = hdr->nrun;                                    \
      for (int c_ = blockIdx.x; c_ < np_; c_ += gridDim.x) {        \
        const SelProblem& P_ = probs[c_];                           \
        call_batch;                                                 \
        __syncthreads();                                            \
      }                                                             \
    } else {                                                        \
      call_one;                                                     \
    }                                                               \
  } while (0)

template <bool BATCH>
__global__ __launch_bounds__(NTB) void k_find1(const SelProblem* probs, SelProblem one, SelTiles T,
                                               const BatchHdr* hdr) {
  BASS_PER_PROBLEM(BATCH, find1_body(P_, T), find1_body(one, T));
}

// Amplitudes of tile gt (slots past its count read its first entry, unused).
__device__ __forceinline__ void ex1_load(const SelProblem& P, int64_t gt, double (&re)[SEL_ITEMS],
                                         double (&im)[SEL_ITEMS], int& cnt) {
  cnt = tile_count(P, gt);
  const int64_t base = tile_base(P, gt);
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int j = it * NT + threadIdx.x;
    const double2 v = P.a[base + (j < cnt ? j : 0)];
    re[it] = v.x;
    im[it] = v.y;
  }
}

// One tile of k_extract1 from its |a|^2 values vp (ex1_tile_core with the
// digit-2 counts).
__device__ __forceinline__ void ex1_tile(const SelProblem& P, const SelTiles& T, int64_t gt, int cnt,
                                         const double (&vp)[SEL_ITEMS], unsigned int b1, unsigned int* s_hist,
                                         int* s_wcnt, double* s_red, double* s_abv) {
  const Ex1Out O{T.c1p, T.c1i, T.cnt1, T.tile_dsum, T.tile_dcnt};
  ex1_tile_core<true, false, EX1_ABV>(O, gt, tile_base(P, gt), cnt, vp, b1, s_hist, s_wcnt, s_red, s_abv);
}

// Per tile (EXT consecutive tiles per block, all of one problem; EXT = 0 takes
// the count from ext_rt): count / sum above the digit-1 bucket, candidates
// inside it, and the block's digit-2 histogram as one row of hist2_rows.
// The tile being processed is held as |a|^2 and only the tile in flight as
// amplitudes (block barriers wait on LDS only).
template <bool BATCH, int EXT>
__device__ __forceinline__ void extract1_body(const SelProblem* probs, int nprob, const SelProblem& one,
                                              const SelTiles& T, int64_t tiles_end, int ext_rt, int64_t vb) {
  __shared__ unsigned int s_hist[HIST_BINS];
  __shared__ int s_wcnt[2 * SEL_ITEMS * WAVES];
  __shared__ double s_red[WAVES];
  __shared__ double s_abv[EX1_ABV];
  const int ext = EXT > 0 ? EXT : ext_rt;
  const int64_t gt0 = vb * ext;
  const SelProblem& P = tile_problem<BATCH>(probs, nprob, one, gt0);
  const SelState* sel = P.sel;
  if (sel->keepall) {
    // every entry is kept: per-tile kept count = tile count, no candidates
    for (int sub = threadIdx.x; sub < ext; sub += NT) {
      const int64_t gt = gt0 + sub;
      if (gt < tiles_end) {
        T.tile_dcnt[gt] = tile_count(P, gt);
        T.tile_dsum[gt] = 0.0;
        T.cnt1[gt] = 0;
      }
    }
    return;
  }
  // a fused select's missed speculation: the lift's tiles are recomputed for
  // this pass (k_lift_ex1)
  if (!BATCH && P.fused && !sel->spec_hit) return;
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  // every bin is cleared before any wave counts into it (a wave past its own
  // bins could otherwise add to a bin another wave has yet to clear)
  __syncthreads();
  if (!BATCH && sel->spec_hit) {
    // the lift already wrote the per-tile counts, sums and candidates: only
    // the row's digit-2 counts, from the candidates' |a|^2 patterns.  The
    // candidates of the block's tiles are walked as one range (tile starts
    // from a scan of their counts), 4 independent loads per thread at a time.
    static_assert(EX1_TILES_MAX < 64, "one wave scans the tile counts");
    __shared__ int s_pre[EX1_TILES_MAX + 2];
    if (threadIdx.x < 64) {
      const int64_t gt = gt0 + threadIdx.x;
      const int c = ((int)threadIdx.x < ext && gt < tiles_end) ? T.cnt1[gt] : 0;
      const int inc = wave_incl_scan(c);
      if (threadIdx.x == 0) s_pre[0] = 0;
      if ((int)threadIdx.x < ext) s_pre[threadIdx.x + 1] = inc;
    }
    __syncthreads();
    const int total = s_pre[ext];
    for (int f0 = 0; f0 < total; f0 += 4 * NT) {
      unsigned long long pb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int f = f0 + u * NT + threadIdx.x;
        pb[u] = 0;
        if (f < total) {
          int lo = 0, hi = ext - 1;  // the last tile t with s_pre[t] <= f
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_pre[mid] <= f)
              lo = mid;
            else
              hi = mid - 1;
          }
          pb[u] = T.c1p[(gt0 + lo) * SEL_TILE + (f - s_pre[lo])];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (f0 + u * NT + (int)threadIdx.x < total)
          atomicAdd(&s_hist[(unsigned int)(pb[u] >> 41) & (HIST_BINS - 1)], 1u);
    }
    __syncthreads();
    unsigned short* row = T.hist2_rows + vb * HIST_BINS;
    for (int b = threadIdx.x; b < HIST_BINS; b += NT) row[b] = (unsigned short)s_hist[b];
    return;
  }
  const unsigned int b1 = sel->b1;
  // the next tile's amplitudes are loaded before the current tile is
  // processed, so one tile's loads are always in flight
  double nr[SEL_ITEMS], ni[SEL_ITEMS];
  int cnt = 0;
  if (gt0 < tiles_end) ex1_load(P, gt0, nr, ni, cnt);
  for (int sub = 0; sub < ext; ++sub) {
    const int64_t gt = gt0 + sub;
    if (gt >= tiles_end) break;  // uniform
    double vp[SEL_ITEMS];
#pragma unroll
    for (int it = 0; it < SEL_ITEMS; ++it) vp[it] = cabs2(make_double2(nr[it], ni[it]));
    const int c = cnt;
    if (sub + 1 < ext && gt + 1 < tiles_end) ex1_load(P, gt + 1, nr, ni, cnt);
    ex1_tile(P, T, gt, c, vp, b1, s_hist, s_wcnt, s_red, s_abv);
  }
  __syncthreads();
  // counts <= EXT * SEL_TILE fit 16 bits; k_hist2_reduce sums the rows (about
  // half of a row's bins are nonzero, too many for per-bin global atomics)
  unsigned short* row = T.hist2_rows + vb * HIST_BINS;
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) row[b] = (unsigned short)s_hist[b];
}

// Per-tile kernels of a batch walk the header's tiles (or waves of WAVES
// tiles) with a persistent grid; a single problem's block is its index.
template <bool BATCH, int EXT>
__global__ __launch_bounds__(NT) void k_extract1(const SelProblem* probs, int nprob, SelProblem one, SelTiles T,
                                                 int64_t tiles_end, int ext_rt, const BatchHdr* hdr) {
  if constexpr (BATCH) {
    const int np = hdr->nrun;
    const int64_t end = hdr->ntiles;  // one tile per row (EXT = 1)
    for (int64_t vb = blockIdx.x; vb < end; vb += gridDim.x) {
      extract1_body<BATCH, EXT>(probs, np, one, T, end, ext_rt, vb);
      __syncthreads();
    }
  } else {
    extract1_body<BATCH, EXT>(probs, nprob, one, T, tiles_end, ext_rt, blockIdx.x);
  }
}

// Column sums of a problem's digit-2 rows: thread = 4 adjacent bins (one
// 8-byte load per row), block = H2_ROWS rows; blockIdx.z = problem.
constexpr int H2_ROWS = 32;
__device__ __forceinline__ void hist2_reduce_body(const SelProblem& P, const SelTiles& T, int bx, int64_t by) {
  if (P.sel->keepall) return;
  const int b = (bx * NT + threadIdx.x) * 4;
  const int64_t r0 = by * H2_ROWS;
  if (r0 >= P.nrows) return;
  const int64_t r1 = r0 + H2_ROWS < P.nrows ? r0 + H2_ROWS : P.nrows;
  unsigned int s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll 8
  for (int64_t r = r0; r < r1; ++r) {
    const ushort4 v = *reinterpret_cast<const ushort4*>(T.hist2_rows + (P.row0 + r) * HIST_BINS + b);
    s0 += v.x;
    s1 += v.y;
    s2 += v.z;
    s3 += v.w;
  }
  if (s0) atomicAdd(&P.hist2[b], s0);
  if (s1) atomicAdd(&P.hist2[b + 1], s1);
  if (s2) atomicAdd(&P.hist2[b + 2], s2);
  if (s3) atomicAdd(&P.hist2[b + 3], s3);
}

// A batch flattens (bin group, row block, problem) into one persistent index.
template <bool BATCH>
__global__ __launch_bounds__(NT) void k_hist2_reduce(const SelProblem* probs, SelProblem one, SelTiles T,
                                                     const BatchHdr* hdr) {
  if constexpr (BATCH) {
    constexpr int BX = HIST_BINS / (4 * NT);
    const int64_t ry = (hdr->max_rows + H2_ROWS - 1) / H2_ROWS;
    const int64_t total = (int64_t)BX * ry * hdr->nrun;
    for (int64_t v = blockIdx.x; v < total; v += gridDim.x) {
      const int bx = (int)(v % BX);
      const int64_t by = (v / BX) % ry;
      const int64_t c = v / (BX * ry);
      hist2_reduce_body(probs[c], T, bx, by);
    }
  } else {
    hist2_reduce_body(one, T, blockIdx.x, blockIdx.y);
  }
}

__device__ __forceinline__ void find2_body(const SelProblem& P) {
  __shared__ int s_scan[WAVES];
  __shared__ unsigned int s_b;
  __shared__ unsigned long long s_above;
  SelState* sel = P.sel;
  if (threadIdx.x == 0) {
    s_b = 0xFFFFFFFFu;
    s_above = 0;
  }
  __syncthreads();
  if (sel->keepall) return;
  const unsigned long long need = sel->need;
  find_from_top(P.hist2, need, &s_b, &s_above, s_scan);
  if (threadIdx.x == 0) {
    if (s_b == 0xFFFFFFFFu) set_err(&P.st->err, ERR_SELECT);
    sel->b2 = s_b;
    sel->need = need - s_above;
    sel->prefix |= (unsigned long long)s_b << 41;
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NT) void k_find2(const SelProblem* probs, SelProblem one, const BatchHdr* hdr) {
  BASS_PER_PROBLEM(BATCH, find2_body(P_), find2_body(one));
}

// Candidates of the digit-2 bucket, appended to the problem's list in any
// order (the threshold does not depend on their order).  One wave per tile (a
// tile holds a few hundred candidates; whole blocks would mostly wait).
template <bool BATCH>
__device__ __forceinline__ void extract2_body(const SelProblem* probs, int nprob, const SelProblem& one,
                                              const SelTiles& T, int64_t tiles_end, int64_t vb) {
  const int64_t gt = vb * WAVES + (threadIdx.x >> 6);
  if (gt >= tiles_end) return;  // uniform per wave
  const SelProblem& P = tile_problem<BATCH>(probs, nprob, one, gt);
  SelState* sel = P.sel;
  if (sel->keepall) return;
  const unsigned int b2 = sel->b2;
  const int lane = threadIdx.x & 63;
  const int cnt = T.cnt1[gt];
  const int64_t base = gt * SEL_TILE;
  constexpr int U = 8;  // candidate loads in flight per lane (clamped to a valid slot)
  for (int j0 = 0; j0 < cnt; j0 += 64 * U) {
    unsigned long long bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * 64 + lane;
      bv[u] = T.c1p[base + (j < cnt ? j : 0)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * 64 + lane;
      const unsigned long long bits = bv[u];
      const bool take = j < cnt && ((unsigned int)(bits >> 41) & (HIST_BINS - 1)) == b2;
      const unsigned long long bal = __ballot(take);
      if (bal) {
        unsigned int wb = 0;
        const int leader = __ffsll((long long)bal) - 1;
        if (lane == leader) wb = atomicAdd(&sel->cand2, (unsigned int)__popcll(bal));
        wb = __shfl(wb, leader, 64);
        if (take) {
          const unsigned int r = (unsigned int)__popcll(bal & ((1ull << lane) - 1ull));
          P.c2p[wb + r] = bits;
          P.c2i[wb + r] = T.c1i[base + j];
        }
      }
    }
  }
}

// Batch form: the header's tiles, WAVES per virtual block (no block barriers:
// a wave leaves a tile past the end on its own).
template <bool BATCH>
__global__ __launch_bounds__(NT) void k_extract2(const SelProblem* probs, int nprob, SelProblem one, SelTiles T,
                                                 int64_t tiles_end, const BatchHdr* hdr) {
  if constexpr (BATCH) {
    const int np = hdr->nrun;
    const int64_t end = hdr->ntiles;
    for (int64_t vb = blockIdx.x; vb * WAVES < end; vb += gridDim.x) extract2_body<BATCH>(probs, np, one, T, end, vb);
  } else {
    extract2_body<BATCH>(probs, nprob, one, T, tiles_end, blockIdx.x);
  }
}

// The same for one problem: each wave takes tpw consecutive tiles and buffers
// its candidates in LDS, reserving space in the list with one atomic per
// buffer flush (usually one per wave) instead of one per 64 candidates with a
// match: the list's single counter, not the candidate reads, bounds the
// per-64 form when most tiles hold a few matches.
constexpr int EX2_CAP = 512;  // buffered candidates per wave
__global__ __launch_bounds__(NT) void k_extract2_agg(SelProblem P, SelTiles T, int64_t tiles_end, int tpw) {
  __shared__ unsigned long long s_bits[WAVES][EX2_CAP];
  __shared__ unsigned int s_idx[WAVES][EX2_CAP];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t t0 = ((int64_t)blockIdx.x * WAVES + w) * tpw;
  if (t0 >= tiles_end) return;  // uniform per wave; no block barriers below
  SelState* sel = P.sel;
  if (sel->keepall) return;
  const unsigned int b2 = sel->b2;
  const int64_t t1 = t0 + tpw < tiles_end ? t0 + tpw : tiles_end;
  const unsigned long long below = (1ull << lane) - 1ull;
  int nbuf = 0;  // wave-uniform
  constexpr int U = 8;  // candidate loads in flight per lane (clamped to a valid slot)
  for (int64_t gt = t0; gt < t1; ++gt) {
    const int cnt = T.cnt1[gt];
    const int64_t base = gt * SEL_TILE;
    for (int j0 = 0; j0 < cnt; j0 += 64 * U) {
      unsigned long long bv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * 64 + lane;
        bv[u] = T.c1p[base + (j < cnt ? j : 0)];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * 64 + lane;
        const unsigned long long bits = bv[u];
        const bool take = j < cnt && ((unsigned int)(bits >> 41) & (HIST_BINS - 1)) == b2;
        const unsigned long long bal = __ballot(take);
        if (!bal) continue;  // uniform
        const int c = __popcll(bal);
        if (nbuf + c > EX2_CAP) {
          // flush: reserve nbuf slots, copy the buffer out (wave-local LDS:
          // the wave's own writes are visible to its later reads)
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          unsigned int wb = 0;
          if (lane == 0) wb = atomicAdd(&sel->cand2, (unsigned int)nbuf);
          wb = __shfl(wb, 0, 64);
          for (int i = lane; i < nbuf; i += 64) {
            P.c2p[wb + i] = s_bits[w][i];
            P.c2i[wb + i] = s_idx[w][i];
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          nbuf = 0;
        }
        if (take) {
          const int r = nbuf + __popcll(bal & below);
          s_bits[w][r] = bits;
          s_idx[w][r] = T.c1i[base + j];
        }
        nbuf += c;
      }
    }
  }
  if (nbuf > 0) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    unsigned int wb = 0;
    if (lane == 0) wb = atomicAdd(&sel->cand2, (unsigned int)nbuf);
    wb = __shfl(wb, 0, 64);
    for (int i = lane; i < nbuf; i += 64) {
      P.c2p[wb + i] = s_bits[w][i];
      P.c2i[wb + i] = s_idx[w][i];
    }
  }
}

// Digits of the multi-block resolutions (staged finish, wide path): |a|^2
// digits at bits 30..40, 19..29, 8..18, 0..7; then, if entries equal to the
// threshold must be split, a radix select of the smallest logical keys over
// the collected tied keys: bits 53..63, 42..52, 31..41, 20..30, 9..19, 0..8.
__device__ __constant__ int kWideShift[WIDE_LEVELS] = {30, 19, 8, 0, 53, 42, 31, 20, 9, 0};
__device__ __constant__ int kWideWidth[WIDE_LEVELS] = {11, 11, 11, 8, 11, 11, 11, 11, 11, 9};

// Resolve the remaining bits of the threshold from the digit-2 candidate
// list, then the key tie-break among entries equal to it.  One block of NTB
// threads per problem, FIN_U candidate loads in flight per thread, histogram
// increments of a wave's lanes that share a bin (exact ties) as one LDS
// atomic.  The logical keys of the tied entries are gathered once: into LDS
// when they fit (FIN_TIES), else into the problem's key list P.tk (through the
// key map's byte tables, built in LDS), whose radix passes take key digits of
// at most 11 bits sized to the key width.
constexpr int FIN_TIES = 2048;
constexpr int FIN_U = 8;
struct FinLds {
  unsigned int hist[HIST_BINS];
  int scan[WAVESB];
  unsigned int b;
  unsigned long long ab;
  uint64_t keys[FIN_TIES];
  unsigned int nk;
  uint64_t ktab[KTAB_WORDS];  // byte tables of the key map (the tied-key list path)
};

// Byte tables of km into LDS tab (the block's threads; km.nbits > 0).
__device__ __forceinline__ void ktab_lds(const KeyMap& km, uint64_t* tab) {
  for (int e = threadIdx.x; e < KTAB_WORDS; e += NTB) {
    const int bb = e >> 8, v = e & 255;
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int p = 8 * bb + i;
      if (((v >> i) & 1) && p < km.nbits) r |= 1ull << km.q[p];
    }
    tab[e] = r;
  }
}

// The single-block resolution of problem P (every thread of the block calls it).
__device__ __forceinline__ void finish_block(const SelProblem& P, FinLds& S) {
  SelState* sel = P.sel;
  OpStats* st = P.st;
  const unsigned int m2 = sel->cand2;
  const unsigned long long* c2p = P.c2p;
  const unsigned int* c2i = P.c2i;
  const int tid = threadIdx.x;
  unsigned long long need = sel->need;
  unsigned long long prefix = sel->prefix;
  const int shifts[4] = {30, 19, 8, 0};
  const int widths[4] = {11, 11, 11, 8};
  unsigned int cnt_eq = 0;
  for (int lv = 0; lv < 4; ++lv) {
    const int sh = shifts[lv], wd = widths[lv];
    for (int b = tid; b < HIST_BINS; b += NTB) S.hist[b] = 0;
    if (tid == 0) {
      S.b = 0xFFFFFFFFu;
      S.ab = 0;
    }
    __syncthreads();
    const unsigned long long hi = prefix >> (sh + wd);
    const unsigned int mask = (1u << wd) - 1u;
    for (unsigned int j0 = 0; j0 < m2; j0 += NTB * FIN_U) {
      unsigned long long bv[FIN_U];
#pragma unroll
      for (int u = 0; u < FIN_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        bv[u] = c2p[j < m2 ? j : 0];
      }
#pragma unroll
      for (int u = 0; u < FIN_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        wave_hist_add(S.hist, (unsigned int)(bv[u] >> sh) & mask, j < m2 && (bv[u] >> (sh + wd)) == hi);
      }
    }
    __syncthreads();
    find_from_top<NTB>(S.hist, need, &S.b, &S.ab, S.scan);
    if (S.b == 0xFFFFFFFFu) {
      if (tid == 0) set_err(&st->err, ERR_SELECT);
      return;
    }
    prefix |= (unsigned long long)S.b << sh;
    need -= S.ab;
    cnt_eq = S.hist[S.b];
    __syncthreads();
  }
  const unsigned long long pstar = prefix;
  unsigned long long keystar = ~0ull;
  if (need < cnt_eq) {
    // need-th smallest logical key among entries whose pattern equals pstar
    const bool in_lds = cnt_eq <= (unsigned int)FIN_TIES;
    const bool tab = !in_lds && P.km.nbits > 0;
    if (tab) ktab_lds(P.km, S.ktab);
    if (tid == 0) S.nk = 0;
    __syncthreads();
    unsigned int base = 0;  // list path: keys written so far (the same in every thread)
    for (unsigned int j0 = 0; j0 < m2; j0 += NTB * FIN_U) {
      unsigned long long bv[FIN_U];
#pragma unroll
      for (int u = 0; u < FIN_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        bv[u] = c2p[j < m2 ? j : 0];
      }
      if (in_lds) {
#pragma unroll
        for (int u = 0; u < FIN_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          if (j < m2 && bv[u] == pstar) {
            const unsigned int slot = atomicAdd(&S.nk, 1u);
            if (slot < (unsigned int)FIN_TIES) S.keys[slot] = lkey(P, P.x[c2i[j]]);
          }
        }
      } else {
        // the round's tied candidates at consecutive list slots (a block scan)
        int c = 0;
#pragma unroll
        for (int u = 0; u < FIN_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          c += (j < m2 && bv[u] == pstar) ? 1 : 0;
        }
        int tot;
        const int pre = block_excl_scan<NTB>(c, tot, S.scan);
        unsigned int off = base + (unsigned int)pre;
#pragma unroll
        for (int u = 0; u < FIN_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          if (j < m2 && bv[u] == pstar) {
            const uint64_t w = P.x[c2i[j]];
            P.tk[off++] = tab ? logical_key_tab(w, S.ktab) : w;
          }
        }
        base += (unsigned int)tot;
      }
    }
    __syncthreads();
    const unsigned int nk = in_lds ? (S.nk < (unsigned int)FIN_TIES ? S.nk : (unsigned int)FIN_TIES) : base;
    unsigned long long kp = 0;
    if (in_lds) {
      for (int sh = 56; sh >= 0; sh -= 8) {
        for (int b = tid; b < HIST_BINS; b += NTB) S.hist[b] = 0;
        if (tid == 0) {
          S.b = 0xFFFFFFFFu;
          S.ab = 0;
        }
        __syncthreads();
        for (unsigned int e = tid; e < nk; e += NTB) {
          const uint64_t key = S.keys[e];
          const bool match = (sh == 56) ? true : ((key >> (sh + 8)) == (kp >> (sh + 8)));
          if (match) atomicAdd(&S.hist[(unsigned int)(key >> sh) & 255u], 1u);
        }
        __syncthreads();
        find_from_bottom<NTB>(S.hist, need, &S.b, &S.ab, S.scan);
        if (S.b == 0xFFFFFFFFu) {
          if (tid == 0) set_err(&st->err, ERR_SELECT);
          return;
        }
        kp |= (unsigned long long)S.b << sh;
        need -= S.ab;
        __syncthreads();
      }
    } else {
      // the list's keys have kb significant bits: the fewest levels of at
      // most 11 bits, from the top
      const int kb = P.km.nbits ? P.km.nbits : 64;
      const int nlev = (kb + 10) / 11;
      const int w = (kb + nlev - 1) / nlev;
      for (int lv = 0; lv < nlev; ++lv) {
        const int hi = kb - lv * w;
        const int lo = hi - w > 0 ? hi - w : 0;
        const unsigned int mask = (1u << (hi - lo)) - 1u;
        for (int b = tid; b < HIST_BINS; b += NTB) S.hist[b] = 0;
        if (tid == 0) {
          S.b = 0xFFFFFFFFu;
          S.ab = 0;
        }
        __syncthreads();
        for (unsigned int j0 = 0; j0 < nk; j0 += NTB * FIN_U) {
          uint64_t kv[FIN_U];
#pragma unroll
          for (int u = 0; u < FIN_U; ++u) {
            const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
            kv[u] = P.tk[j < nk ? j : 0];
          }
#pragma unroll
          for (int u = 0; u < FIN_U; ++u) {
            const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
            wave_hist_add(S.hist, (unsigned int)(kv[u] >> lo) & mask,
                          j < nk && (hi >= 64 || (kv[u] >> hi) == (kp >> hi)));
          }
        }
        __syncthreads();
        find_from_bottom<NTB>(S.hist, need, &S.b, &S.ab, S.scan);
        if (S.b == 0xFFFFFFFFu) {
          if (tid == 0) set_err(&st->err, ERR_SELECT);
          return;
        }
        kp |= (unsigned long long)S.b << lo;
        need -= S.ab;
        __syncthreads();
      }
    }
    keystar = kp;
  }
  if (tid == 0) {
    sel->pstar = pstar;
    sel->keystar = keystar;
  }
}

__device__ __forceinline__ void finish_body(const SelProblem& P) {
  __shared__ FinLds S;
  SelState* sel = P.sel;
  if (sel->keepall) return;  // pstar / keystar set by k_find1
  if (sel->cand2 > FIN_MAX && !P.finish_any) {
    // too many for one block: keep nothing, ask for the wide path
    if (threadIdx.x == 0) {
      sel->pstar = ~0ull;
      sel->keystar = 0;
      atomicOr(&P.st->flags, 2);
    }
    return;
  }
  finish_block(P, S);
}

template <bool BATCH>
__global__ __launch_bounds__(NTB) void k_finish(const SelProblem* probs, SelProblem one, const BatchHdr* hdr) {
  BASS_PER_PROBLEM(BATCH, finish_body(P_), finish_body(one));
}

// ------------------------------------------------------ staged finish ----
// One problem, any digit-2 candidate count, one launch.  Up to FIN_MAX
// candidates block 0 runs finish_block and the other blocks exit.  Beyond
// that (massive |a|^2 ties at the threshold) the blocks run STG_STAGES stages
// of STG_BLOCKS blocks each:
//   stages 0-3: histograms of the |a|^2 digits at bits 30..40, 19..29, 8..18,
//               0..7 over the candidates matching the resolved prefix;
//   stage 4:    the logical keys of the candidates equal to the threshold,
//               appended to P.tk (only when the key tie-break is needed);
//   stages 5-10: histograms of the key digits, from the top of the key width
//               (KeyMap bits, else 64) down, in the fewest levels of at most 11
//               bits, widths balanced (56-bit keys: 6 levels of 10, 10, 10,
//               10, 10, 6 bits); stages past the last level end at once.
// Stage s uses its own zeroed histogram (hist_w + HIST_BINS * level).  The
// last block of a stage to finish (a per-stage counter) resolves its digit,
// publishes the result in the select state and sets the stage's flag; blocks
// of stage s + 1 wait for that flag before they read anything.  A block's
// stage comes from an atomic ticket in dispatch order, so every block of stage
// s had started before any block of stage s + 1 waits on it, and stage-s
// blocks wait only on stage s - 1: every wait ends.  A spin bound turns a
// broken invariant into an error bit instead of a hang.  Data crossing blocks
// inside the launch (histogram atomics, the collected keys, the published
// digits) is ordered by agent-scope release fences before each counter
// increment / flag store and acquire fences after each counter / flag read.
// Tied candidates land in one histogram bin: a wave whose active lanes share
// a bin adds them with one LDS atomic.
constexpr int STG_BLOCKS = 256;
constexpr int STG_STAGES = 11;
constexpr int STG_COLLECT = 4;
constexpr int STG_KEY_LEVELS = 6;
constexpr unsigned int STG_SPIN = 1u << 24;
constexpr int STG_U = 8;  // candidate loads in flight per thread
static_assert(STG_STAGES <= 12 && STG_STAGES == STG_COLLECT + 1 + STG_KEY_LEVELS, "stage flags of SelState");

__device__ __forceinline__ unsigned int ld_acq_u32(unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish the end of stages [from, STG_STAGES) at once (error exits and a
// threshold without a key tie-break release every later waiter).
__device__ __forceinline__ void stg_release_from(SelState* sel, int from) {
  __threadfence();
  for (int s = from; s < STG_STAGES; ++s)
    __hip_atomic_store(&sel->stg_flag[s], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Key digit `lv` (0 = most significant) of keys of kb bits: bits [lo, hi).
__device__ __forceinline__ void stg_key_digit(int kb, int lv, int& lo, int& hi) {
  const int nlev = (kb + 10) / 11;
  const int w = (kb + nlev - 1) / nlev;
  hi = kb - lv * w;
  lo = hi - w > 0 ? hi - w : 0;
  if (hi < 0) hi = 0;
}

__global__ __launch_bounds__(NTB) void k_finish_staged(SelProblem P) {
  __shared__ FinLds S;
  __shared__ unsigned int s_t, s_flag, s_base;
  SelState* sel = P.sel;
  OpStats* st = P.st;
  if (sel->keepall) return;  // pstar / keystar set by k_find1
  const unsigned int m2 = sel->cand2;
  if (m2 <= FIN_MAX) {
    if (blockIdx.x == 0) finish_block(P, S);
    return;
  }
  const int tid = threadIdx.x;
  if (tid == 0) s_t = atomicAdd(&sel->stg_ticket, 1u);
  __syncthreads();
  const unsigned int ticket = s_t;
  const int stage = (int)(ticket / STG_BLOCKS), part = (int)(ticket % STG_BLOCKS);
  if (stage >= STG_STAGES) return;  // the grid holds exactly STG_STAGES * STG_BLOCKS blocks
  if (stage > 0) {
    if (tid == 0) {
      unsigned int spins = 0, ok = 1;
      while (ld_acq_u32(&sel->stg_flag[stage - 1]) == 0u) {
        if (++spins > STG_SPIN) {
          set_err(&st->err, ERR_LOOKBACK);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
      s_flag = ok && ld_acq_u32(&sel->stg_abort) == 0u;
    }
    __syncthreads();
    if (!s_flag) return;
    __threadfence();  // acquire what the earlier stages published
  }
  const unsigned long long prefix = __hip_atomic_load(&sel->prefix, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long need = __hip_atomic_load(&sel->need, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long kp = __hip_atomic_load(&sel->kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long pstar = __hip_atomic_load(&sel->pstar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned int tie = ld_acq_u32(&sel->tie);
  const unsigned int tcount = ld_acq_u32(&sel->tcount);
  // released by stage 3 without a tie-break, or by the last key level
  if (stage >= STG_COLLECT && (!tie || ld_acq_u32(&sel->stg_flag[stage]))) return;
  const int kb = P.km.nbits ? P.km.nbits : 64;
  const int lvh = stage < STG_COLLECT ? stage : stage - 1;  // histogram
  int sh, wd;
  if (stage < STG_COLLECT) {
    sh = kWideShift[lvh];
    wd = kWideWidth[lvh];
  } else {
    int hi;
    stg_key_digit(kb, stage - STG_COLLECT - 1, sh, hi);
    wd = hi - sh;
  }
  unsigned int* hist = P.hist_w + (size_t)lvh * HIST_BINS;
  const unsigned int stride = (unsigned int)STG_BLOCKS * NTB;
  if (stage != STG_COLLECT) {
    for (int b = tid; b < HIST_BINS; b += NTB) S.hist[b] = 0;
    __syncthreads();
    const unsigned int mask = (1u << wd) - 1u;
    const int top = sh + wd;  // may be 64 at the first key level
    if (stage < STG_COLLECT) {
      const unsigned long long hi = prefix >> top;
      for (unsigned int j0 = (unsigned int)part * NTB * STG_U; j0 < m2; j0 += stride * STG_U) {
        unsigned long long bv[STG_U];
#pragma unroll
        for (int u = 0; u < STG_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          bv[u] = P.c2p[j < m2 ? j : 0];
        }
#pragma unroll
        for (int u = 0; u < STG_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          wave_hist_add(S.hist, (unsigned int)(bv[u] >> sh) & mask, j < m2 && (bv[u] >> top) == hi);
        }
      }
    } else {
      for (unsigned int j0 = (unsigned int)part * NTB * STG_U; j0 < tcount; j0 += stride * STG_U) {
        uint64_t kv[STG_U];
#pragma unroll
        for (int u = 0; u < STG_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          kv[u] = P.tk[j < tcount ? j : 0];
        }
#pragma unroll
        for (int u = 0; u < STG_U; ++u) {
          const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
          wave_hist_add(S.hist, (unsigned int)(kv[u] >> sh) & mask,
                       j < tcount && (top >= 64 || (kv[u] >> top) == (kp >> top)));
        }
      }
    }
    __syncthreads();
    for (int b = tid; b < HIST_BINS; b += NTB)
      if (S.hist[b]) atomicAdd(&hist[b], S.hist[b]);
  } else {
    // tied candidates -> logical keys (in any order); one list reservation
    // per round of NTB * STG_U candidates (a block scan of the per-thread
    // counts); the key map's byte tables staged in LDS
    const bool tab = P.km.nbits > 0 && P.ktab;
    if (tab)
      for (int e = tid; e < KTAB_WORDS; e += NTB) S.ktab[e] = P.ktab[e];
    __syncthreads();
    for (unsigned int j0 = (unsigned int)part * NTB * STG_U; j0 < m2; j0 += stride * STG_U) {
      unsigned long long bv[STG_U];
#pragma unroll
      for (int u = 0; u < STG_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        bv[u] = P.c2p[j < m2 ? j : 0];
      }
      int c = 0;
#pragma unroll
      for (int u = 0; u < STG_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        c += (j < m2 && bv[u] == pstar) ? 1 : 0;
      }
      int tot;
      int off = block_excl_scan<NTB>(c, tot, S.scan);
      if (tid == 0) s_base = tot > 0 ? atomicAdd(&sel->tcount, (unsigned int)tot) : 0u;
      __syncthreads();
      const unsigned int base = s_base;
#pragma unroll
      for (int u = 0; u < STG_U; ++u) {
        const unsigned int j = j0 + (unsigned int)(u * NTB + tid);
        if (j < m2 && bv[u] == pstar) {
          const uint64_t w = P.x[P.c2i[j]];
          P.tk[base + (unsigned int)off] = tab ? logical_key_tab(w, S.ktab) : logical_key(w, P.km);
          ++off;
        }
      }
      __syncthreads();
    }
  }
  // this block's share is done: release it, count it; the last block goes on
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    s_flag = atomicAdd(&sel->stg_done[stage], 1u) == (unsigned int)(STG_BLOCKS - 1);
  }
  __syncthreads();
  if (!s_flag) return;
  __threadfence();  // acquire every block's share of the stage
  if (stage == STG_COLLECT) {
    if (tid == 0) {
      __threadfence();
      __hip_atomic_store(&sel->stg_flag[stage], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return;
  }
  for (int b = tid; b < HIST_BINS; b += NTB) S.hist[b] = ld_acq_u32(&hist[b]);
  if (tid == 0) {
    S.b = 0xFFFFFFFFu;
    S.ab = 0;
  }
  __syncthreads();
  if (stage < STG_COLLECT)
    find_from_top<NTB>(S.hist, need, &S.b, &S.ab, S.scan);
  else
    find_from_bottom<NTB>(S.hist, need, &S.b, &S.ab, S.scan);
  if (tid != 0) return;
  const unsigned int b = S.b;
  if (b == 0xFFFFFFFFu) {
    set_err(&st->err, ERR_SELECT);
    __hip_atomic_store(&sel->stg_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    stg_release_from(sel, stage);
    return;
  }
  const unsigned long long nneed = need - S.ab;
  sel->need = nneed;
  if (stage < STG_COLLECT) {
    const unsigned long long np = prefix | ((unsigned long long)b << sh);
    sel->prefix = np;
    if (stage == STG_COLLECT - 1) {
      const unsigned int tied = nneed < (unsigned long long)S.hist[b] ? 1u : 0u;
      sel->pstar = np;
      sel->keystar = ~0ull;
      sel->kp = 0;
      sel->tie = tied;
      if (!tied) {
        stg_release_from(sel, stage);
        return;
      }
    }
  } else {
    const unsigned long long nkp = kp | ((unsigned long long)b << sh);
    sel->kp = nkp;
    if (sh == 0) {
      // the last key level: the need-th smallest tied key; later stages end
      sel->keystar = nkp;
      stg_release_from(sel, stage);
      return;
    }
  }
  __threadfence();
  __hip_atomic_store(&sel->stg_flag[stage], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------- wide path --
// Multi-launch resolution of the digit-2 candidates (P.c2p / P.c2i, count
// sel->cand2) for a batch of problems whose threshold buckets hold very many
// equal |a|^2 (the levels of kWideShift / kWideWidth).  One histogram per
// problem (P.hist_w): each find clears it for the next level.
constexpr int WIDE_BLOCKS = 256;

__device__ __forceinline__ void flush_hist(unsigned int* s_hist, unsigned int* hist) {
  __syncthreads();
  for (int b = threadIdx.x; b < HIST_BINS; b += NT)
    if (s_hist[b]) atomicAdd(&hist[b], s_hist[b]);
}

template <bool BATCH>
__global__ __launch_bounds__(NT) void k_wide_hist(const SelProblem* probs, SelProblem one, int lv) {
  __shared__ unsigned int s_hist[HIST_BINS];
  const SelProblem& P = block_problem<BATCH>(probs, one, blockIdx.y);
  const SelState* sel = P.sel;
  if (sel->keepall) return;
  const int sh = kWideShift[lv], wd = kWideWidth[lv];
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  __syncthreads();
  const unsigned int m2 = sel->cand2;
  const unsigned long long hi = sel->prefix >> (sh + wd);
  for (unsigned int j = blockIdx.x * NT + threadIdx.x; j < m2; j += gridDim.x * NT) {
    const unsigned long long bits = P.c2p[j];
    if ((bits >> (sh + wd)) == hi) atomicAdd(&s_hist[(unsigned int)(bits >> sh) & ((1u << wd) - 1u)], 1u);
  }
  flush_hist(s_hist, P.hist_w);
}

template <bool BATCH>
__global__ __launch_bounds__(NTB) void k_wide_find(const SelProblem* probs, SelProblem one, int lv) {
  __shared__ int s_scan[WAVESB];
  __shared__ unsigned int s_b;
  __shared__ unsigned long long s_ab;
  const SelProblem& P = block_problem<BATCH>(probs, one, blockIdx.x);
  SelState* sel = P.sel;
  if (sel->keepall) return;
  if (lv >= 4 && !sel->tie) return;
  const int sh = kWideShift[lv];
  unsigned int* hist = P.hist_w;
  if (threadIdx.x == 0) {
    s_b = 0xFFFFFFFFu;
    s_ab = 0;
  }
  __syncthreads();
  const unsigned long long need = sel->need;
  if (lv < 4)
    find_from_top<NTB>(hist, need, &s_b, &s_ab, s_scan);
  else
    find_from_bottom<NTB>(hist, need, &s_b, &s_ab, s_scan);
  const unsigned int b = s_b;
  const unsigned int cnt_b = b != 0xFFFFFFFFu ? hist[b] : 0u;
  __syncthreads();
  for (int i = threadIdx.x; i < HIST_BINS; i += NTB) hist[i] = 0;
  if (threadIdx.x != 0) return;
  if (b == 0xFFFFFFFFu) {
    set_err(&P.st->err, ERR_SELECT);
    return;
  }
  sel->need = need - s_ab;
  if (lv < 4) {
    sel->prefix |= (unsigned long long)b << sh;
    sel->cnt_eq = cnt_b;
    if (lv == 3) {
      sel->pstar = sel->prefix;
      sel->tie = sel->need < sel->cnt_eq ? 1u : 0u;
      sel->keystar = ~0ull;
      sel->kp = 0;
    }
  } else {
    sel->kp |= (unsigned long long)b << sh;
    if (lv == WIDE_LEVELS - 1) sel->keystar = sel->kp;
  }
}

// Logical keys of the candidates equal to the threshold (any order).
template <bool BATCH>
__global__ __launch_bounds__(NT) void k_wide_collect(const SelProblem* probs, SelProblem one) {
  const SelProblem& P = block_problem<BATCH>(probs, one, blockIdx.y);
  SelState* sel = P.sel;
  if (sel->keepall || !sel->tie) return;
  const unsigned int m2 = sel->cand2;
  const unsigned long long pstar = sel->pstar;
  const int lane = threadIdx.x & 63;
  for (unsigned int j0 = blockIdx.x * NT; j0 < m2; j0 += gridDim.x * NT) {
    const unsigned int j = j0 + threadIdx.x;
    const bool take = j < m2 && P.c2p[j] == pstar;
    const unsigned long long bal = __ballot(take);
    if (!bal) continue;
    unsigned int wb = 0;
    const int leader = __ffsll((long long)bal) - 1;
    if (lane == leader) wb = atomicAdd(&sel->tcount, (unsigned int)__popcll(bal));
    wb = __shfl(wb, leader, 64);
    if (take) P.tk[wb + __popcll(bal & ((1ull << lane) - 1ull))] = lkey(P, P.x[P.c2i[j]]);
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NT) void k_wide_key_hist(const SelProblem* probs, SelProblem one, int lv) {
  __shared__ unsigned int s_hist[HIST_BINS];
  const SelProblem& P = block_problem<BATCH>(probs, one, blockIdx.y);
  const SelState* sel = P.sel;
  if (sel->keepall || !sel->tie) return;
  const int sh = kWideShift[lv], wd = kWideWidth[lv];
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  __syncthreads();
  const unsigned int m = sel->tcount;
  const unsigned long long kp = sel->kp;
  const int top = sh + wd;  // 64 at the first key level
  for (unsigned int j = blockIdx.x * NT + threadIdx.x; j < m; j += gridDim.x * NT) {
    const uint64_t key = P.tk[j];
    if (top >= 64 || (key >> top) == (kp >> top))
      atomicAdd(&s_hist[(unsigned int)(key >> sh) & ((1u << wd) - 1u)], 1u);
  }
  flush_hist(s_hist, P.hist_w);
}


// Per tile: sum and count of the kept candidates (lane-strided, then a
// fixed shuffle tree).  One wave per tile.
template <bool BATCH>
__device__ __forceinline__ void keptsum_body(const SelProblem* probs, int nprob, const SelProblem& one,
                                             const SelTiles& T, int64_t tiles_end, int64_t vb) {
  const int64_t gt = vb * WAVES + (threadIdx.x >> 6);
  if (gt >= tiles_end) return;  // uniform per wave
  const SelProblem& P = tile_problem<BATCH>(probs, nprob, one, gt);
  const unsigned long long pstar = P.sel->pstar, keystar = P.sel->keystar;
  const int lane = threadIdx.x & 63;
  const int cnt = T.cnt1[gt];
  const int64_t base = gt * SEL_TILE;
  double s = 0.0;
  unsigned long long c = 0;
  constexpr int U = 8;  // candidate loads in flight per lane (clamped to a valid slot)
  for (int j0 = 0; j0 < cnt; j0 += 64 * U) {
    unsigned long long bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * 64 + lane;
      bv[u] = T.c1p[base + (j < cnt ? j : 0)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = j0 + u * 64 + lane;
      if (j >= cnt) continue;
      const unsigned long long bits = bv[u];
      bool keep = bits > pstar;
      if (bits == pstar) keep = lkey(P, P.x[T.c1i[base + j]]) <= keystar;
      if (keep) {
        s += __longlong_as_double((long long)bits);
        ++c;
      }
    }
  }
  const double tot = wave_sum(s);
  const unsigned long long cc = wave_sum_u64(c);
  if (lane == 0) {
    T.ksum[gt] = tot;
    T.kcnt[gt] = (int)cc;
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NT) void k_keptsum(const SelProblem* probs, int nprob, SelProblem one, SelTiles T,
                                                int64_t tiles_end, const BatchHdr* hdr) {
  if constexpr (BATCH) {
    const int np = hdr->nrun;
    const int64_t end = hdr->ntiles;
    for (int64_t vb = blockIdx.x; vb * WAVES < end; vb += gridDim.x) keptsum_body<BATCH>(probs, np, one, T, end, vb);
  } else {
    keptsum_body<BATCH>(probs, nprob, one, T, tiles_end, blockIdx.x);
  }
}

// kept = sum over the problem's tiles in a fixed order; kept-entry offsets
// per tile (exclusive scan).  Thread i owns FIN_CH consecutive tiles per
// round; its loads are issued together.
constexpr int FIN_CH = 4;
__device__ __forceinline__ void sel_finalize_body(const SelProblem& P, const SelTiles& T) {
  __shared__ double s_red[WAVESB];
  __shared__ int s_scan[WAVESB];
  const int64_t t_end = P.tile0 + P.ntiles;
  double s = 0.0;
  int64_t carry = 0;
  for (int64_t r0 = P.tile0; r0 < t_end; r0 += (int64_t)NTB * FIN_CH) {
    const int64_t t0 = r0 + (int64_t)threadIdx.x * FIN_CH;
    double ds[FIN_CH], ks[FIN_CH];
    int c[FIN_CH];
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) {
      const int64_t t = t0 + q;
      const bool in = t < t_end;
      ds[q] = in ? T.tile_dsum[t] : 0.0;
      ks[q] = in ? T.ksum[t] : 0.0;
      c[q] = in ? T.tile_dcnt[t] + T.kcnt[t] : 0;
    }
    int v = 0;
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) {
      s += ds[q] + ks[q];
      v += c[q];
    }
    int tot;
    int64_t run = carry + block_excl_scan<NTB>(v, tot, s_scan);
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) {
      if (t0 + q < t_end) T.koff[t0 + q] = run;
      run += c[q];
    }
    carry += tot;
  }
  const double kept = block_sum<NTB>(s, s_red);
  if (threadIdx.x == 0) {
    OpStats* st = P.st;
    st->nnz = carry;
    if (P.sel->keepall) {
      st->kept = st->total;
      st->norm = 1.0;
    } else if (st->flags & 2) {
      st->kept = 0.0;  // result discarded; the caller repeats with the wide path
      st->norm = 1.0;
    } else {
      if ((unsigned long long)carry != P.k) set_err(&st->err, ERR_SELECT);
      st->kept = kept;
      const double norm = sqrt(kept);
      st->norm = (P.renorm && norm > 1e-30) ? norm : 1.0;
    }
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NTB) void k_sel_finalize(const SelProblem* probs, SelProblem one, SelTiles T,
                                                      const BatchHdr* hdr) {
  BASS_PER_PROBLEM(BATCH, sel_finalize_body(P_, T), sel_finalize_body(one, T));
}

// Chunked form of k_sel_finalize for one large problem, first level: block c
// scans the kept counts of tiles [c * SEL_CHUNK, (c+1) * SEL_CHUNK) into
// chunk-local offsets and sums, in a fixed order, their kept and input |a|^2.
__global__ __launch_bounds__(NTB) void k_sel_chunks(SelProblem P, SelTiles T) {
  static_assert(SEL_CHUNK == NTB * FIN_CH, "one round per chunk");
  __shared__ double s_red[WAVESB];
  __shared__ int s_scan[WAVESB];
  const int64_t c = blockIdx.x;
  const int64_t t_end = P.tile0 + P.ntiles, p_end = P.tile0 + P.npsum;
  const int64_t t0 = P.tile0 + c * SEL_CHUNK + (int64_t)threadIdx.x * FIN_CH;
  double ds[FIN_CH], ks[FIN_CH], ps[FIN_CH];
  int cn[FIN_CH];
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    const int64_t t = t0 + q;
    const bool in = t < t_end;
    ds[q] = in ? T.tile_dsum[t] : 0.0;
    ks[q] = in ? T.ksum[t] : 0.0;
    cn[q] = in ? T.tile_dcnt[t] + T.kcnt[t] : 0;
    ps[q] = t < p_end ? T.tile_psum[t] : 0.0;
  }
  double sk = 0.0, sp = 0.0;
  int v = 0;
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    sk += ds[q] + ks[q];
    sp += ps[q];
    v += cn[q];
  }
  int tot;
  int64_t run = block_excl_scan<NTB>(v, tot, s_scan);
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    if (t0 + q < t_end) T.koff[t0 + q] = run;
    run += cn[q];
  }
  const double kept = block_sum<NTB>(sk, s_red);
  const double inp = block_sum<NTB>(sp, s_red);
  if (threadIdx.x == 0) {
    T.ch_cnt[c] = tot;
    T.ch_sum[c] = kept;
    T.ch_psum[c] = inp;
  }
}

// Second level: chunk offsets (exclusive scan of the chunk counts), kept and
// input totals summed over the chunks in a fixed order; the stats of
// k_sel_finalize.
__global__ __launch_bounds__(NTB) void k_sel_final(SelProblem P, SelTiles T, int64_t nch) {
  __shared__ double s_red[WAVESB];
  __shared__ int s_scan[WAVESB];
  int64_t carry = 0;
  for (int64_t r0 = 0; r0 < nch; r0 += NTB) {
    const int64_t c = r0 + threadIdx.x;
    const int v = c < nch ? (int)T.ch_cnt[c] : 0;
    int tot;
    const int ex = block_excl_scan<NTB>(v, tot, s_scan);
    if (c < nch) T.ch_off[c] = carry + ex;
    carry += tot;
  }
  const double kept = block_sum_array<NTB>(T.ch_sum, nch, s_red);
  const double total = block_sum_array<NTB>(T.ch_psum, nch, s_red);
  if (threadIdx.x == 0) {
    OpStats* st = P.st;
    st->total = total;
    st->nnz = carry;
    if (P.sel->keepall) {
      st->kept = total;
      st->norm = 1.0;
    } else if (st->flags & 2) {
      st->kept = 0.0;  // result discarded; the caller repeats with the wide path
      st->norm = 1.0;
    } else {
      if ((unsigned long long)carry != P.k) set_err(&st->err, ERR_SELECT);
      st->kept = kept;
      const double norm = sqrt(kept);
      st->norm = (P.renorm && norm > 1e-30) ? norm : 1.0;
    }
  }
}

template <bool BATCH>
__device__ __forceinline__ void compact_body(const SelProblem* probs, int nprob, const SelProblem& one,
                                             const SelTiles& T, int64_t gt) {
  __shared__ int s_scan[SEL_ITEMS * WAVES];
  __shared__ double s_red[WAVES];
  const SelProblem& P = tile_problem<BATCH>(probs, nprob, one, gt);
  const unsigned long long pstar = P.sel->pstar, keystar = P.sel->keystar;
  // numpy divides complex128 by a real norm as a product with 1/norm
  const double inv = 1.0 / P.st->norm;
  const int cnt = tile_count(P, gt);
  const int64_t base = tile_base(P, gt);
  uint64_t kk[SEL_ITEMS];
  double2 vv[SEL_ITEMS];
  bool keep[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int j = it * NT + threadIdx.x;
    keep[it] = false;
    if (j < cnt) {
      kk[it] = P.x[base + j];
      vv[it] = P.a[base + j];
      keep[it] = kept_pred(pbits_of(vv[it]), kk[it], pstar, keystar, P);
    }
  }
  const int64_t rb = T.koff[gt] + (P.chunked ? T.ch_off[(gt - P.tile0) / SEL_CHUNK] : 0);
  // kept-entry offsets in slot (item-major) order: per (item, wave) ballot
  // counts, one wave scans them, a lane adds its rank among the set lanes
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    m[it] = __ballot(keep[it]);
    if (lane == 0) s_scan[it * WAVES + w] = __popcll(m[it]);
  }
  __syncthreads();
  const int cw = lane < SEL_ITEMS * WAVES ? s_scan[lane] : 0;
  const int cx = wave_incl_scan(cw) - cw;
  const int ktot = __builtin_amdgcn_readlane(cx + cw, SEL_ITEMS * WAVES - 1);
  const int wu = __builtin_amdgcn_readfirstlane(w);
  // |a|^2 of the written entries by kept rank: the tile's sums in the order of
  // its kept entries, whichever slots held them
  double pv[SEL_ITEMS];
  int kr[SEL_ITEMS];
  bool fits = true;
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const unsigned long long mk = m[it];
    kr[it] = __builtin_amdgcn_readlane(cx, it * WAVES + wu) +
             (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mk, 0u));
    pv[it] = 0.0;
    if (keep[it]) {
      const double2 v = make_double2(vv[it].x * inv, vv[it].y * inv);
      const int64_t pos = rb + kr[it];
      if (pos < P.cap_out) {
        P.xo[pos] = kk[it];
        P.ao[pos] = v;
      } else {
        fits = false;
      }
      pv[it] = cabs2(v);
    }
  }
  if (!fits) set_err(&P.st->err, ERR_OVERFLOW);
  __shared__ double s_kbuf[CMP_BUF];
  double s4;
  const double s2 = rank_ordered_sum<SEL_ITEMS, CMP_BUF, true>(pv, kr, keep, ktot, s_kbuf, &s4);
  const double t2 = block_sum(s2, s_red);
  const double t4 = block_sum(s4, s_red);
  if (threadIdx.x == 0) {
    T.tile_s2[gt] = t2;
    T.tile_s4[gt] = t4;
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NT) void k_compact(const SelProblem* probs, int nprob, SelProblem one, SelTiles T,
                                                const BatchHdr* hdr) {
  if constexpr (BATCH) {
    const int np = hdr->nrun;
    const int64_t end = hdr->ntiles;
    for (int64_t gt = blockIdx.x; gt < end; gt += gridDim.x) {
      compact_body<BATCH>(probs, np, one, T, gt);
      __syncthreads();
    }
  } else {
    compact_body<BATCH>(probs, nprob, one, T, blockIdx.x);
  }
}

// Fixed-order reduction of a problem's per-tile sums into its stats; a batch
// problem with a count cell gets its output count there.
__device__ __forceinline__ void sums_finalize_body(const SelProblem& P, const SelTiles& T) {
  __shared__ double s_red[WAVESB];
  const double s2 = block_sum_array<NTB>(T.tile_s2 + P.tile0, P.ntiles, s_red);
  const double s4 = block_sum_array<NTB>(T.tile_s4 + P.tile0, P.ntiles, s_red);
  if (threadIdx.x == 0) {
    P.st->sum2 = s2;
    P.st->sum4 = s4;
    if (P.cnt_out) *P.cnt_out = P.st->nnz;
  }
}

template <bool BATCH>
__global__ __launch_bounds__(NTB) void k_sums_finalize(const SelProblem* probs, SelProblem one, SelTiles T,
                                                       const BatchHdr* hdr) {
  BASS_PER_PROBLEM(BATCH, sums_finalize_body(P_, T), sums_finalize_body(one, T));
}

// Chunked form of k_sums_finalize: per-chunk sums of the output's |a|^2 and
// |a|^4 in a fixed order, then their fixed-order totals.
__global__ __launch_bounds__(NTB) void k_sums_chunks(SelProblem P, SelTiles T) {
  __shared__ double s_red[WAVESB];
  const int64_t t_end = P.tile0 + P.ntiles;
  const int64_t t0 = P.tile0 + (int64_t)blockIdx.x * SEL_CHUNK + (int64_t)threadIdx.x * FIN_CH;
  double a2[FIN_CH], a4[FIN_CH];
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    const bool in = t0 + q < t_end;
    a2[q] = in ? T.tile_s2[t0 + q] : 0.0;
    a4[q] = in ? T.tile_s4[t0 + q] : 0.0;
  }
  double s2 = 0.0, s4 = 0.0;
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    s2 += a2[q];
    s4 += a4[q];
  }
  const double t2 = block_sum<NTB>(s2, s_red);
  const double t4 = block_sum<NTB>(s4, s_red);
  if (threadIdx.x == 0) {
    T.ch_s2[blockIdx.x] = t2;
    T.ch_s4[blockIdx.x] = t4;
  }
}

__global__ __launch_bounds__(NTB) void k_sums_final(SelProblem P, SelTiles T, int64_t nch) {
  __shared__ double s_red[WAVESB];
  const double s2 = block_sum_array<NTB>(T.ch_s2, nch, s_red);
  const double s4 = block_sum_array<NTB>(T.ch_s4, nch, s_red);
  if (threadIdx.x == 0) {
    P.st->sum2 = s2;
    P.st->sum4 = s4;
  }
}

__global__ __launch_bounds__(NTB) void k_sums_reduce(const double* tile_s2, const double* tile_s4,
                                                      int64_t ntiles, OpStats* st) {
  __shared__ double s_red[WAVESB];
  const double s2 = block_sum_array<NTB>(tile_s2, ntiles, s_red);
  const double s4 = block_sum_array<NTB>(tile_s4, ntiles, s_red);
  if (threadIdx.x == 0) {
    st->sum2 = s2;
    st->sum4 = s4;
  }
}

// Segment counts -> exclusive offsets and total (nnz of a segmented output).
// Single block; ch_off[0] = 0 so readers add ch_off[t / SEL_CHUNK] in both forms.
__global__ __launch_bounds__(NTB) void k_seg_scan(const int* seg_cnt, int64_t ntiles, int64_t* soff,
                                                   int64_t* ch_off, OpStats* st) {
  __shared__ int s_scan[WAVESB];
  int64_t carry = 0;
  for (int64_t r0 = 0; r0 < ntiles; r0 += (int64_t)NTB * FIN_CH) {
    const int64_t t0 = r0 + (int64_t)threadIdx.x * FIN_CH;
    int c[FIN_CH];
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) c[q] = t0 + q < ntiles ? seg_cnt[t0 + q] : 0;
    int v = 0;
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) v += c[q];
    int tot;
    int64_t run = carry + block_excl_scan<NTB>(v, tot, s_scan);
#pragma unroll
    for (int q = 0; q < FIN_CH; ++q) {
      if (t0 + q < ntiles) soff[t0 + q] = run;
      run += c[q];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    st->nnz = carry;
    ch_off[0] = 0;
  }
}

// Chunked segment scan, first level: chunk-local offsets and chunk totals.
__global__ __launch_bounds__(NTB) void k_seg_chunks(const int* seg_cnt, int64_t ntiles, int64_t* soff,
                                                     int64_t* ch_cnt) {
  __shared__ int s_scan[WAVESB];
  const int64_t t0 = (int64_t)blockIdx.x * SEL_CHUNK + (int64_t)threadIdx.x * FIN_CH;
  int c[FIN_CH];
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) c[q] = t0 + q < ntiles ? seg_cnt[t0 + q] : 0;
  int v = 0;
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) v += c[q];
  int tot;
  int64_t run = block_excl_scan<NTB>(v, tot, s_scan);
#pragma unroll
  for (int q = 0; q < FIN_CH; ++q) {
    if (t0 + q < ntiles) soff[t0 + q] = run;
    run += c[q];
  }
  if (threadIdx.x == 0) ch_cnt[blockIdx.x] = tot;
}

// Second level: chunk offsets and the total.
__global__ __launch_bounds__(NTB) void k_seg_final(const int64_t* ch_cnt, int64_t nch, int64_t* ch_off,
                                                    OpStats* st) {
  __shared__ int s_scan[WAVESB];
  int64_t carry = 0;
  for (int64_t r0 = 0; r0 < nch; r0 += NTB) {
    const int64_t c = r0 + threadIdx.x;
    const int v = c < nch ? (int)ch_cnt[c] : 0;
    int tot;
    const int ex = block_excl_scan<NTB>(v, tot, s_scan);
    if (c < nch) ch_off[c] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) st->nnz = carry;
}

// Copy a segmented array to dense positions (soff from k_seg_scan).
__global__ __launch_bounds__(NT) void k_densify(const uint64_t* __restrict__ x, const double2* __restrict__ a,
                                                 const int* seg_cnt, const int64_t* soff, const int64_t* ch_off,
                                                 uint64_t* __restrict__ xo, double2* __restrict__ ao) {
  const int64_t t = blockIdx.x;
  const int cnt = seg_cnt[t];
  const int64_t src = t * SEL_TILE, dst = soff[t] + ch_off[t / SEL_CHUNK];
  for (int j = threadIdx.x; j < cnt; j += NT) {
    xo[dst + j] = x[src + j];
    ao[dst + j] = a[src + j];
  }
}

// sum |a|^2 and sum |a|^4 per tile.
__global__ __launch_bounds__(NT) void k_moments(const double2* __restrict__ a, int64_t n, double* tile_s2,
                                                 double* tile_s4) {
  __shared__ double s_red[WAVES];
  double s2 = 0.0, s4 = 0.0;
  // all loads first (clamped to a valid index), so they are in flight together
  double vr[SEL_ITEMS], vi[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = (int64_t)blockIdx.x * SEL_TILE + (int64_t)it * NT + threadIdx.x;
    const double2 v = a[i < n ? i : 0];
    vr[it] = v.x;
    vi[it] = v.y;
  }
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int64_t i = (int64_t)blockIdx.x * SEL_TILE + (int64_t)it * NT + threadIdx.x;
    if (i < n) {
      const double p = cabs2(make_double2(vr[it], vi[it]));
      s2 += p;
      s4 += p * p;
    }
  }
  const double t2 = block_sum(s2, s_red);
  const double t4 = block_sum(s4, s_red);
  if (threadIdx.x == 0) {
    tile_s2[blockIdx.x] = t2;
    tile_s4[blockIdx.x] = t4;
  }
}

// a[i] /= div for a real div, rounded as numpy does it: a[i] * (1 / div)
__global__ __launch_bounds__(NT) void k_scale(double2* __restrict__ a, int64_t n, double div) {
  const double inv = 1.0 / div;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    const double2 v = a[i];
    a[i] = make_double2(v.x * inv, v.y * inv);
  }
}

// Gather x_out[j] = x[idx[j]], a_out[j] = a[idx[j]].
__global__ __launch_bounds__(NT) void k_gather(const uint64_t* __restrict__ x, const double2* __restrict__ a,
                                                const int64_t* __restrict__ idx, int64_t m,
                                                uint64_t* __restrict__ xo, double2* __restrict__ ao) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t j = (int64_t)blockIdx.x * NT + threadIdx.x; j < m; j += stride) {
    const int64_t s = idx[j];
    xo[j] = x[s];
    ao[j] = a[s];
  }
}

// ----------------------------------------------------------- host side ----

// The select of one problem (tiles_end = its tiles; extract1 groups several
// tiles per block).
static void select_one(const SelProblem& one, const SelTiles& T, int64_t tiles_end, bool hist_ready, bool wide,
                       hipStream_t s, const LiftRecompute* fused, const LiftRecompute* relift) {
  const SelProblem* probs = nullptr;
  const BatchHdr* hdr = nullptr;
  const int nprob = 1;
  if (!hist_ready)
    hipLaunchKernelGGL((k_hist_psum<false>), dim3((unsigned)tiles_end), dim3(NT), 0, s, probs, nprob, one, T);
  hipLaunchKernelGGL((k_find1<false>), dim3(1), dim3(NTB), 0, s, probs, one, T, hdr);
  if (relift) launch_lift_recompute(2, *relift, one, T, 0, s);
  const int ext = extract1_tiles(tiles_end);
  const int64_t rows = (tiles_end + ext - 1) / ext;
  if (ext == EX1_TILES)
    hipLaunchKernelGGL((k_extract1<false, EX1_TILES>), dim3((unsigned)rows), dim3(NT), 0, s, probs, nprob, one,
                       T, tiles_end, ext, hdr);
  else
    hipLaunchKernelGGL((k_extract1<false, 0>), dim3((unsigned)rows), dim3(NT), 0, s, probs, nprob, one, T,
                       tiles_end, ext, hdr);
  if (fused) launch_lift_recompute(0, *fused, one, T, ext, s);
  hipLaunchKernelGGL((k_hist2_reduce<false>),
                     dim3(HIST_BINS / (4 * NT), (unsigned)((one.nrows + H2_ROWS - 1) / H2_ROWS), 1u), dim3(NT), 0,
                     s, probs, one, T, hdr);
  hipLaunchKernelGGL((k_find2<false>), dim3(1), dim3(NT), 0, s, probs, one, hdr);
  const unsigned wave_tiles = (unsigned)((tiles_end + WAVES - 1) / WAVES);
  {
    // 8 tiles per wave from 16384 tiles on (every wave still resident at once)
    const int tpw = tiles_end >= 16384 ? 8 : 1;
    const int64_t waves = (tiles_end + tpw - 1) / tpw;
    hipLaunchKernelGGL(k_extract2_agg, dim3((unsigned)((waves + WAVES - 1) / WAVES)), dim3(NT), 0, s, one, T,
                       tiles_end, tpw);
  }
  if (!wide) {
    hipLaunchKernelGGL(k_finish_staged, dim3(STG_STAGES * STG_BLOCKS), dim3(NTB), 0, s, one);
  } else {
    const dim3 wg(WIDE_BLOCKS, 1u);
    for (int lv = 0; lv < 4; ++lv) {
      hipLaunchKernelGGL((k_wide_hist<false>), wg, dim3(NT), 0, s, probs, one, lv);
      hipLaunchKernelGGL((k_wide_find<false>), dim3(1), dim3(NTB), 0, s, probs, one, lv);
    }
    hipLaunchKernelGGL((k_wide_collect<false>), wg, dim3(NT), 0, s, probs, one);
    for (int lv = 4; lv < WIDE_LEVELS; ++lv) {
      hipLaunchKernelGGL((k_wide_key_hist<false>), wg, dim3(NT), 0, s, probs, one, lv);
      hipLaunchKernelGGL((k_wide_find<false>), dim3(1), dim3(NTB), 0, s, probs, one, lv);
    }
  }
  hipLaunchKernelGGL((k_keptsum<false>), dim3(wave_tiles), dim3(NT), 0, s, probs, nprob, one, T, tiles_end, hdr);
  const int64_t nch = sel_chunks(one.npsum > one.ntiles ? one.npsum : one.ntiles);
  if (one.chunked) {
    hipLaunchKernelGGL(k_sel_chunks, dim3((unsigned)nch), dim3(NTB), 0, s, one, T);
    hipLaunchKernelGGL(k_sel_final, dim3(1), dim3(NTB), 0, s, one, T, nch);
  } else {
    hipLaunchKernelGGL((k_sel_finalize<false>), dim3(1), dim3(NTB), 0, s, probs, one, T, hdr);
  }
  if (fused)
    launch_lift_recompute(1, *fused, one, T, 0, s);
  else
    hipLaunchKernelGGL((k_compact<false>), dim3((unsigned)tiles_end), dim3(NT), 0, s, probs, nprob, one, T, hdr);
  if (one.chunked) {
    hipLaunchKernelGGL(k_sums_chunks, dim3((unsigned)sel_chunks(one.ntiles)), dim3(NTB), 0, s, one, T);
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(NTB), 0, s, one, T, sel_chunks(one.ntiles));
  } else {
    hipLaunchKernelGGL((k_sums_finalize<false>), dim3(1), dim3(NTB), 0, s, probs, one, T, hdr);
  }
}

// Tiles per k_extract1 block of a single problem.  From 16384 tiles on, the
// count in [EX1_TILES_MIN, EX1_TILES_MAX] whose blocks best fill whole rounds
// of the blocks the device holds at once (the last round of a launch otherwise
// runs partly empty; ties go to the larger count: fewer digit-2 rows for
// k_hist2_reduce to sum); else 4.
int extract1_tiles(int64_t ntiles) {
  if (ntiles < 16384) return EX1_TILES;
  static int slots = 0;  // resident k_extract1 blocks on the device (same value from every thread)
  if (slots == 0) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(&k_extract1<false, 0>), NT,
                                                     0) != hipSuccess ||
        cus * per <= 0) {
      (void)hipGetLastError();
      cus = 256;
      per = 4;
    }
    slots = cus * per;
  }
  int best = EX1_TILES_MAX;
  double best_eff = 0.0;
  for (int e = EX1_TILES_MAX; e >= EX1_TILES_MIN; --e) {
    const int64_t rows = (ntiles + e - 1) / e;
    const int64_t rounds = (rows + slots - 1) / slots;
    const double eff = (double)rows / (double)(rounds * slots);
    if (eff > best_eff + 1e-3) {
      best_eff = eff;
      best = e;
    }
  }
  return best;
}

int select_rows(int64_t ntiles, bool batch) {
  const int ext = batch ? 1 : extract1_tiles(ntiles);
  return (int)((ntiles + ext - 1) / ext);
}

int launch_select(const SelProblem& P, const SelTiles& T, bool hist_ready, bool wide, hipStream_t s,
                  const LiftRecompute* fused, const LiftRecompute* relift) {
  select_one(P, T, P.ntiles, hist_ready, wide, s, fused, relift);
  return (int)hipGetLastError();
}

// The select of a batch of problems (digit-1 histograms filled by the gate
// that produced the tiles), geometry from hdr; persistent grids sized by the
// tile / problem bounds.
int launch_select_batch(const SelProblem* d_probs, const BatchHdr* hdr, const SelTiles& T, int64_t tiles, int jobs,
                        hipStream_t s) {
  SelProblem none;
  memset(&none, 0, sizeof(none));
  const int64_t wtiles = (tiles + WAVES - 1) / WAVES;
  const int64_t h2 = (int64_t)(HIST_BINS / (4 * NT)) * (tiles / H2_ROWS + jobs + 1);
#define BASS_PG(fn, thr, bound) dim3(persist_grid(reinterpret_cast<const void*>(&fn), thr, bound))
  hipLaunchKernelGGL((k_find1<true>), BASS_PG(k_find1<true>, NTB, jobs), dim3(NTB), 0, s, d_probs, none, T, hdr);
  hipLaunchKernelGGL((k_extract1<true, 1>), BASS_PG((k_extract1<true, 1>), NT, tiles), dim3(NT), 0, s, d_probs, 0,
                     none, T, (int64_t)0, 1, hdr);
  hipLaunchKernelGGL((k_hist2_reduce<true>), BASS_PG(k_hist2_reduce<true>, NT, h2), dim3(NT), 0, s, d_probs, none, T,
                     hdr);
  hipLaunchKernelGGL((k_find2<true>), BASS_PG(k_find2<true>, NT, jobs), dim3(NT), 0, s, d_probs, none, hdr);
  hipLaunchKernelGGL((k_extract2<true>), BASS_PG(k_extract2<true>, NT, wtiles), dim3(NT), 0, s, d_probs, 0, none, T,
                     (int64_t)0, hdr);
  hipLaunchKernelGGL((k_finish<true>), BASS_PG(k_finish<true>, NTB, jobs), dim3(NTB), 0, s, d_probs, none, hdr);
  hipLaunchKernelGGL((k_keptsum<true>), BASS_PG(k_keptsum<true>, NT, wtiles), dim3(NT), 0, s, d_probs, 0, none, T,
                     (int64_t)0, hdr);
  hipLaunchKernelGGL((k_sel_finalize<true>), BASS_PG(k_sel_finalize<true>, NTB, jobs), dim3(NTB), 0, s, d_probs,
                     none, T, hdr);
  hipLaunchKernelGGL((k_compact<true>), BASS_PG(k_compact<true>, NT, tiles), dim3(NT), 0, s, d_probs, 0, none, T,
                     hdr);
  hipLaunchKernelGGL((k_sums_finalize<true>), BASS_PG(k_sums_finalize<true>, NTB, jobs), dim3(NTB), 0, s, d_probs,
                     none, T, hdr);
#undef BASS_PG
  return (int)hipGetLastError();
}

// Blocks of a kernel the device holds at once (occupancy x compute units),
// measured once per kernel.
int resident_blocks(const void* kernel, int threads) {
  static std::mutex mu;
  static std::unordered_map<const void*, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || cus * per <= 0) {
    (void)hipGetLastError();
    cus = 256;
    per = 2;
  }
  const int r = cus * per;
  cache[kernel] = r;
  return r;
}

int64_t sel_tiles(int64_t n) { return (n + SEL_TILE - 1) / SEL_TILE; }

int launch_seg_scan(const int* seg_cnt, int64_t ntiles, int64_t* soff, int64_t* ch_cnt, int64_t* ch_off,
                    OpStats* st, hipStream_t s) {
  const int64_t nch = sel_chunks(ntiles);
  if (nch <= 1) {
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(NTB), 0, s, seg_cnt, ntiles, soff, ch_off, st);
  } else {
    hipLaunchKernelGGL(k_seg_chunks, dim3((unsigned)nch), dim3(NTB), 0, s, seg_cnt, ntiles, soff, ch_cnt);
    hipLaunchKernelGGL(k_seg_final, dim3(1), dim3(NTB), 0, s, ch_cnt, nch, ch_off, st);
  }
  return (int)hipGetLastError();
}

int launch_densify(const uint64_t* x, const double2* a, const int* seg_cnt, const int64_t* soff,
                   const int64_t* ch_off, int64_t ntiles, uint64_t* xo, double2* ao, hipStream_t s) {
  if (ntiles <= 0) return 0;
  hipLaunchKernelGGL(k_densify, dim3((unsigned)ntiles), dim3(NT), 0, s, x, a, seg_cnt, soff, ch_off, xo, ao);
  return (int)hipGetLastError();
}

int launch_moments(const double2* a, int64_t n, double* tile_s2, double* tile_s4, OpStats* st,
                   hipStream_t s) {
  const int64_t tiles = sel_tiles(n);
  if (tiles > 0)
    hipLaunchKernelGGL(k_moments, dim3((unsigned)tiles), dim3(NT), 0, s, a, n, tile_s2, tile_s4);
  hipLaunchKernelGGL(k_sums_reduce, dim3(1), dim3(NTB), 0, s, tile_s2, tile_s4, tiles, st);
  return (int)hipGetLastError();
}

int launch_scale(double2* a, int64_t n, double div, hipStream_t s) {
  if (n <= 0) return 0;
  int64_t blocks = (n + NT - 1) / NT;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(NT), 0, s, a, n, div);
  return (int)hipGetLastError();
}

int launch_gather(const uint64_t* x, const double2* a, const int64_t* idx, int64_t m, uint64_t* xo,
                  double2* ao, hipStream_t s) {
  if (m <= 0) return 0;
  int64_t blocks = (m + NT - 1) / NT;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)blocks), dim3(NT), 0, s, x, a, idx, m, xo, ao);
  return (int)hipGetLastError();
}

}  // namespace bass
