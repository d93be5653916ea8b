// Gate application on a key-ordered state (the simulator's internal layout).
//
// The simulator keeps its entries sorted by a stored word w: the basis key
// with its bits permuted (the host records which qubit sits at which bit; the
// bit VALUES are the key's).  Applying a gate to the qubit stored at bit j
// "lifts" bit j down to bit t (t = 0 for a 1-qubit gate; t = 1 for a 2-qubit
// gate, whose other qubit must already sit at bit 0):
//
//     T(w) = hi | mid << 1 | b << t | lo
//     hi = bits above j,  b = bit j,  mid = bits t..j-1,  lo = bits below t.
//
// T keeps the order of entries that agree in bit j.  Entries sharing hi form
// a run; inside a run the part with b = 0 (A) precedes the part with b = 1
// (B), and the entries sorted by T are the merge of A and B.  After the merge
// the members of every gate group (entries that differ only in bits 0..t of
// T) are adjacent, so the gate is applied to consecutive entries and its
// outputs come out sorted by T: the output is again an ordered state, with
// the gated qubit(s) at the lowest bits.
//
// Kernels (TM merged positions per tile; runs coincide in input and output):
//   k_lift_part : one wave per tile boundary d finds the run around d and the
//                 merge split at d with 64-ary searches.
//   k_lift_gate : per tile, loads its <= 3 input ranges (jointly sorted by w),
//                 merges the entries with bit j clear and set (merge path in
//                 LDS) into T order, applies the gate per
//                 group and writes the tile's outputs to its own segment
//                 [tile * SEG_ENTRIES, +count) with the digit-1 histogram and
//                 the |a|^2 sum that the top-k select consumes.
//   k_lift_move : the same permutation without a gate (dense output).
//
// Group semantics follow the hash-table kernels (gate_kernels.hip): a gate
// output exists when one of its products is nonzero, products are summed in
// input order (member index order here, which is input order because the
// input is sorted); rotations sum every member (absent = 0) and drop outputs
// with |a| <= 1e-30.
#include "common.h"
#include "extract_core.h"
#include "kernels.h"

namespace bass {

__device__ __forceinline__ uint64_t lift_key(uint64_t w, int j, int t) {
  const uint64_t lowmask = (1ull << t) - 1ull;
  const uint64_t hi = w & ~((2ull << j) - 1ull);
  const uint64_t b = (w >> j) & 1ull;
  const uint64_t mid = w & ((1ull << j) - 1ull) & ~lowmask;
  return hi | (mid << 1) | (b << t) | (w & lowmask);
}

// First index in [lo, hi) where the monotone (true ... false) predicate is
// false, or hi.  Called by all 64 lanes of a wave with equal arguments.
template <class P>
__device__ __forceinline__ int64_t wave_partition_point(int64_t lo, int64_t hi, P pred) {
  const int lane = threadIdx.x & 63;
  while (hi - lo > 64) {
    const int64_t step = (hi - lo + 63) / 64;
    const int64_t p = lo + (int64_t)lane * step;
    const bool t = p < hi && pred(p);
    const int c = __popcll(__ballot(t));
    if (c == 0) return lo;
    const int64_t nlo = lo + (int64_t)(c - 1) * step + 1;
    const int64_t nhi = lo + (int64_t)c * step;
    lo = nlo;
    hi = nhi < hi ? nhi : hi;
  }
  const int64_t p = lo + lane;
  const bool t = p < hi && pred(p);
  return lo + __popcll(__ballot(t));
}

// As wave_partition_point, but with 8 probes per step while the range exceeds
// 64 entries: a step then touches 8 cache lines instead of 64, for about twice
// the steps.  For searches whose range is already narrow (bounded by known
// neighbours), where memory traffic rather than latency sets the cost.
template <class P>
__device__ __forceinline__ int64_t wave_partition_point8(int64_t lo, int64_t hi, P pred) {
  const int lane = threadIdx.x & 63;
  while (hi - lo > 64) {
    const int64_t step = (hi - lo + 7) / 8;
    const int64_t p = lo + (int64_t)lane * step;
    const bool t = lane < 8 && p < hi && pred(p);
    const int c = __popcll(__ballot(t));
    if (c == 0) return lo;
    const int64_t nlo = lo + (int64_t)(c - 1) * step + 1;
    const int64_t nhi = lo + (int64_t)c * step;
    lo = nlo;
    hi = nhi < hi ? nhi : hi;
  }
  const int64_t p = lo + lane;
  const bool t = p < hi && pred(p);
  return lo + __popcll(__ballot(t));
}

// The run around merged position d (clamped to n) and the merge split there.
// Called by all 64 lanes of a wave.
__device__ __forceinline__ LiftBound lift_boundary(const uint64_t* __restrict__ w, int64_t n, int j, int t, int64_t d) {
  if (d > n) d = n;
  LiftBound r;
  if (d >= n) {
    r.s0 = r.sm = r.s1 = r.a = r.b = (int)n;
  } else {
    const uint64_t bj = 1ull << j;
    const uint64_t hi = w[d] & ~((bj << 1) - 1ull);
    // a run that starts at 0 or ends at n (one run over everything when bit j
    // is the top stored bit) is seen from the end entries, without a search
    const int64_t s0 = w[0] < hi ? wave_partition_point(0, d, [&](int64_t i) { return w[i] < hi; }) : 0;
    const int64_t s1 =
        w[n - 1] < hi + (bj << 1) ? n : wave_partition_point(d + 1, n, [&](int64_t i) { return w[i] < hi + (bj << 1); });
    const int64_t sm = wave_partition_point(s0, s1, [&](int64_t i) { return w[i] < (hi | bj); });
    const int64_t ld = d - s0, na = sm - s0, nb = s1 - sm;
    const int64_t lo = ld - nb > 0 ? ld - nb : 0;
    const int64_t hh = ld < na ? ld : na;
    // number of A entries among the first ld merged entries of the run
    const int64_t i = wave_partition_point(
        lo, hh, [&](int64_t q) { return lift_key(w[s0 + q], j, t) < lift_key(w[sm + ld - 1 - q], j, t); });
    r.s0 = (int)s0;
    r.sm = (int)sm;
    r.s1 = (int)s1;
    r.a = (int)(s0 + i);
    r.b = (int)(sm + ld - i);
  }
  r.pad0 = r.pad1 = r.pad2 = 0;
  return r;
}

// Block (0, 0) clears the lift's zero ranges (every thread, before any exit).
__device__ __forceinline__ void lift_zero(const LiftZero& z) {
  if (z.p == nullptr || blockIdx.x != 0 || blockIdx.y != 0) return;
  for (int i = threadIdx.x; i < z.wp + z.wq; i += NT) {
    if (i < z.wp)
      z.p[i] = 0ull;
    else
      z.q[i - z.wp] = 0ull;
  }
}

// Tile boundaries of the lift of bit j0 + blockIdx.y (one bit per grid row;
// row r writes bounds[r * nbound + u]).
__global__ __launch_bounds__(NT) void k_lift_part(const uint64_t* __restrict__ w, int64_t n, int j0, int t,
                                                   int tm, int64_t nbound, LiftBound* bounds,
                                                   const int64_t* __restrict__ n_dev, LiftZero zero) {
  lift_zero(zero);
  n = dev_count(n_dev, n);
  const int lane = threadIdx.x & 63;
  const int j = j0 + (int)blockIdx.y;
  bounds += (int64_t)blockIdx.y * nbound;
  const int64_t u = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (u >= nbound) return;  // uniform per wave
  const LiftBound r = lift_boundary(w, n, j, t, u * (int64_t)tm);
  if (lane == 0) bounds[u] = r;
}

// The same boundary when the boundaries D0 (at a merged position <= d) and
// D1 (>= d) are known: run start and end, A/B split and both merge cursors
// are monotone in the merged position, so every search is confined between
// the two (empty when d lies in the run of both).  The confined searches take
// 8 probes per step (wave_partition_point8).
__device__ __forceinline__ LiftBound lift_boundary_between(const uint64_t* __restrict__ w, int64_t n, int j, int t,
                                                           int64_t d, const LiftBound& D0, const LiftBound& D1) {
  if (d > n) d = n;
  LiftBound r;
  if (d >= n) {
    r.s0 = r.sm = r.s1 = r.a = r.b = (int)n;
  } else {
    const uint64_t bj = 1ull << j;
    const uint64_t hi = w[d] & ~((bj << 1) - 1ull);
    const int64_t top = d < (int64_t)D1.s0 ? d : (int64_t)D1.s0;
    const int64_t s0 = wave_partition_point8(D0.s0, top, [&](int64_t i) { return w[i] < hi; });
    const int64_t lo1 = d + 1 > (int64_t)D0.s1 ? d + 1 : (int64_t)D0.s1;
    const int64_t s1 = wave_partition_point8(lo1, D1.s1, [&](int64_t i) { return w[i] < hi + (bj << 1); });
    const int64_t lom = s0 > (int64_t)D0.sm ? s0 : (int64_t)D0.sm;
    const int64_t him = s1 < (int64_t)D1.sm ? s1 : (int64_t)D1.sm;
    const int64_t sm = wave_partition_point8(lom, him, [&](int64_t i) { return w[i] < (hi | bj); });
    const int64_t ld = d - s0, na = sm - s0, nb = s1 - sm;
    int64_t lo = ld - nb > 0 ? ld - nb : 0;
    int64_t hh = ld < na ? ld : na;
    // A cursor a = s0 + i and B cursor b = sm + ld - i both lie between the
    // neighbours' cursors (a neighbour in another run bounds nothing here).
    if ((int64_t)D0.a - s0 > lo) lo = (int64_t)D0.a - s0;
    if ((int64_t)D1.a - s0 < hh) hh = (int64_t)D1.a - s0;
    if (ld - ((int64_t)D1.b - sm) > lo) lo = ld - ((int64_t)D1.b - sm);
    if (ld - ((int64_t)D0.b - sm) < hh) hh = ld - ((int64_t)D0.b - sm);
    const int64_t i = wave_partition_point8(
        lo, hh, [&](int64_t q) { return lift_key(w[s0 + q], j, t) < lift_key(w[sm + ld - 1 - q], j, t); });
    r.s0 = (int)s0;
    r.sm = (int)sm;
    r.s1 = (int)s1;
    r.a = (int)(s0 + i);
    r.b = (int)(sm + ld - i);
  }
  r.pad0 = r.pad1 = r.pad2 = 0;
  return r;
}

// Lifts of up to LIFT_ONE_MAX boundaries launch the boundary passes as one
// kernel (k_lift_part_one): it saves a launch (about 3 us of a 100 us gate at
// 2^22 entries) but searches each coarse boundary twice and holds 14 of its 16
// waves while the coarse searches run, which costs more than the launch from
// about 2^24 entries on (2^26: 17 us more per 1-qubit gate).
constexpr int64_t LIFT_ONE_MAX = 16385;

// Tile boundaries in two passes: every LIFT_CS-th boundary (and the last) by
// full searches, then every boundary between its two coarse neighbours.  The
// full 64-ary searches touch a cache line per probe; confining them cuts the
// boundary pass's memory traffic about LIFT_CS-fold.  Grid row y = bit j0 + y.
constexpr int LIFT_CS = 16;

__host__ __device__ inline int64_t lift_coarse_count(int64_t nbound) { return (nbound - 1 + LIFT_CS - 1) / LIFT_CS + 1; }

__global__ __launch_bounds__(NT) void k_lift_part_coarse(const uint64_t* __restrict__ w, int64_t n, int j0, int t,
                                                          int tm, int64_t nbound, LiftBound* coarse,
                                                          const int64_t* __restrict__ n_dev, LiftZero zero) {
  lift_zero(zero);
  n = dev_count(n_dev, n);
  const int lane = threadIdx.x & 63;
  const int j = j0 + (int)blockIdx.y;
  const int64_t nc = lift_coarse_count(nbound);
  coarse += (int64_t)blockIdx.y * nc;
  const int64_t c = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (c >= nc) return;  // uniform per wave
  int64_t b = c * LIFT_CS;
  if (b > nbound - 1) b = nbound - 1;
  const LiftBound r = lift_boundary(w, n, j, t, b * (int64_t)tm);
  if (lane == 0) coarse[c] = r;
}

__global__ __launch_bounds__(NT) void k_lift_part_fine(const uint64_t* __restrict__ w, int64_t n, int j0, int t,
                                                        int tm, int64_t nbound, const LiftBound* __restrict__ coarse,
                                                        LiftBound* bounds, const int64_t* __restrict__ n_dev) {
  n = dev_count(n_dev, n);
  const int lane = threadIdx.x & 63;
  const int j = j0 + (int)blockIdx.y;
  const int64_t nc = lift_coarse_count(nbound);
  coarse += (int64_t)blockIdx.y * nc;
  bounds += (int64_t)blockIdx.y * nbound;
  const int64_t u = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (u >= nbound) return;  // uniform per wave
  const int64_t c0 = u / LIFT_CS;
  const int64_t c1 = c0 + 1 < nc ? c0 + 1 : nc - 1;
  const int64_t b1 = c1 * LIFT_CS < nbound - 1 ? c1 * LIFT_CS : nbound - 1;
  LiftBound r;
  if (u == c0 * LIFT_CS)
    r = coarse[c0];
  else if (u == b1)
    r = coarse[c1];
  else
    r = lift_boundary_between(w, n, j, t, u * (int64_t)tm, coarse[c0], coarse[c1]);
  if (lane == 0) bounds[u] = r;
}

// Both passes in one launch: block b (LIFT_CS waves) writes boundaries
// [b * LIFT_CS, min((b + 1) * LIFT_CS, nbound - 1)) -- and the last block the
// final one -- after its waves 0 and 1 have searched the two coarse
// boundaries around them (each coarse boundary is searched by the two blocks
// that share it); the same values as the two-launch form.
__global__ __launch_bounds__(LIFT_CS * 64) void k_lift_part_one(const uint64_t* __restrict__ w, int64_t n, int j0,
                                                                int t, int tm, int64_t nbound, LiftBound* bounds,
                                                                const int64_t* __restrict__ n_dev, LiftZero zero) {
  __shared__ LiftBound s_c[2];
  lift_zero(zero);
  n = dev_count(n_dev, n);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j = j0 + (int)blockIdx.y;
  bounds += (int64_t)blockIdx.y * nbound;
  const int64_t u0 = (int64_t)blockIdx.x * LIFT_CS;
  const int64_t u1 = u0 + LIFT_CS < nbound - 1 ? u0 + LIFT_CS : nbound - 1;
  if (wv < 2) {
    const LiftBound r = lift_boundary(w, n, j, t, (wv == 0 ? u0 : u1) * (int64_t)tm);
    if (lane == 0) s_c[wv] = r;
  }
  __syncthreads();
  const int64_t u = u0 + wv;
  if (u < u1) {
    const LiftBound r = wv == 0 ? s_c[0] : lift_boundary_between(w, n, j, t, u * (int64_t)tm, s_c[0], s_c[1]);
    if (lane == 0) bounds[u] = r;
  }
  if (u1 == nbound - 1 && wv == 0 && lane == 0) bounds[u1] = s_c[1];
}

// Boundaries of rows j0 .. j0 + rows - 1 (nbound per row) into bounds; the
// coarse pass uses `coarse` (lift_coarse_count(nbound) per row).  Small
// lifts take the direct pass.
static void lift_partition(const uint64_t* x, int64_t n, int j0, int rows, int t, int tm, int64_t nbound,
                           LiftBound* bounds, LiftBound* coarse, hipStream_t s, const int64_t* n_dev = nullptr,
                           const LiftZero& zero = LiftZero()) {
  if (nbound <= 4 * LIFT_CS || coarse == nullptr) {
    hipLaunchKernelGGL(k_lift_part, dim3((unsigned)((nbound + WAVES - 1) / WAVES), (unsigned)rows), dim3(NT), 0, s,
                       x, n, j0, t, tm, nbound, bounds, n_dev, zero);
    return;
  }
  if (nbound <= LIFT_ONE_MAX) {
    hipLaunchKernelGGL(k_lift_part_one, dim3((unsigned)((nbound - 1 + LIFT_CS - 1) / LIFT_CS), (unsigned)rows),
                       dim3(LIFT_CS * 64), 0, s, x, n, j0, t, tm, nbound, bounds, n_dev, zero);
    return;
  }
  const int64_t nc = lift_coarse_count(nbound);
  hipLaunchKernelGGL(k_lift_part_coarse, dim3((unsigned)((nc + WAVES - 1) / WAVES), (unsigned)rows), dim3(NT), 0, s,
                     x, n, j0, t, tm, nbound, coarse, n_dev, zero);
  hipLaunchKernelGGL(k_lift_part_fine, dim3((unsigned)((nbound + WAVES - 1) / WAVES), (unsigned)rows), dim3(NT), 0,
                     s, x, n, j0, t, tm, nbound, (const LiftBound*)coarse, bounds, n_dev);
}

__device__ __forceinline__ int lds_lower_bound(const uint64_t* s, int cnt, uint64_t v) {
  int lo = 0, len = cnt;
  while (len > 0) {
    const int half = len >> 1;
    const bool go = s[lo + half] < v;
    lo = go ? lo + half + 1 : lo;
    len = go ? len - half - 1 : half;
  }
  return lo;
}

// Load the tile's entries (merged positions [tile*TM, +cnt)) and place them in
// merged order: s_t = lifted keys, s_a = amplitudes.  Returns cnt.  Ends with
// a barrier.
//
// The tile's entries are loaded in w order.  Those with bit j clear (A) and
// those with it set (B) are each already ordered by the lifted key T (T keeps
// the order of entries that agree in bit j), so the merged order is the merge
// of the two lists: a stable partition (one scan) writes A then B into s_t,
// each thread finds the split of its ITEMS merged positions on the merge path
// (one binary search) and merges them, recording every entry's merged rank in
// s_rank; the loading threads then scatter keys and amplitudes to their ranks.
// s_scan holds ITEMS * WAVES ints.
template <int TM>
__device__ __forceinline__ int lift_stage(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                          int64_t n, int j, int t, const LiftBound& B0, const LiftBound& B1,
                                          int64_t tile, unsigned short* s_rank, int* s_scan, uint64_t* s_t,
                                          double2* s_a) {
  constexpr int ITEMS = TM / NT;
  const int64_t d = tile * TM;
  const int cnt = (int)((n - d) < TM ? (n - d) : TM);
  int64_t p1lo, p1n, p2lo, p2n, p3lo;
  if (B0.s0 == B1.s0) {  // the tile lies inside one run
    p1lo = B0.a;
    p1n = B1.a - B0.a;
    p2lo = B0.b;
    p2n = B1.b - B0.b;
    p3lo = 0;
  } else {
    p1lo = B0.a;
    p1n = B0.sm - B0.a;
    p2lo = B0.b;
    p2n = B1.a - B0.b;
    p3lo = B1.sm;
  }
  // private arrays as plain doubles: arrays of the vector type double2 stay
  // in scratch memory
  uint64_t tk[ITEMS];
  double ar_re[ITEMS], ar_im[ITEMS];
  int fb[ITEMS], pb[ITEMS], e[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int idx = it * NT + threadIdx.x;
    tk[it] = 0;
    ar_re[it] = ar_im[it] = 0.0;
    fb[it] = 0;
    if (idx < cnt) {
      const int64_t g = idx < p1n ? p1lo + idx : (idx < p1n + p2n ? p2lo + (idx - p1n) : p3lo + (idx - p1n - p2n));
      const uint64_t x = w[g];
      const double2 v = a[g];
      ar_re[it] = v.x;
      ar_im[it] = v.y;
      fb[it] = (int)((x >> j) & 1ull);
      tk[it] = lift_key(x, j, t);
    }
  }
  int nb;
  block_excl_scan_items<ITEMS>(fb, pb, nb, s_scan);
  const int na = cnt - nb;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int idx = it * NT + threadIdx.x;
    e[it] = fb[it] ? na + pb[it] : idx - pb[it];
    if (idx < cnt) s_t[e[it]] = tk[it];
  }
  __syncthreads();
  {
    const uint64_t* A = s_t;
    const uint64_t* B = s_t + na;
    const int d0 = threadIdx.x * ITEMS;
    if (d0 < cnt) {
      int lo = d0 - nb > 0 ? d0 - nb : 0;
      int hi = d0 < na ? d0 : na;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] < B[d0 - 1 - mid])
          lo = mid + 1;
        else
          hi = mid;
      }
      int ia = lo, ib = d0 - lo;
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int m = d0 + k;
        if (m < cnt) {
          if (ib >= nb || (ia < na && A[ia] < B[ib]))
            s_rank[ia++] = (unsigned short)m;
          else
            s_rank[na + ib++] = (unsigned short)m;
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int idx = it * NT + threadIdx.x;
    if (idx < cnt) {
      const int r = s_rank[e[it]];
      s_t[r] = tk[it];
      s_a[r] = make_double2(ar_re[it], ar_im[it]);
    }
  }
  __syncthreads();
  return cnt;
}

// Groups of a staged tile (QB gate bits at the bottom of the lifted key):
// the groups of the entries merged just before the tile, whose members the
// tile does not own, and the members of the tile's last group merged after
// the tile.  Written by thread 0 of the block.
struct LiftHalo {
  uint64_t prev[2];
  uint64_t t[3];
  double2 a[3];
  int n;
};

template <int QB>
__device__ __forceinline__ uint64_t lift_group_of(uint64_t tk) {
  return tk >> QB;
}

// Thread 0 fills H; the caller synchronises.
template <int QB>
__device__ __forceinline__ void lift_halo_setup(const uint64_t* __restrict__ w, const double2* __restrict__ a, int j,
                                                const LiftBound& B0, const LiftBound& B1, const uint64_t* s_t,
                                                int cnt, LiftHalo& H) {
  constexpr int T = QB - 1;
  constexpr uint64_t NONE = ~0ull;
  if (threadIdx.x != 0) return;
  // group of the entry merged just before the tile (A or B tail of the run)
  H.prev[0] = B0.a > B0.s0 ? lift_group_of<QB>(lift_key(w[B0.a - 1], j, T)) : NONE;
  H.prev[1] = B0.b > B0.sm ? lift_group_of<QB>(lift_key(w[B0.b - 1], j, T)) : NONE;
  // members of the tile's last group merged after the tile
  int nh = 0;
  const uint64_t glast = cnt > 0 ? lift_group_of<QB>(s_t[cnt - 1]) : NONE;
  const int64_t cand[3] = {B1.a < B1.sm ? (int64_t)B1.a : -1, B1.b < B1.s1 ? (int64_t)B1.b : -1,
                           B1.b + 1 < B1.s1 ? (int64_t)B1.b + 1 : -1};
  for (int c = 0; c < 3; ++c) {
    if (cand[c] < 0) continue;
    const uint64_t tk = lift_key(w[cand[c]], j, T);
    if (lift_group_of<QB>(tk) == glast) {
      H.t[nh] = tk;
      H.a[nh] = a[cand[c]];
      ++nh;
    }
  }
  H.n = nh;
}

// If merged position m heads its group (the tile owns the group), gathers the
// member amplitudes by member index (absent members are 0) and their presence
// mask (bit c set if member c exists) and returns true.  The amplitudes are
// written through unrolled selects and presence is a bit mask, so everything
// stays in registers.
template <int QB>
__device__ __forceinline__ bool lift_group_gather(int m, int cnt, const uint64_t* s_t, const double2* s_a,
                                                  const LiftHalo& H, double (&avr)[1 << QB], double (&avi)[1 << QB],
                                                  unsigned int& pres) {
  constexpr int M = 1 << QB;
  const uint64_t g = lift_group_of<QB>(s_t[m]);
  const bool head = m == 0 ? (g != H.prev[0] && g != H.prev[1]) : g != lift_group_of<QB>(s_t[m - 1]);
  if (!head) return false;
#pragma unroll
  for (int c = 0; c < M; ++c) avr[c] = avi[c] = 0.0;
  pres = 0;
  int k = m;
#pragma unroll
  for (int s = 0; s < M; ++s) {
    if (k < cnt && lift_group_of<QB>(s_t[k]) == g) {
      const int c = (int)(s_t[k] & (M - 1));
      const double2 v = s_a[k];
#pragma unroll
      for (int cc = 0; cc < M; ++cc) {
        if (c == cc) {
          avr[cc] = v.x;
          avi[cc] = v.y;
        }
      }
      pres |= 1u << c;
      ++k;
    }
  }
  if (k == cnt) {
#pragma unroll
    for (int h = 0; h < 3; ++h) {
      if (h < H.n) {
        const int c = (int)(H.t[h] & (M - 1));
        const double2 v = H.a[h];
#pragma unroll
        for (int cc = 0; cc < M; ++cc) {
          if (c == cc) {
            avr[cc] = v.x;
            avi[cc] = v.y;
          }
        }
        pres |= 1u << c;
      }
    }
  }
  return true;
}

// LDS of one lift tile (QB gate bits, TM = SEG_ENTRIES >> QB merged input
// positions): merge ranks, staged lifted keys and amplitudes, scan and
// reduction scratch, the halo.
template <int QB>
struct LiftTileLds {
  static constexpr int TM = SEG_ENTRIES >> QB;
  static constexpr int ITEMS = TM / NT;
  unsigned short rank[TM];
  uint64_t t[TM];
  double2 a[TM];
  int scan[ITEMS * WAVES];
  double red[WAVES];
  LiftHalo halo;
};

// A lift tile's outputs in registers.  Item it (merged position it * NT +
// threadIdx.x, when it heads its group) holds the M = 2^QB outputs of its
// group; output o exists when bit o of mask[it] is set.  The existing outputs
// of item it take the tile's output slots off[it], off[it] + 1, ... in o order
// (slots run item-major, then thread order); total = the tile's output count.
template <int QB>
struct LiftTileOut {
  static constexpr int M = 1 << QB;
  static constexpr int ITEMS = (SEG_ENTRIES >> QB) / NT;
  uint64_t key[ITEMS][M];
  double re[ITEMS][M], im[ITEMS][M];
  unsigned int mask[ITEMS];
  int off[ITEMS];
  int total;
};

// Stages, merges and gates local tile `tile` of the input (merged positions
// [tile * TM, ...)) into Q.  Ends with a barrier; L.t and L.a are free after it.
// PRUNE: every output's |a|^2 is added to psum (in the stored lift's order)
// and to the digit-1 histogram (LDS bins below LBINS, else hist), and only
// outputs in bucket b1p - margin or above (all when b1p is 0xFFFFFFFF)
// enter Q's mask and offsets.
template <int QB, int MODE, bool PRUNE = false>
__device__ __forceinline__ void lift_tile_compute(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                                  int64_t n, int j, const LiftBound& B0, const LiftBound& B1,
                                                  const GateMat& G, int64_t tile, LiftTileLds<QB>& L,
                                                  LiftTileOut<QB>& Q, unsigned int b1p = 0, unsigned int margin = 0,
                                                  double* psum = nullptr,
                                                  unsigned int* s_hist = nullptr, unsigned int* hist = nullptr) {
  constexpr int M = 1 << QB;
  constexpr int T = QB - 1;
  constexpr int TM = SEG_ENTRIES / M;
  constexpr int ITEMS = TM / NT;
  const int cnt = lift_stage<TM>(w, a, n, j, T, B0, B1, tile, L.rank, L.scan, L.t, L.a);
  lift_halo_setup<QB>(w, a, j, B0, B1, L.t, cnt, L.halo);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int m = it * NT + threadIdx.x;
    unsigned int mask = 0;
    double avr[M], avi[M];
    unsigned int pres;
    if (m < cnt && lift_group_gather<QB>(m, cnt, L.t, L.a, L.halo, avr, avi, pres)) {
      const uint64_t base = L.t[m] & ~(uint64_t)(M - 1);
#pragma unroll
      for (int o = 0; o < M; ++o) {
        double2 acc = make_double2(0.0, 0.0);
        bool any = false;
        if constexpr (MODE == MODE_GATE) {
#pragma unroll
          for (int c = 0; c < M; ++c) {
            const double2 term = cmul(G.m[o * M + c], make_double2(avr[c], avi[c]));
            if (((pres >> c) & 1u) && cnonzero(term)) {
              acc = any ? cadd(acc, term) : term;
              any = true;
            }
          }
        } else {
          acc = cmul(G.m[o * M + 0], make_double2(avr[0], avi[0]));
#pragma unroll
          for (int c = 1; c < M; ++c) acc = cadd(acc, cmul(G.m[o * M + c], make_double2(avr[c], avi[c])));
          any = hypot(acc.x, acc.y) > 1e-30;
        }
        Q.key[it][o] = base | (uint64_t)o;
        Q.re[it][o] = acc.x;
        Q.im[it][o] = acc.y;
        if constexpr (PRUNE) {
          if (any) {
            const double p = cabs2(acc);
            *psum += p;
            const unsigned int bin = (unsigned int)(__double_as_longlong(p) >> 52) & (HIST_BINS - 1);
            if (bin < HIST_BINS / 2)
              atomicAdd(&s_hist[bin], 1u);
            else
              atomicAdd(&hist[bin], 1u);
            if (b1p == 0xFFFFFFFFu || bin + margin >= b1p) mask |= 1u << o;
          }
        } else {
          if (any) mask |= 1u << o;
        }
      }
    }
    Q.mask[it] = mask;
  }
  // output offsets in merged (item-major) order: one scan for all items
  int ocnt[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) ocnt[it] = __popc(Q.mask[it]);
  block_excl_scan_items<ITEMS>(ocnt, Q.off, Q.total, L.scan);
}

// Per-output values v[it][o] (existing outputs only) read back in slot order:
// out[s] = the value of slot s * NT + threadIdx.x, 0 past the tile's outputs.
// buf: LDS for CAP = 2 * TM values of 8 bytes (the staged amplitudes' space,
// free once the outputs are computed); SEG_ENTRIES / CAP passes, each ending
// with a barrier.
template <int QB, class V>
__device__ __forceinline__ void lift_to_slots(const LiftTileOut<QB>& Q,
                                              const V (&v)[LiftTileOut<QB>::ITEMS][LiftTileOut<QB>::M], V* buf,
                                              V (&out)[SEL_ITEMS]) {
  constexpr int M = 1 << QB;
  constexpr int TM = SEG_ENTRIES / M;
  constexpr int ITEMS = TM / NT;
  constexpr int CAP = 2 * TM;
  static_assert(SEG_ENTRIES % CAP == 0 && CAP % NT == 0, "passes cover whole items");
#pragma unroll
  for (int pass = 0; pass < SEG_ENTRIES / CAP; ++pass) {
    const int lo = pass * CAP;
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      int pos = Q.off[it];
#pragma unroll
      for (int o = 0; o < M; ++o) {
        if ((Q.mask[it] >> o) & 1u) {
          if (pos >= lo && pos < lo + CAP) buf[pos - lo] = v[it][o];
          ++pos;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = lo / NT; it < (lo + CAP) / NT; ++it) {
      const int jj = it * NT + threadIdx.x;
      out[it] = jj < Q.total ? buf[jj - lo] : V(0);
    }
    __syncthreads();
  }
}

// One tile of a gate: local tile `tile` of the input (merged positions
// [tile*TM, ...)), outputs to segment `otile` of (xo, ao), seg_cnt[otile],
// tile_psum[otile]; digit-1 histogram into hist.  SPEC: also the select's
// digit-1 pass on the tile (ex1_tile_core, bucket *S.pred) from its |a|^2
// values staged in LDS in slot order.  FUSE: the outputs are not stored (a
// fused select recomputes them); the speculative pass keeps the candidates'
// keys (S.out.c1k).  PRUNE 1: the histogram and |a|^2 sum over every output,
// but only outputs in bucket *S.pred - S.margin or above take a slot (in
// the same order); PRUNE 2 (a relift): every output stored, no histogram or
// sum (the pruned lift made them).
template <int QB, int MODE, bool SPEC = false, bool FUSE = false, int PRUNE = 0>
__device__ __forceinline__ void lift_gate_tile(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                               int64_t n, int j, const LiftBound& B0, const LiftBound& B1,
                                               const GateMat& G, int64_t tile, int64_t otile,
                                               uint64_t* __restrict__ xo, double2* __restrict__ ao,
                                               unsigned int* hist, double* tile_psum, int* seg_cnt,
                                               LiftSpec S = LiftSpec()) {
  constexpr int M = 1 << QB;
  constexpr int TM = SEG_ENTRIES / M;
  constexpr int ITEMS = TM / NT;
  __shared__ LiftTileLds<QB> L;
  // digit-1 bins below 1024 (|a|^2 < 2, every normalised state) count in LDS;
  // higher bins go straight to the global histogram
  constexpr int LBINS = HIST_BINS / 2;
  __shared__ unsigned int s_hist[LBINS];

  for (int b = threadIdx.x; b < LBINS; b += NT) s_hist[b] = 0;
  LiftTileOut<QB> Q;
  double psum = 0.0;
  if constexpr (PRUNE == 1)
    lift_tile_compute<QB, MODE, true>(w, a, n, j, B0, B1, G, tile, L, Q, *S.pred, S.margin, &psum, s_hist, hist);
  else
    lift_tile_compute<QB, MODE>(w, a, n, j, B0, B1, G, tile, L, Q);
  if (threadIdx.x == 0) seg_cnt[otile] = Q.total;

  // SPEC: |a|^2 of the outputs, staged in L.a (free: every gather ended
  // before the offset scan's barrier) in slot order, CAP slots per pass; one
  // pass (1 qubit) stages them here, two (2 qubits) keep them in pv
  constexpr int CAP = 2 * TM;
  constexpr bool ONE_PASS = CAP == SEG_ENTRIES;
  double* s_p = reinterpret_cast<double*>(L.a);
  double pv[ITEMS][M];
  const int64_t rb = otile * SEG_ENTRIES;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    int64_t pos = rb + Q.off[it];
#pragma unroll
    for (int o = 0; o < M; ++o) {
      if ((Q.mask[it] >> o) & 1u) {
        const double2 v = make_double2(Q.re[it][o], Q.im[it][o]);
        if constexpr (!FUSE) {
          xo[pos] = Q.key[it][o];
          ao[pos] = v;
        }
        const double p = cabs2(v);
        if constexpr (SPEC && ONE_PASS) s_p[pos - rb] = p;
        if constexpr (SPEC && !ONE_PASS) pv[it][o] = p;
        if constexpr (PRUNE == 0) {
          psum += p;
          const unsigned int bin = (unsigned int)(__double_as_longlong(p) >> 52) & (HIST_BINS - 1);
          if (bin < LBINS)
            atomicAdd(&s_hist[bin], 1u);
          else
            atomicAdd(&hist[bin], 1u);
        }
        ++pos;
      }
    }
  }
  const double s = block_sum(psum, L.red);
  if (threadIdx.x == 0 && PRUNE != 2) tile_psum[otile] = s;
  __syncthreads();
  if constexpr (PRUNE != 2)
    for (int b = threadIdx.x; b < LBINS; b += NT)
      if (s_hist[b]) atomicAdd(&hist[b], s_hist[b]);
  if constexpr (SPEC) {
    // the staged |a|^2 read back in the select's slot-to-thread order
    __shared__ int s_wcnt[2 * SEL_ITEMS * WAVES];
    const unsigned int b1 = *S.pred;
    double vp[SEL_ITEMS];
    if constexpr (ONE_PASS) {
      __syncthreads();
#pragma unroll
      for (int it = 0; it < SEL_ITEMS; ++it) {
        const int jj = it * NT + threadIdx.x;
        vp[it] = jj < Q.total ? s_p[jj] : 0.0;
      }
      __syncthreads();
    } else {
      lift_to_slots<QB>(Q, pv, s_p, vp);
    }
    if constexpr (FUSE) {
      uint64_t vk[SEL_ITEMS];
      lift_to_slots<QB>(Q, Q.key, reinterpret_cast<uint64_t*>(L.a), vk);
      ex1_tile_core<false, true, CAP>(S.out, otile, rb, Q.total, vp, b1, nullptr, s_wcnt, L.red, s_p, vk);
    } else {
      ex1_tile_core<false, false, CAP>(S.out, otile, rb, Q.total, vp, b1, nullptr, s_wcnt, L.red, s_p);
    }
  }
}

// (the speculative variants are held to the plain kernel's 5 waves per SIMD)
template <int QB, int MODE, bool SPEC, bool FUSE, int PRUNE = 0>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SPEC ? 5 : 1))) void k_lift_gate(
    const uint64_t* __restrict__ w, const double2* __restrict__ a, int64_t n, int j,
    const LiftBound* __restrict__ bounds, GateMat G, uint64_t* __restrict__ xo, double2* __restrict__ ao,
    unsigned int* hist, double* tile_psum, int* seg_cnt, const int64_t* __restrict__ n_dev, LiftSpec S) {
  // tiles past a counted input's end have empty bounds: they write a zero
  // count and sum
  n = dev_count(n_dev, n);
  const int64_t tile = blockIdx.x;
  lift_gate_tile<QB, MODE, SPEC, FUSE, PRUNE>(w, a, n, j, bounds[tile], bounds[tile + 1], G, tile, tile, xo, ao,
                                              hist, tile_psum, seg_cnt, S);
}

// The relift of a pruned lift whose select needs outputs it dropped (flag set
// by k_find1; otherwise every block returns at once): persistent blocks
// recompute every tile and store all its outputs into the select's input
// tiles, with the speculative digit-1 pass against the select's own bucket,
// so the select goes on as after a held prediction.
template <int QB, int MODE>
__global__ __launch_bounds__(NT) void k_lift_relift(LiftRecompute R, SelProblem P, SelTiles T) {
  if (!P.sel->relift) return;
  const int64_t n = dev_count(R.n_dev, R.n);
  LiftSpec S;
  S.pred = &P.sel->b1;
  S.out = Ex1Out{T.c1p, T.c1i, T.cnt1, T.tile_dsum, T.tile_dcnt, nullptr};
  uint64_t* xo = const_cast<uint64_t*>(P.x);
  double2* ao = const_cast<double2*>(P.a);
  int* seg = const_cast<int*>(P.seg_cnt);
  for (int64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x) {
    lift_gate_tile<QB, MODE, true, false, 2>(R.x, R.a, n, R.j, R.bounds[tile], R.bounds[tile + 1], R.G, tile, tile,
                                             xo, ao, nullptr, nullptr, seg, S);
    __syncthreads();
  }
}

// ------------------------------------------------- fused gate select ------
// A gate whose output goes straight into a top-k select need not store it.
// The lift (k_lift_gate with FUSE) keeps what the select reads before its
// compaction -- the digit-1 histogram, per-tile |a|^2 sums and output counts,
// and from the speculative pass the candidates with their keys -- and the
// select recomputes the lift's tiles where it needs the outputs themselves:
// k_lift_ex1 (the digit-1 pass, when the speculation missed) and
// k_lift_compact (the compaction).  A recomputed tile is the lift's tile bit
// for bit (same code on the same input), so the fused select writes what
// the stored path writes, for about half the memory traffic.

// The select's digit-1 pass when the speculative bucket was not the select's:
// block vb recomputes tiles [vb * ext, +ext) and runs ex1_tile_core on each
// (keys into c1k), its digit-2 counts into row vb (as k_extract1's rows).
template <int QB, int MODE>
__global__ __launch_bounds__(NT) void k_lift_ex1(LiftRecompute R, SelProblem P, SelTiles T, int ext) {
  constexpr int M = 1 << QB;
  constexpr int ITEMS = (SEG_ENTRIES >> QB) / NT;
  const SelState* sel = P.sel;
  if (sel->keepall || sel->spec_hit) return;
  __shared__ LiftTileLds<QB> L;
  __shared__ unsigned int s_hist[HIST_BINS];
  __shared__ int s_wcnt[2 * SEL_ITEMS * WAVES];
  const int64_t n = dev_count(R.n_dev, R.n);
  const unsigned int b1 = sel->b1;
  const Ex1Out O{T.c1p, T.c1i, T.cnt1, T.tile_dsum, T.tile_dcnt, T.c1k};
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  const int64_t vb = blockIdx.x;
  for (int sub = 0; sub < ext; ++sub) {
    const int64_t gt = vb * ext + sub;
    if (gt >= P.ntiles) break;  // uniform
    LiftTileOut<QB> Q;
    lift_tile_compute<QB, MODE>(R.x, R.a, n, R.j, R.bounds[gt], R.bounds[gt + 1], R.G, gt, L, Q);
    double pv[ITEMS][M];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it)
#pragma unroll
      for (int o = 0; o < M; ++o)
        pv[it][o] = ((Q.mask[it] >> o) & 1u) ? cabs2(make_double2(Q.re[it][o], Q.im[it][o])) : 0.0;
    double vp[SEL_ITEMS];
    uint64_t vk[SEL_ITEMS];
    lift_to_slots<QB>(Q, pv, reinterpret_cast<double*>(L.a), vp);
    lift_to_slots<QB>(Q, Q.key, reinterpret_cast<uint64_t*>(L.a), vk);
    ex1_tile_core<true, true, 2 * (SEG_ENTRIES >> QB)>(O, gt, gt * SEG_ENTRIES, Q.total, vp, b1, s_hist, s_wcnt,
                                                        L.red, reinterpret_cast<double*>(L.a), vk);
  }
  __syncthreads();
  unsigned short* row = T.hist2_rows + vb * HIST_BINS;
  for (int b = threadIdx.x; b < HIST_BINS; b += NT) row[b] = (unsigned short)s_hist[b];
}

// The compaction: block gt recomputes tile gt and writes its kept outputs
// (kept_pred against the resolved threshold, as k_compact) from the tile's
// kept offset on, in slot order, divided by the norm; the tile's sums of the
// written |a|^2 and |a|^4 follow k_compact's order (slot it * NT + thread,
// items in order; a slot not kept adds +0.0 to a sum of non-negative terms,
// which leaves it unchanged).
template <int QB, int MODE>
__global__ __launch_bounds__(NT) void k_lift_compact(LiftRecompute R, SelProblem P, SelTiles T) {
  constexpr int M = 1 << QB;
  constexpr int ITEMS = (SEG_ENTRIES >> QB) / NT;
  __shared__ LiftTileLds<QB> L;
  const int64_t gt = blockIdx.x;
  const int64_t n = dev_count(R.n_dev, R.n);
  LiftTileOut<QB> Q;
  lift_tile_compute<QB, MODE>(R.x, R.a, n, R.j, R.bounds[gt], R.bounds[gt + 1], R.G, gt, L, Q);
  const unsigned long long pstar = P.sel->pstar, keystar = P.sel->keystar;
  // numpy divides complex128 by a real norm as a product with 1/norm
  const double inv = 1.0 / P.st->norm;
  unsigned int km[ITEMS];
  int kc[ITEMS], ko[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    unsigned int mk = 0;
#pragma unroll
    for (int o = 0; o < M; ++o) {
      if (((Q.mask[it] >> o) & 1u) &&
          kept_pred(pbits_of(make_double2(Q.re[it][o], Q.im[it][o])), Q.key[it][o], pstar, keystar, P))
        mk |= 1u << o;
    }
    km[it] = mk;
    kc[it] = __popc(mk);
  }
  int ktot;
  block_excl_scan_items<ITEMS>(kc, ko, ktot, L.scan);
  const int64_t rb = T.koff[gt] + (P.chunked ? T.ch_off[(gt - P.tile0) / SEL_CHUNK] : 0);
  // the written |a|^2 by kept rank (k_compact's sums, from the same entries)
  double p2[ITEMS * M];
  int kr[ITEMS * M];
  bool kok[ITEMS * M];
  bool fits = true;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    int r = ko[it];
#pragma unroll
    for (int o = 0; o < M; ++o) {
      const int e = it * M + o;
      p2[e] = 0.0;
      kr[e] = r;
      kok[e] = (km[it] >> o) & 1u;
      if (kok[e]) {
        const double2 v = make_double2(Q.re[it][o] * inv, Q.im[it][o] * inv);
        const int64_t pos = rb + r;
        if (pos < P.cap_out) {
          P.xo[pos] = Q.key[it][o];
          P.ao[pos] = v;
        } else {
          fits = false;
        }
        p2[e] = cabs2(v);
        ++r;
      }
    }
  }
  if (!fits) set_err(&P.st->err, ERR_OVERFLOW);
  double s4;
  const double s2 = rank_ordered_sum<ITEMS * M, 2 * (SEG_ENTRIES >> QB), true>(p2, kr, kok, ktot,
                                                                            reinterpret_cast<double*>(L.a), &s4);
  const double t2 = block_sum(s2, L.red);
  const double t4 = block_sum(s4, L.red);
  if (threadIdx.x == 0) {
    T.tile_s2[gt] = t2;
    T.tile_s4[gt] = t4;
  }
}

// ------------------------------------------------- two-qubit RDM ------
// RDM of the qubit pair stored at bits (1, 0) after the lift of bit j to bit 1
// (the second qubit already at bit 0; src/simulation/bass_simulator.py:202-234):
// rho[p][q] = sum over groups (entries equal outside the pair) of
// a_p conj(a_q), member index p = (bit 1 << 1) | bit 0, absent members 0.
// Per tile the 16 real values of the upper triangle (4 diagonal, 6 complex
// off-diagonal (p < q) in row order) are summed in a fixed order; a finishing
// kernel reduces the tiles, also in a fixed order.
constexpr int RDM2_VALS = 16;

__device__ __forceinline__ void lift_rdm2_tile(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                               int64_t n, int j, const LiftBound& B0, const LiftBound& B1,
                                               int64_t tile, double* __restrict__ part) {
  constexpr int QB = 2, M = 4, T = 1;
  constexpr int TM = SEG_ENTRIES / M;
  constexpr int ITEMS = TM / NT;
  __shared__ unsigned short s_rank[TM];
  __shared__ uint64_t s_t[TM];
  __shared__ double2 s_a[TM];
  __shared__ int s_scan[ITEMS * WAVES];
  __shared__ double s_red[RDM2_VALS * WAVES];
  __shared__ LiftHalo s_halo;
  const int cnt = lift_stage<TM>(w, a, n, j, T, B0, B1, tile, s_rank, s_scan, s_t, s_a);
  lift_halo_setup<QB>(w, a, j, B0, B1, s_t, cnt, s_halo);
  __syncthreads();
  double acc[RDM2_VALS];
#pragma unroll
  for (int v = 0; v < RDM2_VALS; ++v) acc[v] = 0.0;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int m = it * NT + threadIdx.x;
    double avr[M], avi[M];
    unsigned int pres;
    if (m < cnt && lift_group_gather<QB>(m, cnt, s_t, s_a, s_halo, avr, avi, pres)) {
#pragma unroll
      for (int p = 0; p < M; ++p) acc[p] += cabs2(make_double2(avr[p], avi[p]));
      int v = M;
#pragma unroll
      for (int p = 0; p < M; ++p) {
#pragma unroll
        for (int q = p + 1; q < M; ++q) {
          const double2 z = cmul(make_double2(avr[p], avi[p]), make_double2(avr[q], -avi[q]));
          acc[v] += z.x;
          acc[v + 1] += z.y;
          v += 2;
        }
      }
    }
  }
  // block sums of the 16 values: shuffle tree per wave, then waves in order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < RDM2_VALS; ++v) {
    const double t = wave_sum(acc[v]);
    if (lane == 0) s_red[v * WAVES + wv] = t;
  }
  __syncthreads();
  if (threadIdx.x < RDM2_VALS) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) t += s_red[threadIdx.x * WAVES + i];
    part[tile * RDM2_VALS + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(NT) void k_lift_rdm2(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                                   int64_t n, int j, const LiftBound* __restrict__ bounds,
                                                   double* __restrict__ part) {
  const int64_t tile = blockIdx.x;
  lift_rdm2_tile(w, a, n, j, bounds[tile], bounds[tile + 1], tile, part);
}

// Value v (block v) of the tiles' partials, summed in a fixed order.
__global__ __launch_bounds__(NTB) void k_rdm2_fin(const double* __restrict__ part, int64_t ntiles, double* out) {
  __shared__ double s_red[WAVESB];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < ntiles; t += NTB) s += part[t * RDM2_VALS + v];
  const double tot = block_sum<NTB>(s, s_red);
  if (threadIdx.x == 0) out[v] = tot;
}

size_t rdm2_scratch(int64_t n) {
  const int64_t tiles = (n + SEG_ENTRIES / 4 - 1) / (SEG_ENTRIES / 4);
  return (((size_t)(tiles + 1 + lift_coarse_count(tiles + 1)) * sizeof(LiftBound) + 255) & ~(size_t)255) +
         (size_t)tiles * RDM2_VALS * 8 + 256;
}

// out: RDM2_VALS doubles (device).
int launch_rdm2(const uint64_t* x, const double2* a, int64_t n, int j, void* scratch, double* out, hipStream_t s) {
  constexpr int TM = SEG_ENTRIES / 4;
  const int64_t tiles = (n + TM - 1) / TM, nbound = tiles + 1;
  char* base = (char*)scratch;
  LiftBound* bounds = (LiftBound*)base;
  base += ((size_t)(nbound + lift_coarse_count(nbound)) * sizeof(LiftBound) + 255) & ~(size_t)255;
  double* part = (double*)base;
  lift_partition(x, n, j, 1, 1, TM, nbound, bounds, bounds + nbound, s);
  hipLaunchKernelGGL(k_lift_rdm2, dim3((unsigned)tiles), dim3(NT), 0, s, x, a, n, j, bounds, part);
  hipLaunchKernelGGL(k_rdm2_fin, dim3(RDM2_VALS), dim3(NTB), 0, s, part, tiles, out);
  return (int)hipGetLastError();
}

template <int TM>
__device__ __forceinline__ void lift_move_tile(const uint64_t* __restrict__ w, const double2* __restrict__ a, int64_t n, int j,
                               const LiftBound& B0, const LiftBound& B1, int64_t tile, uint64_t* __restrict__ xo,
                               double2* __restrict__ ao) {
  __shared__ unsigned short s_rank[TM];
  __shared__ int s_scan[(TM / NT) * WAVES];
  __shared__ uint64_t s_t[TM];
  __shared__ double2 s_a[TM];
  const int cnt = lift_stage<TM>(w, a, n, j, 0, B0, B1, tile, s_rank, s_scan, s_t, s_a);
  const int64_t d = tile * TM;
  for (int m = threadIdx.x; m < cnt; m += NT) {
    xo[d + m] = s_t[m];
    ao[d + m] = s_a[m];
  }
}

template <int TM>
__global__ __launch_bounds__(NT) void k_lift_move(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                                   int64_t n, int j, const LiftBound* __restrict__ bounds,
                                                   uint64_t* __restrict__ xo, double2* __restrict__ ao,
                                                   const int64_t* __restrict__ n_dev) {
  n = dev_count(n_dev, n);
  lift_move_tile<TM>(w, a, n, j, bounds[blockIdx.x], bounds[blockIdx.x + 1], blockIdx.x, xo, ao);
}

// ------------------------------------------------------------- batches ----
// A batch of independent states (one job each) in one launch sequence: job c
// owns output tiles [tile0_c, tile0_c + ntiles_c) and boundaries [b0_c, b0_c +
// ntiles_c] (b0_c = tile0_c + c); blocks find their job by binary search.
// k_bprep fixes that geometry on the device (BatchHdr): every kernel after it
// walks the totals of the header with a persistent grid of any size.
// Job of a block-uniform tile index.  The result is made explicitly
// wave-uniform so the job's fields (pointers, gate matrix) are read with
// scalar loads instead of occupying vector registers.
__device__ __forceinline__ int job_of_tile(const LiftJob* jobs, int njobs, int64_t gt) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile0 <= gt)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Prep of a batch lift / step.  Block 0: per job the input count (its count
// cell, else the staged n), tiles = ceil(n / tm), tile0 by an exclusive scan
// in job order, the select problem's tile geometry and candidate arrays, the
// header totals; it also clears the jobs' stats and select states.  Other
// blocks clear the jobs' digit-1 / digit-2 histograms.  A total above the
// staged tile capacity runs nothing (nrun = 0) and reports ERR_OVERFLOW in
// every job's stats.
__global__ __launch_bounds__(NTB) void k_bprep(LiftJob* __restrict__ jobs, SelProblem* __restrict__ probs,
                                               BatchHdr* hdr, SelBase sb, BatchZero z) {
  const int njobs = hdr->njobs;
  if (blockIdx.x > 0) {
    if (!probs) return;
    const int64_t words = (int64_t)njobs * HIST_BINS;  // per histogram array
    const int64_t stride = (int64_t)(gridDim.x - 1) * NTB;
    for (int64_t i = (int64_t)(blockIdx.x - 1) * NTB + threadIdx.x; i < words; i += stride) {
      z.hist1[i] = 0u;
      z.hist2[i] = 0u;
    }
    return;
  }
  __shared__ int s_scan[WAVESB];
  __shared__ int s_max;
  if (probs) {
    const int64_t stw = (int64_t)njobs * (int64_t)(sizeof(OpStats) / 8);
    const int64_t slw = (int64_t)njobs * (int64_t)(sizeof(SelState) / 4);
    for (int64_t i = threadIdx.x; i < stw; i += NTB) ((unsigned long long*)z.st)[i] = 0ull;
    for (int64_t i = threadIdx.x; i < slw; i += NTB) ((unsigned int*)z.sel)[i] = 0u;
  }
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  int64_t carry = 0;
  for (int c0 = 0; c0 < njobs; c0 += NTB) {
    const int c = c0 + (int)threadIdx.x;
    int64_t n = 0;
    int tiles = 0;
    if (c < njobs) {
      LiftJob& J = jobs[c];
      n = J.cnt_in ? *J.cnt_in : J.n;
      if (n < 0) n = 0;
      J.n = n;
      const int tm = lift_job_tm(J.qb);
      tiles = (int)((n + tm - 1) / tm);
      if (!probs && J.cnt_out) *J.cnt_out = n;  // a move keeps the count
      atomicMax(&s_max, tiles);
    }
    int tot;
    const int pre = block_excl_scan<NTB>(tiles, tot, s_scan);
    if (c < njobs) {
      const int64_t t0 = carry + pre;
      jobs[c].tile0 = t0;
      if (probs) {
        SelProblem& P = probs[c];
        P.tile0 = t0;
        P.ntiles = tiles;
        P.row0 = t0;
        P.nrows = tiles;
        P.npsum = tiles;
        P.c2p = sb.c2p + t0 * SEG_ENTRIES;
        P.c2i = sb.c2i + t0 * SEG_ENTRIES;
        P.tk = sb.tk + t0 * SEG_ENTRIES;
      }
    }
    carry += tot;
  }
  __syncthreads();
  const bool over = carry > hdr->cap_tiles;
  if (over && probs)
    for (int c = threadIdx.x; c < njobs; c += NTB) z.st[c].err = ERR_OVERFLOW;
  if (threadIdx.x == 0) {
    hdr->nrun = over ? 0 : njobs;
    hdr->ntiles = over ? 0 : carry;
    hdr->nbound = over ? 0 : carry + njobs;
    hdr->max_rows = over ? 0 : s_max;
  }
}

__global__ __launch_bounds__(NT) void k_blift_part(const LiftJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr,
                                                    LiftBound* bounds) {
  const int lane = threadIdx.x & 63;
  const int njobs = hdr->nrun;
  const int64_t nbound = hdr->nbound;
  // one wave per boundary (uniform per wave)
  for (int64_t gb = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6); gb < nbound;
       gb += (int64_t)gridDim.x * WAVES) {
    // job: the last one with b0 = tile0 + index <= gb
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].tile0 + mid <= gb)
        lo = mid;
      else
        hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const LiftJob& J = jobs[lo];
    const int64_t u = gb - (J.tile0 + lo);
    const LiftBound r = lift_boundary(J.x, J.n, J.j, lift_job_t(J.qb), u * (int64_t)lift_job_tm(J.qb));
    if (lane == 0) bounds[gb] = r;
  }
}

template <int QB, int MODE>
__global__ __launch_bounds__(NT) void k_blift_gate(const LiftJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr,
                                                    const LiftBound* __restrict__ bounds, uint64_t* __restrict__ xo,
                                                    double2* __restrict__ ao, unsigned int* hist, double* tile_psum,
                                                    int* seg_cnt) {
  const int njobs = hdr->nrun;
  const int64_t ntiles = hdr->ntiles;
  for (int64_t gt = blockIdx.x; gt < ntiles; gt += gridDim.x) {
    const int c = job_of_tile(jobs, njobs, gt);
    const LiftJob& J = jobs[c];
    if (J.qb != QB) continue;  // another kind's tile (uniform per block)
    const int64_t u = gt - J.tile0;
    const int64_t b = J.tile0 + c + u;
    lift_gate_tile<QB, MODE>(J.x, J.a, J.n, J.j, bounds[b], bounds[b + 1], J.g, u, gt, xo, ao,
                             hist + (int64_t)c * HIST_BINS, tile_psum, seg_cnt);
    __syncthreads();  // the next tile reuses the LDS
  }
}

__global__ __launch_bounds__(NT) void k_blift_move(const LiftJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr,
                                                    const LiftBound* __restrict__ bounds) {
  const int njobs = hdr->nrun;
  const int64_t ntiles = hdr->ntiles;
  for (int64_t gt = blockIdx.x; gt < ntiles; gt += gridDim.x) {
    const int c = job_of_tile(jobs, njobs, gt);
    const LiftJob& J = jobs[c];
    const int64_t u = gt - J.tile0;
    const int64_t b = J.tile0 + c + u;
    lift_move_tile<1024>(J.x, J.a, J.n, J.j, bounds[b], bounds[b + 1], u, J.xo, J.ao);
    __syncthreads();
  }
}

// ------------------------------------------------ single-qubit RDMs ------
// RDMs of every stored bit of a key-ordered state (src/simulation/bass_simulator.py:441-470):
//   rho_p[0][0] = sum |a_x|^2 over bit p of x clear, rho_p[1][1] over bit p set,
//   rho_p[0][1] = sum a_x * conj(a_{x | 2^p}) over pairs present.
// Pairs: the lift of bit p (t = 0) makes x and x | 2^p adjacent in merged
// order, so every pair lies in the tile holding x, or its partner is the first
// entry merged after the tile (B cursor of the next boundary).  A tile loads
// only keys (and their indices), finds each partner with one LDS search and
// gathers the two amplitudes of each pair.
constexpr int RDM_TM = 1024;

// rho01 partial of stored bit p over merged tile u (bounds B0, B1).
__device__ __forceinline__ double2 rdm_pairs_tile(const uint64_t* __restrict__ w, const double2* __restrict__ a, int64_t n, int p,
                                  int64_t u, const LiftBound& B0, const LiftBound& B1) {
  constexpr int ITEMS = RDM_TM / NT;
  __shared__ uint64_t s_w[RDM_TM];
  __shared__ int s_g[RDM_TM];
  __shared__ double s_red[WAVES];
  __shared__ uint64_t s_hw;
  __shared__ int s_hg;
  const int64_t d = u * RDM_TM;
  const int cnt = (int)((n - d) < RDM_TM ? (n - d) : RDM_TM);
  int64_t p1lo, p1n, p2lo, p2n, p3lo;
  if (B0.s0 == B1.s0) {
    p1lo = B0.a;
    p1n = B1.a - B0.a;
    p2lo = B0.b;
    p2n = B1.b - B0.b;
    p3lo = 0;
  } else {
    p1lo = B0.a;
    p1n = B0.sm - B0.a;
    p2lo = B0.b;
    p2n = B1.a - B0.b;
    p3lo = B1.sm;
  }
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int idx = it * NT + threadIdx.x;
    if (idx < cnt) {
      const int64_t g = idx < p1n ? p1lo + idx : (idx < p1n + p2n ? p2lo + (idx - p1n) : p3lo + (idx - p1n - p2n));
      s_w[idx] = w[g];
      s_g[idx] = (int)g;
    }
  }
  if (threadIdx.x == 0) {
    const bool h = B1.b < B1.s1;
    s_hw = h ? w[B1.b] : ~0ull;  // ~0 never equals x | 2^p for keys < 2^63
    s_hg = h ? B1.b : -1;
  }
  __syncthreads();
  const uint64_t bp = 1ull << p;
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int idx = it * NT + threadIdx.x;
    if (idx < cnt) {
      const uint64_t x = s_w[idx];
      if (!(x & bp)) {
        const uint64_t y = x | bp;
        const int q = lds_lower_bound(s_w, cnt, y);
        const int gb = (q < cnt && s_w[q] == y) ? s_g[q] : (s_hw == y ? s_hg : -1);
        if (gb >= 0) {
          const double2 v = cmul(a[s_g[idx]], cconj(a[gb]));
          re += v.x;
          im += v.y;
        }
      }
    }
  }
  const double tr = block_sum(re, s_red);
  const double ti = block_sum(im, s_red);
  return make_double2(tr, ti);
}

__global__ __launch_bounds__(NT) void k_rdm_pairs(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                                   int64_t n, int64_t nbound, const LiftBound* __restrict__ bounds,
                                                   double2* __restrict__ part) {
  const int p = blockIdx.y;
  const int64_t u = blockIdx.x;
  const double2 v = rdm_pairs_tile(w, a, n, p, u, bounds[(int64_t)p * nbound + u], bounds[(int64_t)p * nbound + u + 1]);
  if (threadIdx.x == 0) part[(int64_t)p * gridDim.x + u] = v;
}

// Diagonal RDM elements of every bit in one pass: per tile of SEG_ENTRIES
// entries, lane b of each wave owns bit b and walks the wave's entries in
// order (LDS broadcast reads).
__device__ __forceinline__ void rdm_diag_tile(const uint64_t* __restrict__ w, const double2* __restrict__ a, int64_t n, int nbits,
                              int64_t tile, double2* __restrict__ out) {
  constexpr int TILE = SEG_ENTRIES;
  constexpr int PER_WAVE = TILE / WAVES;
  __shared__ uint64_t s_w[TILE];
  __shared__ double s_p[TILE];
  __shared__ double2 s_acc[WAVES][64];
  const int64_t d = tile * TILE;
  const int cnt = (int)((n - d) < TILE ? (n - d) : TILE);
  for (int j = threadIdx.x; j < cnt; j += NT) {
    s_w[j] = w[d + j];
    s_p[j] = cabs2(a[d + j]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double z = 0.0, o = 0.0;
  const int e0 = wv * PER_WAVE;
  const int e1 = e0 + PER_WAVE < cnt ? e0 + PER_WAVE : cnt;
  for (int e = e0; e < e1; ++e) {
    const double pe = s_p[e];
    if ((s_w[e] >> lane) & 1ull)
      o += pe;
    else
      z += pe;
  }
  s_acc[wv][lane] = make_double2(z, o);
  __syncthreads();
  if (threadIdx.x < nbits) {
    double2 t = s_acc[0][threadIdx.x];
    for (int k = 1; k < WAVES; ++k) {
      t.x += s_acc[k][threadIdx.x].x;
      t.y += s_acc[k][threadIdx.x].y;
    }
    out[threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(NT) void k_rdm_diag(const uint64_t* __restrict__ w, const double2* __restrict__ a,
                                                  int64_t n, int nbits, double2* __restrict__ dpart) {
  rdm_diag_tile(w, a, n, nbits, blockIdx.x, dpart + (int64_t)blockIdx.x * 64);
}

// rho[p] = (rho00, rho11, re rho01, im rho01): one wave per bit, partials
// summed in a fixed order.
__global__ __launch_bounds__(NT) void k_rdm_ordered_fin(const double2* __restrict__ part, int64_t ptiles,
                                                         const double2* __restrict__ dpart, int64_t dtiles, int nbits,
                                                         double* rho) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (p >= nbits) return;
  double z = 0.0, o = 0.0, re = 0.0, im = 0.0;
  for (int64_t u = lane; u < dtiles; u += 64) {
    const double2 v = dpart[u * 64 + p];
    z += v.x;
    o += v.y;
  }
  for (int64_t u = lane; u < ptiles; u += 64) {
    const double2 v = part[(int64_t)p * ptiles + u];
    re += v.x;
    im += v.y;
  }
  z = wave_sum(z);
  o = wave_sum(o);
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) {
    rho[4 * p + 0] = z;
    rho[4 * p + 1] = o;
    rho[4 * p + 2] = re;
    rho[4 * p + 3] = im;
  }
}

// ---- batched RDMs: job c covers nbits_c bits; flattened index spaces
// boundaries [bound_off_c, + nbits_c (ptiles_c + 1)), pair tiles
// [ptile_off_c, + nbits_c ptiles_c), diagonal tiles [dtile_off_c, + dtiles_c).
template <int FIELD>
__device__ __forceinline__ int rdm_job_of(const RdmJob* jobs, int njobs, int64_t g) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t off = FIELD == 0 ? jobs[mid].bound_off : (FIELD == 1 ? jobs[mid].ptile_off : jobs[mid].dtile_off);
    if (off <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(NT) void k_brdm_part(const RdmJob* __restrict__ jobs, int njobs, int64_t nbound,
                                                   LiftBound* bounds) {
  const int lane = threadIdx.x & 63;
  const int64_t gb = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (gb >= nbound) return;  // uniform per wave
  const RdmJob& J = jobs[rdm_job_of<0>(jobs, njobs, gb)];
  const int64_t r = gb - J.bound_off;
  const int p = (int)(r / (J.ptiles + 1));
  const int64_t u = r % (J.ptiles + 1);
  const LiftBound b = lift_boundary(J.x, J.n, p, 0, u * RDM_TM);
  if (lane == 0) bounds[gb] = b;
}

__global__ __launch_bounds__(NT) void k_brdm_pairs(const RdmJob* __restrict__ jobs, int njobs,
                                                    const LiftBound* __restrict__ bounds, double2* __restrict__ part) {
  const int64_t gt = blockIdx.x;
  const RdmJob& J = jobs[rdm_job_of<1>(jobs, njobs, gt)];
  const int64_t r = gt - J.ptile_off;
  const int p = (int)(r / J.ptiles);
  const int64_t u = r % J.ptiles;
  const int64_t b = J.bound_off + (int64_t)p * (J.ptiles + 1) + u;
  const double2 v = rdm_pairs_tile(J.x, J.a, J.n, p, u, bounds[b], bounds[b + 1]);
  if (threadIdx.x == 0) part[gt] = v;
}

__global__ __launch_bounds__(NT) void k_brdm_diag(const RdmJob* __restrict__ jobs, int njobs,
                                                   double2* __restrict__ dpart) {
  const int64_t gt = blockIdx.x;
  const RdmJob& J = jobs[rdm_job_of<2>(jobs, njobs, gt)];
  rdm_diag_tile(J.x, J.a, J.n, J.nbits, gt - J.dtile_off, dpart + gt * 64);
}

// rho[c][p] for p < nbits_c: one wave per (job, bit); blockIdx.y = job.
__global__ __launch_bounds__(NT) void k_brdm_fin(const RdmJob* __restrict__ jobs, const double2* __restrict__ part,
                                                  const double2* __restrict__ dpart, double* rho) {
  const int lane = threadIdx.x & 63;
  const RdmJob& J = jobs[blockIdx.y];
  const int p = blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (p >= J.nbits) return;
  double z = 0.0, o = 0.0, re = 0.0, im = 0.0;
  for (int64_t u = lane; u < J.dtiles; u += 64) {
    const double2 v = dpart[(J.dtile_off + u) * 64 + p];
    z += v.x;
    o += v.y;
  }
  for (int64_t u = lane; u < J.ptiles; u += 64) {
    const double2 v = part[J.ptile_off + (int64_t)p * J.ptiles + u];
    re += v.x;
    im += v.y;
  }
  z = wave_sum(z);
  o = wave_sum(o);
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) {
    double* out = rho + ((int64_t)blockIdx.y * 64 + p) * 4;
    out[0] = z;
    out[1] = o;
    out[2] = re;
    out[3] = im;
  }
}

int launch_rdm_batch(const RdmJob* d_jobs, int njobs, int64_t nbound, int64_t nptiles, int64_t ndtiles, void* scratch,
                     double* rho, hipStream_t s) {
  char* base = (char*)scratch;
  LiftBound* bounds = (LiftBound*)base;
  base += ((size_t)nbound * sizeof(LiftBound) + 255) & ~(size_t)255;
  double2* part = (double2*)base;
  base += ((size_t)nptiles * 16 + 255) & ~(size_t)255;
  double2* dpart = (double2*)base;
  if (nbound > 0)
    hipLaunchKernelGGL(k_brdm_part, dim3((unsigned)((nbound + WAVES - 1) / WAVES)), dim3(NT), 0, s, d_jobs, njobs,
                       nbound, bounds);
  if (nptiles > 0)
    hipLaunchKernelGGL(k_brdm_pairs, dim3((unsigned)nptiles), dim3(NT), 0, s, d_jobs, njobs, bounds, part);
  if (ndtiles > 0) hipLaunchKernelGGL(k_brdm_diag, dim3((unsigned)ndtiles), dim3(NT), 0, s, d_jobs, njobs, dpart);
  hipLaunchKernelGGL(k_brdm_fin, dim3(64 / WAVES, (unsigned)njobs), dim3(NT), 0, s, d_jobs, part, dpart, rho);
  return (int)hipGetLastError();
}

size_t rdm_batch_scratch(int64_t nbound, int64_t nptiles, int64_t ndtiles) {
  return (size_t)nbound * sizeof(LiftBound) + (size_t)nptiles * 16 + (size_t)ndtiles * 64 * 16 + 3 * 256;
}

int rdm_tile_entries() { return RDM_TM; }

size_t rdm_ordered_scratch(int64_t n, int nbits) {
  const int64_t ptiles = (n + RDM_TM - 1) / RDM_TM, dtiles = (n + SEG_ENTRIES - 1) / SEG_ENTRIES;
  return (size_t)nbits * (ptiles + 1 + lift_coarse_count(ptiles + 1)) * sizeof(LiftBound) +
         (size_t)nbits * ptiles * 16 + (size_t)dtiles * 64 * 16 + 3 * 256;
}

int launch_rdm_ordered(const uint64_t* x, const double2* a, int64_t n, int nbits, void* scratch, double* rho,
                       hipStream_t s) {
  const int64_t ptiles = (n + RDM_TM - 1) / RDM_TM, dtiles = (n + SEG_ENTRIES - 1) / SEG_ENTRIES;
  const int64_t nbound = ptiles + 1;
  char* base = (char*)scratch;
  LiftBound* bounds = (LiftBound*)base;
  LiftBound* coarse = bounds + (size_t)nbits * nbound;
  base += ((size_t)nbits * (nbound + lift_coarse_count(nbound)) * sizeof(LiftBound) + 255) & ~(size_t)255;
  double2* part = (double2*)base;
  base += ((size_t)nbits * ptiles * 16 + 255) & ~(size_t)255;
  double2* dpart = (double2*)base;
  lift_partition(x, n, 0, nbits, 0, RDM_TM, nbound, bounds, coarse, s);
  hipLaunchKernelGGL(k_rdm_pairs, dim3((unsigned)ptiles, (unsigned)nbits), dim3(NT), 0, s, x, a, n, nbound, bounds,
                     part);
  hipLaunchKernelGGL(k_rdm_diag, dim3((unsigned)dtiles), dim3(NT), 0, s, x, a, n, nbits, dpart);
  hipLaunchKernelGGL(k_rdm_ordered_fin, dim3((unsigned)((nbits + WAVES - 1) / WAVES)), dim3(NT), 0, s, part, ptiles,
                     dpart, dtiles, nbits, rho);
  return (int)hipGetLastError();
}

// ----------------------------------------------------------- host launch ---

// LiftBound entries a lift of n entries needs (boundaries + coarse boundaries).
int64_t lift_bounds_entries(int64_t nbound) { return nbound + lift_coarse_count(nbound); }

int lift_tile_entries(int qb, int mode) { return mode == MODE_MOVE ? 1024 : SEG_ENTRIES / (1 << qb); }

// Batched lift: geometry in hdr (k_bprep first), persistent grids.
int launch_batch_prep(LiftJob* d_jobs, SelProblem* d_probs, BatchHdr* hdr, const SelBase& sb, const BatchZero& z,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_bprep, dim3(d_probs ? 65 : 1), dim3(NTB), 0, s, d_jobs, d_probs, hdr, sb, z);
  return (int)hipGetLastError();
}

int launch_lift_batch(int qmask, int mode, const LiftJob* d_jobs, const BatchHdr* hdr, int64_t tiles, int jobs,
                      LiftBound* bounds, uint64_t* xo, double2* ao, unsigned int* hist, double* tile_psum,
                      int* seg_cnt, hipStream_t s) {
  // boundaries: one wave each, WAVES per block
  const dim3 gp(persist_grid(reinterpret_cast<const void*>(&k_blift_part), NT, (tiles + jobs + WAVES) / WAVES));
  hipLaunchKernelGGL(k_blift_part, gp, dim3(NT), 0, s, d_jobs, hdr, bounds);
#define BASS_BLG(QB_, MODE_)                                                                                  \
  hipLaunchKernelGGL((k_blift_gate<QB_, MODE_>),                                                               \
                     dim3(persist_grid(reinterpret_cast<const void*>(&k_blift_gate<QB_, MODE_>), NT, tiles)), \
                     dim3(NT), 0, s, d_jobs, hdr, bounds, xo, ao, hist, tile_psum, seg_cnt)
  if (mode == MODE_MOVE) {
    hipLaunchKernelGGL(k_blift_move, dim3(persist_grid(reinterpret_cast<const void*>(&k_blift_move), NT, tiles)),
                       dim3(NT), 0, s, d_jobs, hdr, bounds);
  } else if (mode == MODE_GATE) {
    if (qmask & 1) BASS_BLG(1, MODE_GATE);
    if (qmask & 2) BASS_BLG(2, MODE_GATE);
  } else {
    if (qmask & 1) BASS_BLG(1, MODE_ROT);
    if (qmask & 2) BASS_BLG(2, MODE_ROT);
  }
#undef BASS_BLG
  return (int)hipGetLastError();
}

// One gate kernel instance (QB, MODE) with the SPEC / FUSE variant chosen at run time.
template <int QB, int MODE>
static void launch_gate_kernel(bool spec, bool fuse, bool prune, dim3 grid, hipStream_t s, const uint64_t* x,
                               const double2* a, int64_t n, int j, const LiftBound* bounds, const GateMat& G,
                               uint64_t* xo, double2* ao, unsigned int* hist, double* tile_psum, int* seg_cnt,
                               const int64_t* n_dev, const LiftSpec& S) {
#define BASS_LG(SPEC_, FUSE_)                                                                               \
  hipLaunchKernelGGL((k_lift_gate<QB, MODE, SPEC_, FUSE_>), grid, dim3(NT), 0, s, x, a, n, j, bounds, G, xo, ao, \
                     hist, tile_psum, seg_cnt, n_dev, S)
  if (prune && spec && !fuse)
    hipLaunchKernelGGL((k_lift_gate<QB, MODE, true, false, 1>), grid, dim3(NT), 0, s, x, a, n, j, bounds, G, xo, ao,
                       hist, tile_psum, seg_cnt, n_dev, S);
  else if (spec && fuse)
    BASS_LG(true, true);
  else if (spec)
    BASS_LG(true, false);
  else if (fuse)
    BASS_LG(false, true);
  else
    BASS_LG(false, false);
#undef BASS_LG
}

int launch_lift(int qb, int mode, const uint64_t* x, const double2* a, int64_t n, int j, const GateMat& G,
                uint64_t* xo, double2* ao, LiftBound* bounds, unsigned int* hist, double* tile_psum, int* seg_cnt,
                hipStream_t s, const int64_t* n_dev, const LiftZero& zero, const LiftSpec* spec, bool fuse,
                bool prune) {
  if (n <= 0) return 0;
  const int tm = lift_tile_entries(qb, mode);
  const int t = (mode == MODE_MOVE) ? 0 : qb - 1;
  const int64_t tiles = (n + tm - 1) / tm;
  const int64_t nbound = tiles + 1;
  // the coarse boundaries follow the nbound boundaries (lift_bounds_entries)
  lift_partition(x, n, j, 1, t, tm, nbound, bounds, bounds + nbound, s, n_dev, zero);
  const dim3 grid((unsigned)tiles);
  if (mode == MODE_MOVE) {
    hipLaunchKernelGGL(k_lift_move<1024>, grid, dim3(NT), 0, s, x, a, n, j, bounds, xo, ao, n_dev);
  } else {
    const LiftSpec S = spec ? *spec : LiftSpec();
    const bool sp = spec != nullptr;
    if (qb == 1 && mode == MODE_GATE)
      launch_gate_kernel<1, MODE_GATE>(sp, fuse, prune, grid, s, x, a, n, j, bounds, G, xo, ao, hist, tile_psum, seg_cnt,
                                       n_dev, S);
    else if (qb == 1)
      launch_gate_kernel<1, MODE_ROT>(sp, fuse, prune, grid, s, x, a, n, j, bounds, G, xo, ao, hist, tile_psum, seg_cnt,
                                      n_dev, S);
    else if (mode == MODE_GATE)
      launch_gate_kernel<2, MODE_GATE>(sp, fuse, prune, grid, s, x, a, n, j, bounds, G, xo, ao, hist, tile_psum, seg_cnt,
                                       n_dev, S);
    else
      launch_gate_kernel<2, MODE_ROT>(sp, fuse, prune, grid, s, x, a, n, j, bounds, G, xo, ao, hist, tile_psum, seg_cnt,
                                      n_dev, S);
  }
  return (int)hipGetLastError();
}

int launch_lift_recompute(int which, const LiftRecompute& R, const SelProblem& P, const SelTiles& T, int ext,
                          hipStream_t s) {
  const int64_t tiles = P.ntiles;
  if (tiles <= 0) return 0;
#define BASS_LR(QB_, MODE_)                                                                                     \
  do {                                                                                                         \
    if (which == 0)                                                                                            \
      hipLaunchKernelGGL((k_lift_ex1<QB_, MODE_>), dim3((unsigned)((tiles + ext - 1) / ext)), dim3(NT), 0, s, R, \
                         P, T, ext);                                                                           \
    else if (which == 1)                                                                                       \
      hipLaunchKernelGGL((k_lift_compact<QB_, MODE_>), dim3((unsigned)tiles), dim3(NT), 0, s, R, P, T);        \
    else                                                                                                       \
      hipLaunchKernelGGL((k_lift_relift<QB_, MODE_>),                                                          \
                         dim3(persist_grid(reinterpret_cast<const void*>(&k_lift_relift<QB_, MODE_>), NT, tiles)), \
                         dim3(NT), 0, s, R, P, T);                                                             \
  } while (0)
  if (R.qb == 1 && R.mode == MODE_GATE)
    BASS_LR(1, MODE_GATE);
  else if (R.qb == 1)
    BASS_LR(1, MODE_ROT);
  else if (R.mode == MODE_GATE)
    BASS_LR(2, MODE_GATE);
  else
    BASS_LR(2, MODE_ROT);
#undef BASS_LR
  return (int)hipGetLastError();
}

}  // namespace bass
