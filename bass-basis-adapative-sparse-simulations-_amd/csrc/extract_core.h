// The digit-1 pass of the top-k select over one tile of SEG_ENTRIES slots:
// count and sum of the entries above the digit-1 bucket, the candidates
// inside it, and (optionally) their digit-2 counts.  Shared by the select's
// k_extract1 (reading the tile from memory) and by a speculative gate lift
// (k_lift_gate with a predicted bucket, reading the tile it has just produced
// from LDS), so both write the same values in the same order.
#pragma once
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int SEL_ITEMS = SEG_ENTRIES / NT;  // 8
constexpr int SEL_TILE = SEG_ENTRIES;

__device__ __forceinline__ unsigned long long pbits_of(double2 v) {
  return (unsigned long long)__double_as_longlong(cabs2(v));
}

// Logical key of a stored word (ties are ordered by it).
__device__ __forceinline__ uint64_t logical_key(uint64_t w, const KeyMap& km) {
  if (km.nbits == 0) return w;
  uint64_t r = 0;
  for (int p = 0; p < km.nbits; ++p) r |= ((w >> p) & 1ull) << km.q[p];
  return r;
}

// The same through byte tables (tab: KTAB_WORDS words, global or LDS).
__device__ __forceinline__ uint64_t logical_key_tab(uint64_t w, const uint64_t* tab) {
  uint64_t r = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b) r |= tab[b * 256 + (int)((w >> (8 * b)) & 255u)];
  return r;
}

// Logical key of stored word w of problem P.
__device__ __forceinline__ uint64_t lkey(const SelProblem& P, uint64_t w) {
  if (P.km.nbits == 0) return w;
  if (P.ktab) return logical_key_tab(w, P.ktab);
  return logical_key(w, P.km);
}

__device__ __forceinline__ bool kept_pred(unsigned long long bits, uint64_t w, unsigned long long pstar,
                                          unsigned long long keystar, const SelProblem& P) {
  return bits > pstar || (bits == pstar && lkey(P, w) <= keystar);
}

// s_hist[bin] += 1 for the active lanes of the wave: the lanes sharing the
// first remaining lane's bin are added by one atomic, for up to 4 distinct
// bins; lanes left after that add one each.
__device__ __forceinline__ void wave_hist_add(unsigned int* s_hist, unsigned int bin, bool act) {
  unsigned long long am = __ballot(act);
  const int lane = (int)(threadIdx.x & 63);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (!am) return;  // uniform
    const int first = __ffsll((long long)am) - 1;
    const unsigned int b0 = (unsigned int)__shfl((int)bin, first, 64);
    const unsigned long long same = __ballot(((am >> lane) & 1ull) && bin == b0);
    if (lane == first) atomicAdd(&s_hist[b0], (unsigned int)__popcll(same));
    am &= ~same;
  }
  if ((am >> lane) & 1ull) atomicAdd(&s_hist[bin], 1u);
}

// Fixed-order sum of values given by rank (r in [0, total), total the same
// in every thread): thread t adds ranks t, t + NT, t + 2 NT, ... in
// increasing order, so the result depends on the ranked sequence alone, not
// on which thread held which value (a tile whose slots moved -- fewer of them,
// the same entries in the same order -- sums to the same bits).  The values
// go through buf (CAP doubles) in passes of CAP ranks; SQ: also the sum of
// squares into *sq.  Returns the thread's partial sums (for block_sum).
template <int R, int CAP, bool SQ = false>
__device__ __forceinline__ double rank_ordered_sum(const double* val, const int* rank, const bool* ok, int total,
                                                   double* buf, double* sq = nullptr) {
  double acc = 0.0, acc2 = 0.0;
  for (int lo = 0; lo < total; lo += CAP) {
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (ok[i] && rank[i] >= lo && rank[i] < lo + CAP) buf[rank[i] - lo] = val[i];
    __syncthreads();
    const int hi = total < lo + CAP ? total : lo + CAP;
    for (int r = lo + (int)threadIdx.x; r < hi; r += NT) {
      const double v = buf[r - lo];
      acc += v;
      if constexpr (SQ) acc2 += v * v;
    }
    __syncthreads();
  }
  if constexpr (SQ) *sq = acc2;
  return acc;
}

// One tile from its |a|^2 values vp (slot it * NT + threadIdx.x holds
// vp[it]; slots >= cnt are ignored): count / sum above bucket b1 into
// tile_dcnt[gt] / tile_dsum[gt] (the sum in the entries' order, by
// rank_ordered_sum), candidates inside it (in slot order: item-major, then
// thread order) at c1p / c1i[gt * SEL_TILE + ...] (|a|^2 bit pattern, input
// slot base + slot), their count into cnt1[gt], and with HIST their digit-2
// counts into s_hist.  Candidate and above flags are per-item wave ballots
// (scalar registers); an entry's rank is its item's and wave's base (one wave
// scans the (item, wave) counts) plus its rank among the wave's set lanes.
// s_wcnt: 2 * SEL_ITEMS * WAVES ints; s_abv: CAPA doubles.
// KEYS: vk holds the slots' stored words in the same order; candidates get
// their word in c1k and their own list index in c1i.
template <bool HIST, bool KEYS = false, int CAPA = 1024>
__device__ __forceinline__ void ex1_tile_core(const Ex1Out& O, int64_t gt, int64_t base, int cnt,
                                              const double (&vp)[SEL_ITEMS], unsigned int b1,
                                              unsigned int* s_hist, int* s_wcnt, double* s_red, double* s_abv,
                                              const uint64_t* vk = nullptr) {
  constexpr int IW = SEL_ITEMS * WAVES;
  static_assert(IW <= 64, "one wave scans the (item, wave) counts");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t cbase = gt * SEL_TILE;
  unsigned long long m[SEL_ITEMS], ma[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int j = it * NT + threadIdx.x;
    bool cand = false, above = false;
    if (j < cnt) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(vp[it]);
      const unsigned int d = (unsigned int)(bits >> 52);
      if (d > b1) {
        above = true;
      } else if (d == b1) {
        cand = true;
        if constexpr (HIST) atomicAdd(&s_hist[(unsigned int)(bits >> 41) & (HIST_BINS - 1)], 1u);
      }
    }
    m[it] = __ballot(cand);
    ma[it] = __ballot(above);
    if (lane == 0) {
      s_wcnt[it * WAVES + w] = __popcll(m[it]);
      s_wcnt[IW + it * WAVES + w] = __popcll(ma[it]);
    }
  }
  __syncthreads();
  const int cw = lane < IW ? s_wcnt[lane] : 0;
  const int cx = wave_incl_scan(cw) - cw;
  const int rb = __builtin_amdgcn_readlane(cx + cw, IW - 1);
  const int aw = lane < IW ? s_wcnt[IW + lane] : 0;
  const int ax = wave_incl_scan(aw) - aw;
  const int atot = __builtin_amdgcn_readlane(ax + aw, IW - 1);
  const int wu = __builtin_amdgcn_readfirstlane(w);
  int ar[SEL_ITEMS];
  bool aok[SEL_ITEMS];
#pragma unroll
  for (int it = 0; it < SEL_ITEMS; ++it) {
    const int o = __builtin_amdgcn_readlane(cx, it * WAVES + wu);
    const unsigned long long mk = m[it];
    if ((mk >> lane) & 1ull) {
      const int off =
          o + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mk, 0u));
      O.c1p[cbase + off] = (unsigned long long)__double_as_longlong(vp[it]);
      if constexpr (KEYS) {
        O.c1k[cbase + off] = vk[it];
        O.c1i[cbase + off] = (unsigned int)(cbase + off);
      } else {
        O.c1i[cbase + off] = (unsigned int)(base + it * NT + threadIdx.x);
      }
    }
    const unsigned long long mka = ma[it];
    aok[it] = (mka >> lane) & 1ull;
    ar[it] = __builtin_amdgcn_readlane(ax, it * WAVES + wu) +
             (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mka >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mka, 0u));
  }
  const double dsum = rank_ordered_sum<SEL_ITEMS, CAPA>(vp, ar, aok, atot, s_abv);
  const double s = block_sum(dsum, s_red);
  if (threadIdx.x == 0) {
    O.tile_dsum[gt] = s;
    O.tile_dcnt[gt] = atot;
    O.cnt1[gt] = rb;
  }
}

}  // namespace bass
