// Gate application on the SoA sparse state.
//
// One templated emission kernel covers every gate-like operation of the
// reference (src/core/sparse_state.py, src/simulation/bass_simulator.py):
//
//   DEDUP=false, MODE=GATE : apply_1qubit_gate_kernel_opt / apply_2qubit_gate_kernel_opt
//                            (each entry emits alone, exact-zero products dropped,
//                             output in input order, out_idx ascending)
//   DEDUP=true,  MODE=GATE : the deduplicated result of emit -> sort -> merge
//                            (1q) and of apply_2qubit_gate_ht (2q)
//   DEDUP=true,  MODE=ROT  : apply_1qubit_rotation_sparse / apply_2qubit_rotation_sparse
//                            (absent members read as 0, outputs with |a| <= 1e-30 dropped)
//
// Deduplication does not sort.  Entries sharing the bits outside the gate
// qubits ("rest") form a group of at most 2^QB members.  A group table
// (open addressing, linear probing, 64-bit CAS insert of the rest key) records
// each member's input index.  The member with the smallest input index is the
// group leader: it gathers the other members' amplitudes and emits every
// output of the group, summing contributions in increasing input index -- the
// order in which the reference's sequential loop accumulates them.
//
// Outputs are compacted in input order of their leaders with a decoupled
// look-back scan, so results are deterministic run to run.  While writing,
// each tile adds |a|^2 of its outputs to a 2048-bin histogram of the top 12
// bits of the IEEE pattern (sign + exponent) -- the first digit of the top-k
// radix select -- and a per-tile sum of |a|^2.
#include "common.h"
#include "kernels.h"

namespace bass {

template <int QB>
__device__ __forceinline__ int member_index(uint64_t key, int q0, int q1) {
  if constexpr (QB == 1) {
    return (int)((key >> q0) & 1ull);
  } else {
    return (int)((((key >> q0) & 1ull) << 1) | ((key >> q1) & 1ull));
  }
}

template <int QB>
__device__ __forceinline__ uint64_t out_bits(int o, int q0, int q1) {
  if constexpr (QB == 1) {
    return (uint64_t)(o & 1) << q0;
  } else {
    return ((uint64_t)((o >> 1) & 1) << q0) | ((uint64_t)(o & 1) << q1);
  }
}

// Group table slot: the rest key and the member indices share one slot
// (16 B for 1 qubit, 32 B for 2) so a probe and its member reads touch one line.
template <int QB>
struct GSlot {
  static constexpr int WORDS = QB == 1 ? 2 : 4;  // u64 words per slot
};

// Insert: lanes of a wave whose entries share a rest key (members of one
// group) are matched by ballots first -- one round per distinct key of the
// wave, the lowest unmatched lane's key each round -- and only the lowest lane
// of each match probes and CASes the table; the others take its slot by a
// shuffle and write their member index there.  Colliding inserts of one group
// inside a wave thus cost one probe sequence instead of up to 2^QB.
template <int QB>
__global__ __launch_bounds__(NT) void k_group_build(const uint64_t* __restrict__ x, int64_t n, int q0,
                                                     int q1, uint64_t* __restrict__ gslot, int shift,
                                                     uint64_t capmask, int* err) {
  constexpr int M = 1 << QB;
  constexpr int W = GSlot<QB>::WORDS;
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  const bool valid = i < n;  // every lane takes part in the wave-level match
  const int lane = threadIdx.x & 63;
  uint64_t r = 0;
  int c = 0;
  if (valid) {
    const uint64_t key = x[i];
    r = key & ~out_bits<QB>(M - 1, q0, q1);
    c = member_index<QB>(key, q0, q1);
  }
  int leader = lane;
  unsigned long long pend = __ballot(valid);
  while (pend) {  // wave-uniform
    const int f = __ffsll((long long)pend) - 1;
    const uint64_t rf = (uint64_t)__shfl((long long)r, f, 64);
    const unsigned long long same = __ballot(valid && r == rf) & pend;
    if ((same >> lane) & 1ull) leader = f;
    pend &= ~same;
  }
  long long hs = -1;  // slot of the group (leaders), -1: none
  if (valid && leader == lane) {
    uint64_t h = hslot(r, shift);
    for (uint64_t probe = 0; probe <= capmask; ++probe) {
      uint64_t* slot = gslot + h * W;
      uint64_t cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == EMPTY) {
        unsigned long long prev = atomicCAS((unsigned long long*)slot, (unsigned long long)EMPTY,
                                            (unsigned long long)r);
        cur = (prev == (unsigned long long)EMPTY) ? r : (uint64_t)prev;
      }
      if (cur == r) {
        hs = (long long)h;
        break;
      }
      h = (h + 1) & capmask;
    }
    if (hs < 0) set_err(err, ERR_PROBE);
  }
  hs = __shfl(hs, leader, 64);
  if (valid && hs >= 0) reinterpret_cast<int32_t*>(gslot + (uint64_t)hs * W + 1)[c] = (int32_t)i;
}

template <int QB>
__device__ __forceinline__ const uint64_t* group_find(const uint64_t* __restrict__ gslot, uint64_t r,
                                                      int shift, uint64_t capmask) {
  constexpr int W = GSlot<QB>::WORDS;
  uint64_t h = hslot(r, shift);
  for (uint64_t probe = 0; probe <= capmask; ++probe) {
    const uint64_t* slot = gslot + h * W;
    const uint64_t cur = slot[0];
    if (cur == r) return slot;
    if (cur == EMPTY) return nullptr;
    h = (h + 1) & capmask;
  }
  return nullptr;
}

template <int QB, int MODE, bool DEDUP, int ITEMS>
__global__ __launch_bounds__(NT) void k_emit(const uint64_t* __restrict__ x,
                                              const double2* __restrict__ a, int64_t n, int q0, int q1,
                                              const uint64_t* __restrict__ gslot, int shift,
                                              uint64_t capmask, GateMat G, uint64_t* __restrict__ xo,
                                              double2* __restrict__ ao, int64_t cap_out,
                                              unsigned long long* lb_status, unsigned int* ticket,
                                              unsigned int* hist, double* tile_psum, int64_t* n_out,
                                              int* seg_cnt, int* err) {
  constexpr int M = 1 << QB;
  constexpr int TILE = NT * ITEMS;
  __shared__ unsigned int s_hist[HIST_BINS];
  __shared__ int s_scan[WAVES];
  __shared__ double s_red[WAVES];
  __shared__ unsigned int s_tile;
  __shared__ unsigned long long s_base;

  // Segmented output (seg_cnt != null): tile t writes its outputs to the fixed
  // segment [t * M * TILE, ...) and its count to seg_cnt[t]; no inter-tile
  // coordination.  Dense output: tiles in dispatch order + decoupled look-back.
  const bool segmented = seg_cnt != nullptr;
  const unsigned int tile = segmented ? blockIdx.x : take_ticket(ticket, &s_tile);
  if (hist) {
    for (int b = threadIdx.x; b < HIST_BINS; b += NT) s_hist[b] = 0;
  }
  __syncthreads();

  const uint64_t gmask = out_bits<QB>(M - 1, q0, q1);
  uint64_t okey[ITEMS][M];
  double2 oamp[ITEMS][M];
  unsigned int omask[ITEMS];
  int ooff[ITEMS];
  int rtot[ITEMS];

#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int64_t i = (int64_t)tile * TILE + (int64_t)it * NT + threadIdx.x;
    unsigned int mask = 0;
    if (i < n) {
      const uint64_t key = x[i];
      const int c = member_index<QB>(key, q0, q1);
      const uint64_t r = key & ~gmask;
      int32_t mem[M];
#pragma unroll
      for (int cc = 0; cc < M; ++cc) mem[cc] = -1;
      if constexpr (DEDUP) {
        const uint64_t* slot = group_find<QB>(gslot, r, shift, capmask);
        if (!slot) {
          set_err(err, ERR_MISSING);
          mem[c] = (int32_t)i;
        } else {
          const int32_t* sm = reinterpret_cast<const int32_t*>(slot + 1);
#pragma unroll
          for (int cc = 0; cc < M; ++cc) mem[cc] = sm[cc];
        }
      } else {
        mem[c] = (int32_t)i;
      }
      int32_t lead = 0x7FFFFFFF;
#pragma unroll
      for (int cc = 0; cc < M; ++cc)
        if (mem[cc] >= 0 && mem[cc] < lead) lead = mem[cc];
      if (lead == (int32_t)i) {
        double2 av[M];
#pragma unroll
        for (int cc = 0; cc < M; ++cc) {
          if (mem[cc] < 0)
            av[cc] = make_double2(0.0, 0.0);
          else if (cc == c)
            av[cc] = a[i];
          else
            av[cc] = a[mem[cc]];
        }
        // rank of each present member by input index (summation order)
        int rank[M];
#pragma unroll
        for (int cc = 0; cc < M; ++cc) {
          int rk = 0;
#pragma unroll
          for (int c2 = 0; c2 < M; ++c2) rk += (mem[c2] >= 0 && mem[c2] < mem[cc]) ? 1 : 0;
          rank[cc] = rk;
        }
#pragma unroll
        for (int o = 0; o < M; ++o) {
          double2 acc = make_double2(0.0, 0.0);
          bool any = false;
          if constexpr (MODE == MODE_GATE) {
            double2 term[M];
            bool live[M];
#pragma unroll
            for (int cc = 0; cc < M; ++cc) {
              term[cc] = cmul(G.m[o * M + cc], av[cc]);
              live[cc] = mem[cc] >= 0 && cnonzero(term[cc]);
            }
#pragma unroll
            for (int rr = 0; rr < M; ++rr) {
#pragma unroll
              for (int cc = 0; cc < M; ++cc) {
                if (live[cc] && rank[cc] == rr) {
                  acc = any ? cadd(acc, term[cc]) : term[cc];
                  any = true;
                }
              }
            }
          } else {
            acc = cmul(G.m[o * M + 0], av[0]);
#pragma unroll
            for (int cc = 1; cc < M; ++cc) acc = cadd(acc, cmul(G.m[o * M + cc], av[cc]));
            any = hypot(acc.x, acc.y) > 1e-30;
          }
          if (any) {
            mask |= 1u << o;
            okey[it][o] = r | out_bits<QB>(o, q0, q1);
            oamp[it][o] = acc;
          }
        }
      }
    }
    omask[it] = mask;
  }
  // scans after every item's loads and probes are issued
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    int tot;
    ooff[it] = block_excl_scan(__popc(omask[it]), tot, s_scan);
    rtot[it] = tot;
  }

  unsigned long long tile_total = 0;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) tile_total += (unsigned long long)rtot[it];
  if (segmented) {
    if (threadIdx.x == 0) {
      s_base = (unsigned long long)tile * (unsigned long long)(M * TILE);
      seg_cnt[tile] = (int)tile_total;
    }
  } else if (threadIdx.x < 64) {
    const unsigned long long e = lookback_prefix_wave(lb_status, tile, tile_total, err);
    if (threadIdx.x == 0) s_base = e;
  }
  __syncthreads();
  const unsigned long long base = s_base;
  const bool fits = base + tile_total <= (unsigned long long)cap_out;
  if (!fits && threadIdx.x == 0) set_err(err, ERR_OVERFLOW);

  double psum = 0.0;
  unsigned long long rb = base;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    unsigned long long pos = rb + (unsigned long long)ooff[it];
#pragma unroll
    for (int o = 0; o < M; ++o) {
      if ((omask[it] >> o) & 1u) {
        const double2 v = oamp[it][o];
        if (fits) {
          xo[pos] = okey[it][o];
          ao[pos] = v;
        }
        const double p = cabs2(v);
        psum += p;
        if (hist) atomicAdd(&s_hist[(unsigned int)(__double_as_longlong(p) >> 52) & (HIST_BINS - 1)], 1u);
        ++pos;
      }
    }
    rb += (unsigned long long)rtot[it];
  }
  if (!segmented && tile == gridDim.x - 1 && threadIdx.x == 0) *n_out = (int64_t)(base + tile_total);
  if (tile_psum) {
    const double s = block_sum(psum, s_red);
    if (threadIdx.x == 0) tile_psum[tile] = s;
  }
  if (hist) {
    __syncthreads();
    for (int b = threadIdx.x; b < HIST_BINS; b += NT)
      if (s_hist[b]) atomicAdd(&hist[b], s_hist[b]);
  }
}

// In-place diagonal 2-qubit gate, src/core/sparse_state.py:178-190:
// alpha[i] *= diag[(b_q1 << 1) | b_q2].
__global__ __launch_bounds__(NT) void k_diag2q(const uint64_t* __restrict__ x, double2* __restrict__ a,
                                                int64_t n, int q1, int q2, double2 d0, double2 d1,
                                                double2 d2, double2 d3, const int64_t* __restrict__ n_dev) {
  n = dev_count(n_dev, n);
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
    const uint64_t key = x[i];
    const int idx = (int)((((key >> q1) & 1ull) << 1) | ((key >> q2) & 1ull));
    const double2 d = idx == 0 ? d0 : idx == 1 ? d1 : idx == 2 ? d2 : d3;
    a[i] = cmul(a[i], d);
  }
}

// A run of diagonal 2-qubit gates in one pass: per entry, gate g = 0, 1, ...
// in order multiplies the amplitude by its diagonal entry -- the products of
// k_diag2q, in the same order, so the result equals one launch per gate bit
// for bit, from one read and one write of the state.  The gate tables sit in
// LDS (16 B per entry, indexed by the entry's two bits); each thread keeps
// DIAG_ITEMS entries in flight.
constexpr int DIAG_ITEMS = 4;
__global__ __launch_bounds__(NT) void k_diag2q_run(const uint64_t* __restrict__ x, double2* __restrict__ a,
                                                    int64_t n, DiagRun R, const int64_t* __restrict__ n_dev) {
  __shared__ double2 s_d[DIAG_RUN_MAX * 4];
  for (int e = threadIdx.x; e < R.ngates * 4; e += NT) s_d[e] = R.d[e >> 2][e & 3];
  __syncthreads();
  n = dev_count(n_dev, n);
  const int64_t per = (int64_t)NT * DIAG_ITEMS;
  for (int64_t b = (int64_t)blockIdx.x * per; b < n; b += (int64_t)gridDim.x * per) {
    uint64_t key[DIAG_ITEMS];
    double2 v[DIAG_ITEMS];
#pragma unroll
    for (int it = 0; it < DIAG_ITEMS; ++it) {
      const int64_t i = b + it * NT + threadIdx.x;
      if (i < n) {
        key[it] = x[i];
        v[it] = a[i];
      }
    }
    for (int g = 0; g < R.ngates; ++g) {
      const int q1 = R.q[g][0], q2 = R.q[g][1];
#pragma unroll
      for (int it = 0; it < DIAG_ITEMS; ++it) {
        const int idx = (int)((((key[it] >> q1) & 1ull) << 1) | ((key[it] >> q2) & 1ull));
        v[it] = cmul(v[it], s_d[g * 4 + idx]);
      }
    }
#pragma unroll
    for (int it = 0; it < DIAG_ITEMS; ++it) {
      const int64_t i = b + it * NT + threadIdx.x;
      if (i < n) a[i] = v[it];
    }
  }
}

// ----------------------------------------------------------- host launch ---

int group_slot_bytes(int qb) { return qb == 1 ? 8 * GSlot<1>::WORDS : 8 * GSlot<2>::WORDS; }

int launch_group_build(int qb, const uint64_t* x, int64_t n, int q0, int q1, uint64_t* gslot, int log2cap,
                       int* err, hipStream_t s) {
  if (n <= 0) return 0;
  const int shift = 64 - log2cap;
  const uint64_t capmask = (1ull << log2cap) - 1;
  const dim3 grid((unsigned)((n + NT - 1) / NT));
  if (qb == 1)
    hipLaunchKernelGGL(k_group_build<1>, grid, dim3(NT), 0, s, x, n, q0, q1, gslot, shift, capmask, err);
  else
    hipLaunchKernelGGL(k_group_build<2>, grid, dim3(NT), 0, s, x, n, q0, q1, gslot, shift, capmask, err);
  return (int)hipGetLastError();
}

// Input entries per emission tile: NT x items, items per thread 2, 4 (default)
// or 8 for 1 qubit and 1, 2 (default) or 4 for 2 qubits.  A segmented output
// needs the default (a tile then fills SEG_ENTRIES output slots).
int emit_default_items(int qb) { return qb == 1 ? 4 : 2; }
bool emit_items_ok(int qb, int items) {
  return qb == 1 ? (items == 2 || items == 4 || items == 8) : (items == 1 || items == 2 || items == 4);
}
int emit_tile_entries(int qb, int items) { return NT * (items > 0 ? items : emit_default_items(qb)); }

template <int QB, int ITEMS>
static void emit_variant(int mode, bool dedup, dim3 grid, hipStream_t s, const uint64_t* x, const double2* a,
                         int64_t n, int q0, int q1, const uint64_t* gslot, int shift, uint64_t capmask,
                         const GateMat& G, uint64_t* xo, double2* ao, int64_t cap_out,
                         unsigned long long* lb_status, unsigned int* ticket, unsigned int* hist,
                         double* tile_psum, int64_t* n_out, int* seg_cnt, int* err) {
#define BASS_EMIT(MODE_, DEDUP_)                                                                     \
  hipLaunchKernelGGL((k_emit<QB, MODE_, DEDUP_, ITEMS>), grid, dim3(NT), 0, s, x, a, n, q0, q1,     \
                     gslot, shift, capmask, G, xo, ao, cap_out, lb_status, ticket, hist, tile_psum,  \
                     n_out, seg_cnt, err)
  if (!dedup)
    BASS_EMIT(MODE_GATE, false);
  else if (mode == MODE_GATE)
    BASS_EMIT(MODE_GATE, true);
  else
    BASS_EMIT(MODE_ROT, true);
#undef BASS_EMIT
}

int launch_emit(int qb, int mode, bool dedup, const uint64_t* x, const double2* a, int64_t n, int q0,
                int q1, const uint64_t* gslot, int log2cap, const GateMat& G,
                uint64_t* xo, double2* ao, int64_t cap_out, unsigned long long* lb_status,
                unsigned int* ticket, unsigned int* hist, double* tile_psum, int64_t* n_out, int* seg_cnt,
                int* err, hipStream_t s, int items) {
  if (n <= 0) return 0;
  if (items <= 0 || seg_cnt) items = emit_default_items(qb);
  const int shift = 64 - log2cap;
  const uint64_t capmask = log2cap > 0 ? ((1ull << log2cap) - 1) : 0;
  const int tile = emit_tile_entries(qb, items);
  const dim3 grid((unsigned)((n + tile - 1) / tile));
#define BASS_EV(QB_, IT_)                                                                           \
  emit_variant<QB_, IT_>(mode, dedup, grid, s, x, a, n, q0, q1, gslot, shift, capmask, G, xo, ao,  \
                         cap_out, lb_status, ticket, hist, tile_psum, n_out, seg_cnt, err)
  if (qb == 1) {
    if (items == 2)
      BASS_EV(1, 2);
    else if (items == 8)
      BASS_EV(1, 8);
    else
      BASS_EV(1, 4);
  } else {
    if (items == 1)
      BASS_EV(2, 1);
    else if (items == 4)
      BASS_EV(2, 4);
    else
      BASS_EV(2, 2);
  }
#undef BASS_EV
  return (int)hipGetLastError();
}

int launch_diag2q_run(const uint64_t* x, double2* a, int64_t n, const DiagRun& R, hipStream_t s,
                      const int64_t* n_dev) {
  if (n <= 0 || R.ngates <= 0) return 0;
  int64_t blocks = (n + (int64_t)NT * DIAG_ITEMS - 1) / ((int64_t)NT * DIAG_ITEMS);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_diag2q_run, dim3((unsigned)blocks), dim3(NT), 0, s, x, a, n, R, n_dev);
  return (int)hipGetLastError();
}

int launch_diag2q(const uint64_t* x, double2* a, int64_t n, int q1, int q2, const double2* d,
                  hipStream_t s, const int64_t* n_dev) {
  if (n <= 0) return 0;
  int64_t blocks = (n + NT - 1) / NT;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_diag2q, dim3((unsigned)blocks), dim3(NT), 0, s, x, a, n, q1, q2, d[0], d[1], d[2],
                     d[3], n_dev);
  return (int)hipGetLastError();
}

}  // namespace bass
