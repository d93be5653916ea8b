// All single-qubit reduced density matrices in one pass,
// src/simulation/bass_simulator.py:441-470 (compute_all_1qubit_rdms), and the
// key -> index hash table it probes (bass_simulator.py:100-153).
//
// rho_q[b][b] = sum |a_x|^2 over entries with bit q = b
// rho_q[0][1] = sum over bit-q-0 entries x whose partner x | 2^q is present of
//               a_x * conj(a_partner)      (rho_q[1][0] is its exact conjugate)
//
// Each block owns a fixed contiguous chunk of entries and keeps one FP64
// accumulator slot per (wave, qubit); slots and blocks are combined in a fixed
// order, so the result does not depend on scheduling.
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int RDM_ITEMS = 4;
constexpr int RDM_BLOCKS = 512;
constexpr int RDM_MAXQ = 64;

__global__ __launch_bounds__(NT) void k_table_build(const uint64_t* __restrict__ x, int64_t n,
                                                     uint64_t* __restrict__ tkeys,
                                                     int32_t* __restrict__ tvals, int shift,
                                                     uint64_t capmask, int overwrite, int* err) {
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = x[i];
  uint64_t h = hslot(key, shift);
  for (uint64_t probe = 0; probe <= capmask; ++probe) {
    uint64_t cur = __hip_atomic_load(&tkeys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == EMPTY) {
      unsigned long long prev = atomicCAS((unsigned long long*)&tkeys[h], (unsigned long long)EMPTY,
                                          (unsigned long long)key);
      if (prev == (unsigned long long)EMPTY) {
        tvals[h] = (int32_t)i;
        return;
      }
      cur = (uint64_t)prev;
    }
    if (cur == key) {
      // duplicate key in one batch: the largest index wins (the sequential
      // reference overwrites in index order, ht_insert :140-153)
      if (overwrite) atomicMax(&tvals[h], (int32_t)i);
      return;
    }
    h = (h + 1) & capmask;
  }
  set_err(err, ERR_PROBE);
}

__device__ __forceinline__ int32_t table_get(const uint64_t* __restrict__ tkeys,
                                             const int32_t* __restrict__ tvals, uint64_t key, int shift,
                                             uint64_t capmask) {
  uint64_t h = hslot(key, shift);
  for (uint64_t probe = 0; probe <= capmask; ++probe) {
    const uint64_t cur = tkeys[h];
    if (cur == key) return tvals[h];
    if (cur == EMPTY) return -1;
    h = (h + 1) & capmask;
  }
  return -1;
}

__global__ __launch_bounds__(NT) void k_table_lookup(const uint64_t* __restrict__ tkeys,
                                                      const int32_t* __restrict__ tvals, int shift,
                                                      uint64_t capmask, const uint64_t* __restrict__ q,
                                                      int64_t m, int64_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t j = (int64_t)blockIdx.x * NT + threadIdx.x; j < m; j += stride)
    out[j] = (int64_t)table_get(tkeys, tvals, q[j], shift, capmask);
}

__global__ __launch_bounds__(NT) void k_rdm(const uint64_t* __restrict__ x, const double2* __restrict__ a,
                                             int64_t n, int N, const uint64_t* __restrict__ tkeys,
                                             const int32_t* __restrict__ tvals, int shift, uint64_t capmask,
                                             double* __restrict__ part) {
  __shared__ double s_acc[WAVES][RDM_MAXQ][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < WAVES * RDM_MAXQ * 4; j += NT) (&s_acc[0][0][0])[j] = 0.0;
  __syncthreads();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  for (int64_t t0 = lo; t0 < hi; t0 += (int64_t)NT * RDM_ITEMS) {
    uint64_t kx[RDM_ITEMS];
    double2 av[RDM_ITEMS];
    double pr[RDM_ITEMS];
    bool ok[RDM_ITEMS];
#pragma unroll
    for (int it = 0; it < RDM_ITEMS; ++it) {
      const int64_t i = t0 + (int64_t)it * NT + threadIdx.x;
      ok[it] = i < hi;
      kx[it] = ok[it] ? x[i] : 0ull;
      av[it] = ok[it] ? a[i] : make_double2(0.0, 0.0);
      pr[it] = cabs2(av[it]);
    }
    for (int q = 0; q < N; ++q) {
      double d0 = 0.0, d1 = 0.0, ore = 0.0, oim = 0.0;
#pragma unroll
      for (int it = 0; it < RDM_ITEMS; ++it) {
        if (!ok[it]) continue;
        const bool bit = (kx[it] >> q) & 1ull;
        if (bit) {
          d1 += pr[it];
        } else {
          d0 += pr[it];
          const int32_t j = table_get(tkeys, tvals, kx[it] | (1ull << q), shift, capmask);
          if (j >= 0) {
            const double2 t = cmul(av[it], cconj(a[j]));  // a_x * conj(a_j)
            ore += t.x;
            oim += t.y;
          }
        }
      }
      d0 = wave_sum(d0);
      d1 = wave_sum(d1);
      ore = wave_sum(ore);
      oim = wave_sum(oim);
      if (lane == 0) {
        s_acc[w][q][0] += d0;
        s_acc[w][q][1] += d1;
        s_acc[w][q][2] += ore;
        s_acc[w][q][3] += oim;
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N * 4; j += NT) {
    const int q = j >> 2, c = j & 3;
    double s = 0.0;
    for (int ww = 0; ww < WAVES; ++ww) s += s_acc[ww][q][c];
    part[(int64_t)blockIdx.x * N * 4 + j] = s;
  }
}

__global__ __launch_bounds__(NT) void k_rdm_finalize(const double* __restrict__ part, int nblocks, int N,
                                                      double* __restrict__ rho) {
  for (int j = threadIdx.x; j < N * 4; j += NT) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * N * 4 + j];
    rho[j] = s;
  }
}

int rdm_blocks() { return RDM_BLOCKS; }

int launch_table_build(const uint64_t* x, int64_t n, uint64_t* tkeys, int32_t* tvals, int log2cap,
                       int overwrite, int* err, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_table_build, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, s, x, n, tkeys, tvals,
                     64 - log2cap, (1ull << log2cap) - 1, overwrite, err);
  return (int)hipGetLastError();
}

int launch_table_lookup(const uint64_t* tkeys, const int32_t* tvals, int log2cap, const uint64_t* q,
                        int64_t m, int64_t* out, hipStream_t s) {
  if (m <= 0) return 0;
  int64_t blocks = (m + NT - 1) / NT;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_table_lookup, dim3((unsigned)blocks), dim3(NT), 0, s, tkeys, tvals, 64 - log2cap,
                     (1ull << log2cap) - 1, q, m, out);
  return (int)hipGetLastError();
}

int launch_rdm(const uint64_t* x, const double2* a, int64_t n, int N, const uint64_t* tkeys,
               const int32_t* tvals, int log2cap, double* part, double* rho, hipStream_t s) {
  int blocks = RDM_BLOCKS;
  if (n < (int64_t)blocks * NT) {
    blocks = (int)((n + NT - 1) / NT);
    if (blocks < 1) blocks = 1;
  }
  hipLaunchKernelGGL(k_rdm, dim3(blocks), dim3(NT), 0, s, x, a, n, N, tkeys, tvals, 64 - log2cap,
                     (1ull << log2cap) - 1, part);
  hipLaunchKernelGGL(k_rdm_finalize, dim3(1), dim3(NT), 0, s, part, blocks, N, rho);
  return (int)hipGetLastError();
}

}  // namespace bass
