// Dense state vectors on the device: the exact reference simulation and the
// dense reconstruction of a sparse state used for fidelities
// (src/simulation/exact_simulator.py:66-144, BASS.to_statevector
// src/simulation/bass_simulator.py:608-621, fidelity_bass
// src/experimentation/runner.py:615-626).
//
// psi: complex128[2^N], index = basis key (qubit q = bit q).  Gates are
// applied in place, one thread per group of 2 (1-qubit) or 4 (2-qubit)
// amplitudes; products are summed in matrix-column order.  All kernels stream
// the vector once (32 B per amplitude for a gate: read + write), so they are
// HBM-bound; the dot product reduces per-block partials in a fixed order.
#include "common.h"
#include "kernels.h"

namespace bass {

__global__ __launch_bounds__(NT) void k_dense_zero(double2* psi, int64_t len) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < len; i += stride) psi[i] = make_double2(0.0, 0.0);
}

// psi[key(x_i)] = a_i; key = stored word mapped back through km (identity when
// km.nbits == 0).
__global__ __launch_bounds__(NT) void k_dense_scatter(const uint64_t* __restrict__ x, const double2* __restrict__ a,
                                                      int64_t n, KeyMap km, double2* __restrict__ psi) {
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const uint64_t w = x[i];
  uint64_t key = w;
  if (km.nbits) {
    key = 0;
    for (int p = 0; p < km.nbits; ++p) key |= ((w >> p) & 1ull) << km.q[p];
  }
  psi[key] = a[i];
}

// Index with a zero bit inserted at position q.
__device__ __forceinline__ int64_t insert_zero(int64_t i, int q) {
  const int64_t lo = i & ((1ll << q) - 1);
  return ((i >> q) << (q + 1)) | lo;
}

// out_b = sum_c U[b][c] in_c over the pair (bit q clear, bit q set).
__global__ __launch_bounds__(NT) void k_dense_1q(double2* __restrict__ psi, int64_t half, int q, GateMat G) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < half; i += stride) {
    const int64_t i0 = insert_zero(i, q), i1 = i0 | (1ll << q);
    const double2 a0 = psi[i0], a1 = psi[i1];
    psi[i0] = cadd(cmul(G.m[0], a0), cmul(G.m[1], a1));
    psi[i1] = cadd(cmul(G.m[2], a0), cmul(G.m[3], a1));
  }
}

// 4x4 on qubits (q1, q2), matrix index (b_q1 << 1) | b_q2; qa < qb are the
// same qubits in bit order.
__global__ __launch_bounds__(NT) void k_dense_2q(double2* __restrict__ psi, int64_t quarter, int q1, int q2,
                                                 GateMat G) {
  const int qa = q1 < q2 ? q1 : q2, qb = q1 < q2 ? q2 : q1;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < quarter; i += stride) {
    const int64_t base = insert_zero(insert_zero(i, qa), qb);
    int64_t idx[4];
    double2 in[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      idx[c] = base | ((int64_t)((c >> 1) & 1) << q1) | ((int64_t)(c & 1) << q2);
      in[c] = psi[idx[c]];
    }
    double2 out[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      double2 acc = cmul(G.m[o * 4 + 0], in[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(G.m[o * 4 + c], in[c]));
      out[o] = acc;
    }
#pragma unroll
    for (int o = 0; o < 4; ++o) psi[idx[o]] = out[o];
  }
}

// Partial sums of conj(u_i) v_i: block b covers [b * DOT_CHUNK, +DOT_CHUNK).
constexpr int DOT_CHUNK = NT * 16;

__global__ __launch_bounds__(NT) void k_dense_vdot(const double2* __restrict__ u, const double2* __restrict__ v,
                                                   int64_t len, double2* __restrict__ part) {
  __shared__ double s_re[WAVES], s_im[WAVES];
  const int64_t b0 = (int64_t)blockIdx.x * DOT_CHUNK;
  double re = 0.0, im = 0.0;
  constexpr int G = 4;  // loads of G items issued together, clamped to valid indices
  for (int k0 = 0; k0 < DOT_CHUNK / NT; k0 += G) {
    double2 uu[G], vv[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int64_t i = b0 + (int64_t)(k0 + g) * NT + threadIdx.x;
      const int64_t c = i < len ? i : 0;
      uu[g] = u[c];
      vv[g] = v[c];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int64_t i = b0 + (int64_t)(k0 + g) * NT + threadIdx.x;
      if (i < len) {
        const double2 z = cmul(cconj(uu[g]), vv[g]);
        re += z.x;
        im += z.y;
      }
    }
  }
  const double tr = block_sum(re, s_re);
  const double ti = block_sum(im, s_im);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(tr, ti);
}

__global__ __launch_bounds__(NTB) void k_dense_vdot_fin(const double2* __restrict__ part, int64_t nb, double2* out) {
  __shared__ double s_re[WAVESB], s_im[WAVESB];
  double re = 0.0, im = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += NTB) {
    re += part[b].x;
    im += part[b].y;
  }
  const double tr = block_sum<NTB>(re, s_re);
  const double ti = block_sum<NTB>(im, s_im);
  if (threadIdx.x == 0) *out = make_double2(tr, ti);
}

static unsigned grid_for(int64_t items) {
  int64_t b = (items + NT - 1) / NT;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

int launch_dense_scatter(const uint64_t* x, const double2* a, int64_t n, const KeyMap& km, double2* psi, int64_t len,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_dense_zero, dim3(grid_for(len)), dim3(NT), 0, s, psi, len);
  if (n > 0) hipLaunchKernelGGL(k_dense_scatter, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, s, x, a, n, km, psi);
  return (int)hipGetLastError();
}

int launch_dense_gate(double2* psi, int N, int nq, int q0, int q1, const GateMat& G, hipStream_t s) {
  if (nq == 1) {
    const int64_t half = 1ll << (N - 1);
    hipLaunchKernelGGL(k_dense_1q, dim3(grid_for(half)), dim3(NT), 0, s, psi, half, q0, G);
  } else {
    const int64_t quarter = 1ll << (N - 2);
    hipLaunchKernelGGL(k_dense_2q, dim3(grid_for(quarter)), dim3(NT), 0, s, psi, quarter, q0, q1, G);
  }
  return (int)hipGetLastError();
}

size_t dense_vdot_scratch(int64_t len) { return (size_t)((len + DOT_CHUNK - 1) / DOT_CHUNK + 1) * 16 + 256; }

int launch_dense_vdot(const double2* u, const double2* v, int64_t len, void* scratch, double2* out, hipStream_t s) {
  const int64_t nb = (len + DOT_CHUNK - 1) / DOT_CHUNK;
  double2* part = (double2*)scratch;
  hipLaunchKernelGGL(k_dense_vdot, dim3((unsigned)nb), dim3(NT), 0, s, u, v, len, part);
  hipLaunchKernelGGL(k_dense_vdot_fin, dim3(1), dim3(NTB), 0, s, part, nb, out);
  return (int)hipGetLastError();
}

}  // namespace bass
