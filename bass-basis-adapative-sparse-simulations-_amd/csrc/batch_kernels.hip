// Elementwise operations on a batch of independent states (one job each):
// diagonal 2-qubit gate, scaling, |a|^2 / |a|^4 moments.  blockIdx.y = job;
// blocks of a job stride over its entries.  Same arithmetic as the
// single-state kernels (k_diag2q, k_scale, k_moments).
#include "common.h"
#include "kernels.h"

namespace bass {

constexpr int EW_BLOCKS = 64;  // blocks per job

// Jobs of the header (hdr->njobs, staged): blockIdx.y walks them with a
// persistent stride; a job's input count is its count cell when set.
__device__ __forceinline__ int64_t ew_count(const EwJob& J) { return J.cnt_in ? *J.cnt_in : J.n; }

__global__ __launch_bounds__(NT) void k_bdiag2q(const EwJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr) {
  const int nj = hdr->njobs;
  for (int c = blockIdx.y; c < nj; c += gridDim.y) {
    const EwJob& J = jobs[c];
    const int64_t n = ew_count(J);
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
      const uint64_t key = J.x[i];
      const int idx = (int)((((key >> J.j1) & 1ull) << 1) | ((key >> J.j2) & 1ull));
      J.a[i] = cmul(J.a[i], J.d[idx]);
    }
  }
}

__global__ __launch_bounds__(NT) void k_bscale(const EwJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr) {
  const int nj = hdr->njobs;
  for (int c = blockIdx.y; c < nj; c += gridDim.y) {
    const EwJob& J = jobs[c];
    const int64_t n = ew_count(J);
    const double inv = 1.0 / J.div;
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
      const double2 v = J.a[i];
      J.a[i] = make_double2(v.x * inv, v.y * inv);
    }
  }
}

// Moments in the order of the single-state path (k_moments + k_sums_reduce):
// per tile of SEG_ENTRIES entries a block sum, then per job a fixed-order sum
// over its tiles.  Blocks of a job take tiles blockIdx.x, + gridDim.x, ...
// (toff: the host's tile offsets, from staged counts).
__global__ __launch_bounds__(NT) void k_bmoments(const EwJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr,
                                                  double* __restrict__ ts2, double* __restrict__ ts4) {
  __shared__ double s_red[WAVES];
  const int nj = hdr->njobs;
  for (int c = blockIdx.y; c < nj; c += gridDim.y) {
    const EwJob& J = jobs[c];
    const int64_t n = J.n;
    const int64_t ntiles = (n + SEG_ENTRIES - 1) / SEG_ENTRIES;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
      double s2 = 0.0, s4 = 0.0;
      for (int it = 0; it < SEG_ENTRIES / NT; ++it) {
        const int64_t i = t * SEG_ENTRIES + (int64_t)it * NT + threadIdx.x;
        if (i < n) {
          const double p = cabs2(J.a[i]);
          s2 += p;
          s4 += p * p;
        }
      }
      const double t2 = block_sum(s2, s_red);
      const double t4 = block_sum(s4, s_red);
      if (threadIdx.x == 0) {
        ts2[J.toff + t] = t2;
        ts4[J.toff + t] = t4;
      }
    }
  }
}

__global__ __launch_bounds__(NTB) void k_bmoments_fin(const EwJob* __restrict__ jobs, const BatchHdr* __restrict__ hdr,
                                                       const double* __restrict__ ts2, const double* __restrict__ ts4,
                                                       OpStats* st) {
  __shared__ double s_red[WAVESB];
  const int nj = hdr->njobs;
  for (int c = blockIdx.x; c < nj; c += gridDim.x) {
    const EwJob& J = jobs[c];
    const int64_t ntiles = (J.n + SEG_ENTRIES - 1) / SEG_ENTRIES;
    double s2 = 0.0, s4 = 0.0;
    for (int64_t j = threadIdx.x; j < ntiles; j += NTB) s2 += ts2[J.toff + j];
    for (int64_t j = threadIdx.x; j < ntiles; j += NTB) s4 += ts4[J.toff + j];
    s2 = block_sum<NTB>(s2, s_red);
    s4 = block_sum<NTB>(s4, s_red);
    if (threadIdx.x == 0) {
      st[c].sum2 = s2;
      st[c].sum4 = s4;
    }
    __syncthreads();
  }
}

int launch_ew_batch(int op, const EwJob* d_jobs, const BatchHdr* hdr, int jobs, int64_t part_tiles,
                    double* part, OpStats* st, hipStream_t s) {
  // EW_BLOCKS blocks per job row; rows stride over the jobs
  auto rows = [&](const void* fn) {
    const unsigned g = persist_grid(fn, NT, (int64_t)jobs * EW_BLOCKS);
    const unsigned r = (g + EW_BLOCKS - 1) / EW_BLOCKS;
    return dim3(EW_BLOCKS, r < 1 ? 1 : (r > 65535 ? 65535 : r));
  };
  if (op == 0) {
    hipLaunchKernelGGL(k_bdiag2q, rows(reinterpret_cast<const void*>(&k_bdiag2q)), dim3(NT), 0, s, d_jobs, hdr);
  } else if (op == 1) {
    hipLaunchKernelGGL(k_bscale, rows(reinterpret_cast<const void*>(&k_bscale)), dim3(NT), 0, s, d_jobs, hdr);
  } else {
    // moments: staged counts only (the per-tile offsets toff follow from them)
    double* ts2 = part;
    double* ts4 = part + part_tiles;
    hipLaunchKernelGGL(k_bmoments, rows(reinterpret_cast<const void*>(&k_bmoments)), dim3(NT), 0, s, d_jobs, hdr,
                       ts2, ts4);
    hipLaunchKernelGGL(k_bmoments_fin, dim3(persist_grid(reinterpret_cast<const void*>(&k_bmoments_fin), NTB, jobs)),
                       dim3(NTB), 0, s, d_jobs, hdr, ts2, ts4, st);
  }
  return (int)hipGetLastError();
}

size_t ew_batch_scratch(int64_t ntiles_total) { return (size_t)ntiles_total * 16 + 256; }

}  // namespace bass
