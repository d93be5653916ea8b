// Single-thread C++ restatement of the reference's 1-qubit BASS step: the CPU
// baseline timed beside the GPU.  Oracle code (test infrastructure): only the
// tests and the benchmark's CPU baseline load it; the product never does.
//
//   emit   apply_1qubit_gate_kernel_opt (src/core/sparse_state.py:15-67):
//          per entry (x & ~m, G[0][b] a) then (x | m, G[1][b] a), exact zeros
//          dropped; complex products with one fused multiply-add per
//          component, the contraction Numba's fastmath applies;
//   sort   radix_sort_uint64 (:150-172): 4 LSD passes of 16 bits, counts in
//          65536 buckets, stable scatter;
//   merge  merge_sorted_kernel (:127-147): run-length sum of equal keys;
//   top-k  the driver's truncation (src/simulation/bass_simulator.py:899-916):
//          the k largest |a|^2 by selection (np.argpartition there,
//          std::nth_element here; the order of the kept entries and the choice
//          among exact ties at the threshold are implementation defined in
//          both), gamma factor kept / total, then the renormalisation
//          a / sqrt(sum |a|^2) (:551-554).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline void cmul(double xr, double xi, double yr, double yi, double& re, double& im) {
  re = std::fma(xr, yr, -(xi * yi));
  im = std::fma(xr, yi, xi * yr);
}

int64_t emit(const uint64_t* x, const double* a, int64_t n, int q, const double* G, uint64_t* xo, double* ao) {
  const uint64_t m = 1ull << q;
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t xi = x[i];
    const double ar = a[2 * i], ai = a[2 * i + 1];
    const int b = (int)((xi >> q) & 1ull);
    double r0, i0, r1, i1;
    cmul(G[2 * (0 * 2 + b)], G[2 * (0 * 2 + b) + 1], ar, ai, r0, i0);
    cmul(G[2 * (1 * 2 + b)], G[2 * (1 * 2 + b) + 1], ar, ai, r1, i1);
    const uint64_t base = xi & ~m;
    if (r0 != 0.0 || i0 != 0.0) {
      xo[k] = base;
      ao[2 * k] = r0;
      ao[2 * k + 1] = i0;
      ++k;
    }
    if (r1 != 0.0 || i1 != 0.0) {
      xo[k] = base | m;
      ao[2 * k] = r1;
      ao[2 * k + 1] = i1;
      ++k;
    }
  }
  return k;
}

void radix_sort(uint64_t* keys, double* vals, int64_t n, std::vector<uint64_t>& tk, std::vector<double>& tv) {
  tk.resize(n);
  tv.resize(2 * n);
  std::vector<int64_t> cnt(65536);
  for (int shift = 0; shift < 64; shift += 16) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < n; ++i) ++cnt[(keys[i] >> shift) & 0xFFFFull];
    int64_t run = 0;
    for (int b = 0; b < 65536; ++b) {
      const int64_t c = cnt[b];
      cnt[b] = run;
      run += c;
    }
    for (int64_t i = 0; i < n; ++i) {
      const int64_t d = cnt[(keys[i] >> shift) & 0xFFFFull]++;
      tk[d] = keys[i];
      tv[2 * d] = vals[2 * i];
      tv[2 * d + 1] = vals[2 * i + 1];
    }
    std::memcpy(keys, tk.data(), sizeof(uint64_t) * n);
    std::memcpy(vals, tv.data(), sizeof(double) * 2 * n);
  }
}

int64_t merge(uint64_t* x, double* a, int64_t n) {
  if (n == 0) return 0;
  int64_t w = 0;
  for (int64_t r = 1; r < n; ++r) {
    if (x[r] == x[w]) {
      a[2 * w] += a[2 * r];
      a[2 * w + 1] += a[2 * r + 1];
    } else {
      ++w;
      x[w] = x[r];
      a[2 * w] = a[2 * r];
      a[2 * w + 1] = a[2 * r + 1];
    }
  }
  return w + 1;
}

}  // namespace

extern "C" {

// One 1-qubit gate on (x, a) of n entries (a: re, im interleaved), then the
// truncation to k (k <= 0: none) and renormalisation.  x / a need room for 2n
// entries; tx / ta are scratch of 2n entries.  Returns the new count; *gamma_factor
// receives kept / total (1 without truncation), *merged (if not null) the count
// after the merge, before the truncation.
int64_t cpu_step_1q(uint64_t* x, double* a, int64_t n, int q, const double* G, int64_t k, uint64_t* tx, double* ta,
                    double* gamma_factor, int64_t* merged) {
  static thread_local std::vector<uint64_t> sk;
  static thread_local std::vector<double> sv;
  int64_t m = emit(x, a, n, q, G, tx, ta);
  radix_sort(tx, ta, m, sk, sv);
  m = merge(tx, ta, m);
  if (merged) *merged = m;
  *gamma_factor = 1.0;
  if (k > 0 && m > k) {
    std::vector<double> p(m);
    double total = 0.0;
    for (int64_t i = 0; i < m; ++i) {
      p[i] = ta[2 * i] * ta[2 * i] + ta[2 * i + 1] * ta[2 * i + 1];
      total += p[i];
    }
    std::vector<int64_t> idx(m);
    for (int64_t i = 0; i < m; ++i) idx[i] = i;
    std::nth_element(idx.begin(), idx.begin() + (k - 1), idx.end(),
                     [&](int64_t u, int64_t v) { return p[u] > p[v] || (p[u] == p[v] && tx[u] < tx[v]); });
    double kept = 0.0;
    for (int64_t j = 0; j < k; ++j) {
      const int64_t s = idx[j];
      x[j] = tx[s];
      a[2 * j] = ta[2 * s];
      a[2 * j + 1] = ta[2 * s + 1];
      kept += p[s];
    }
    *gamma_factor = total > 1e-30 ? kept / total : 0.0;
    m = k;
  } else {
    std::memcpy(x, tx, sizeof(uint64_t) * m);
    std::memcpy(a, ta, sizeof(double) * 2 * m);
  }
  double s2 = 0.0;
  for (int64_t i = 0; i < m; ++i) s2 += a[2 * i] * a[2 * i] + a[2 * i + 1] * a[2 * i + 1];
  const double norm = std::sqrt(s2);
  if (norm > 1e-30) {
    const double inv = 1.0 / norm;
    for (int64_t i = 0; i < 2 * m; ++i) a[i] *= inv;
  }
  return m;
}

// The gate part of cpu_step_1q alone: emit, sort and merge of (x, a) of n
// entries into (xo, ao) (room for 2n entries each), no truncation and no
// renormalisation.  Returns the merged count.
int64_t cpu_gate_1q(const uint64_t* x, const double* a, int64_t n, int q, const double* G, uint64_t* xo, double* ao) {
  static thread_local std::vector<uint64_t> sk;
  static thread_local std::vector<double> sv;
  const int64_t m = emit(x, a, n, q, G, xo, ao);
  radix_sort(xo, ao, m, sk, sv);
  return merge(xo, ao, m);
}

// The reference's radix_sort_uint64 alone (src/core/sparse_state.py:150-172),
// in place on (keys, a: re, im interleaved).
void cpu_radix_sort(uint64_t* keys, double* a, int64_t n) {
  static thread_local std::vector<uint64_t> sk;
  static thread_local std::vector<double> sv;
  radix_sort(keys, a, n, sk, sv);
}

}  // extern "C"
