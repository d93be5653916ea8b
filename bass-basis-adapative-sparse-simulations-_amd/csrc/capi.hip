// extern "C" entry points (declared in include/bass_hip.h) and workspace layout.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/bass_hip.h"
#include "common.h"
#include "kernels.h"

namespace bass {

__global__ void k_fill_u64(unsigned long long* p, int64_t words, unsigned long long v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = v;
}

static void fill(void* p, size_t bytes, unsigned long long v, hipStream_t s) {
  const int64_t words = (int64_t)(bytes / 8);
  if (words <= 0) return;
  int64_t blocks = (words + NT - 1) / NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(NT), 0, s, (unsigned long long*)p, words, v);
}

// ----------------------------------------------------------- workspace ----
// Everything is laid out from S = the maximum entry count of any input or
// output of a call.  bytes(S) = FIXED + PER * S bounds the layout below.
constexpr size_t WS_FIXED = (size_t)8 << 20;
constexpr size_t WS_PER = 64;

static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
  int64_t S;
  // header A (zeroed by every gate)
  size_t hist1, tick_gate, nout;
  size_t hdrA_end;
  // header B (zeroed by every truncate / auxiliary call)
  size_t hist2, sel, tick_sel, hist_w, hdrB_end;
  // persistent
  size_t st, tile_psum, tile_dsum, tile_s2, tile_s4, rdm_part, rho, sort_hist, sort_tick;
  size_t lb_gate, lb1, sort_status;
  size_t seg_cnt, soff, cnt1, tile_dcnt, ksum, kcnt, koff, lift_b, hist2_rows;
  size_t big;
  size_t total;
};

static Layout layout(int64_t S) {
  Layout L;
  L.S = S;
  const int64_t etiles = S / 512 + 2;   // emission tiles hold >= 512 entries
  const int64_t stiles = S / 2048 + 2;  // select / sort tiles hold 2048 entries
  size_t o = 0;
  L.hist1 = o; o += al(HIST_BINS * 4);
  L.tick_gate = o; o += al(16 * 4);
  L.nout = o; o += al(8);
  L.hdrA_end = o;
  L.hist2 = o; o += al(HIST_BINS * 4);
  L.sel = o; o += al(sizeof(SelState));
  L.tick_sel = o; o += al(16 * 4);
  L.hist_w = o; o += al((size_t)WIDE_LEVELS * HIST_BINS * 4);
  L.hdrB_end = o;
  L.st = o; o += al(sizeof(OpStats));
  L.tile_psum = o; o += al(etiles * 8);
  L.tile_dsum = o; o += al(stiles * 8);
  L.tile_s2 = o; o += al(stiles * 8);
  L.tile_s4 = o; o += al(stiles * 8);
  L.rdm_part = o; o += al((size_t)1024 * 64 * 4 * 8);
  L.rho = o; o += al(64 * 4 * 8);
  L.sort_hist = o; o += al(8 * 256 * 4);
  L.sort_tick = o; o += al(16 * 4);
  L.lb_gate = o; o += al(etiles * 8);
  L.lb1 = o; o += al(stiles * 8);
  L.sort_status = o; o += al((size_t)stiles * 256 * 4);
  L.seg_cnt = o; o += al(etiles * 4);
  L.soff = o; o += al(etiles * 8);
  L.cnt1 = o; o += al(stiles * 4);
  L.tile_dcnt = o; o += al(stiles * 4);
  L.ksum = o; o += al(stiles * 8);
  L.kcnt = o; o += al(stiles * 4);
  L.koff = o; o += al(stiles * 8);
  L.lift_b = o; o += al((etiles + 1) * sizeof(LiftBound));
  L.hist2_rows = o; o += al((size_t)stiles * HIST_BINS * 2);
  L.big = o;
  // big scratch union:
  //   gate group table: cap <= 4 n_in slots x <= 32 B, n_in <= S/4 (2q) or S/2 (1q) -> <= 32 S
  //   select candidates: 24 B per slot of (S + one tile);   rdm / ht table: <= 4 S slots x 12 B = 48 S
  //   radix sort ping-pong: 24 S
  o += al((size_t)48 * S + ((size_t)1 << 17));
  L.total = o;
  return L;
}

static int64_t ws_capacity(size_t ws_bytes) {
  if (ws_bytes < WS_FIXED) return -1;
  int64_t S = (int64_t)((ws_bytes - WS_FIXED) / WS_PER);
  while (S > 0 && layout(S).total > ws_bytes) --S;
  return S;
}

template <typename T>
static T* at(void* ws, size_t off) {
  return (T*)((char*)ws + off);
}

static int log2_ceil(int64_t v) {
  int l = 0;
  while ((1ll << l) < v) ++l;
  return l;
}

static SelTiles tiles_view(void* ws, const Layout& L, unsigned long long* c1p, unsigned int* c1i) {
  SelTiles T;
  T.tile_psum = at<double>(ws, L.tile_psum);
  T.c1p = c1p;
  T.c1i = c1i;
  T.cnt1 = at<int>(ws, L.cnt1);
  T.tile_dsum = at<double>(ws, L.tile_dsum);
  T.tile_dcnt = at<int>(ws, L.tile_dcnt);
  T.ksum = at<double>(ws, L.ksum);
  T.kcnt = at<int>(ws, L.kcnt);
  T.koff = at<int64_t>(ws, L.koff);
  T.tile_s2 = at<double>(ws, L.tile_s2);
  T.tile_s4 = at<double>(ws, L.tile_s4);
  T.hist2_rows = at<unsigned short>(ws, L.hist2_rows);
  return T;
}

// ------------------------------------------------------------- timing ----
struct TimerSlot {
  std::string name;
  int64_t count = 0;
  double total_ms = 0.0;
};
struct PendingEv {
  int slot;
  hipEvent_t a, b;
};
static std::mutex g_tmu;
static bool g_timing = false;
static std::vector<TimerSlot> g_slots;
static std::vector<PendingEv> g_pending;
static std::vector<hipEvent_t> g_evpool;

static hipEvent_t ev_take() {
  if (!g_evpool.empty()) {
    hipEvent_t e = g_evpool.back();
    g_evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

int ktimer_begin(const char* name, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_tmu);
  if (!g_timing) return -1;
  int slot = -1;
  for (size_t i = 0; i < g_slots.size(); ++i)
    if (g_slots[i].name == name) slot = (int)i;
  if (slot < 0) {
    g_slots.push_back(TimerSlot{name, 0, 0.0});
    slot = (int)g_slots.size() - 1;
  }
  PendingEv p{slot, ev_take(), ev_take()};
  (void)hipEventRecord(p.a, s);
  g_pending.push_back(p);
  return (int)g_pending.size() - 1;
}

void ktimer_end(int token, hipStream_t s) {
  if (token < 0) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  (void)hipEventRecord(g_pending[token].b, s);
}

static void timer_drain() {
  for (auto& p : g_pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    g_slots[p.slot].count += 1;
    g_slots[p.slot].total_ms += ms;
    g_evpool.push_back(p.a);
    g_evpool.push_back(p.b);
  }
  g_pending.clear();
}

}  // namespace bass

using namespace bass;

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

static int hip_check(const char* where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(BASS_E_HIP, "%s: %s", where, hipGetErrorString(e));
  return BASS_OK;
}

#define WS_REQUIRE(need)                                                                       \
  int64_t S = ws_capacity(ws_bytes);                                                           \
  if (!ws || S < (need))                                                                       \
    return fail(BASS_E_WORKSPACE, "workspace holds %lld entries, call needs %lld", (long long)S, \
                (long long)(need));                                                            \
  const Layout L = layout(S)

extern "C" {

int bass_abi_version(void) { return 2; }

const char* bass_last_error(void) { return g_err; }

int bass_timing(int enable) {
  std::lock_guard<std::mutex> lk(g_tmu);
  timer_drain();
  g_slots.clear();
  g_timing = enable != 0;
  return BASS_OK;
}

int bass_timing_read(const char* name, int64_t* count, double* total_ms) {
  std::lock_guard<std::mutex> lk(g_tmu);
  timer_drain();
  *count = 0;
  *total_ms = 0.0;
  for (auto& s : g_slots)
    if (s.name == name) {
      *count = s.count;
      *total_ms = s.total_ms;
    }
  return BASS_OK;
}

size_t bass_workspace_bytes(int64_t max_entries) {
  if (max_entries < 1) max_entries = 1;
  const size_t b = WS_FIXED + WS_PER * (size_t)max_entries;
  return b;
}

int bass_read_stats(void* ws, size_t ws_bytes, bass_stats* out, void* stream) {
  WS_REQUIRE(0);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  if (hipMemcpyAsync(out, st, sizeof(bass_stats), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_read_stats: %s", hipGetErrorString(hipGetLastError()));
  if (out->err) {
    fill(&st->err, 8, 0ull, s);
    return fail(BASS_E_DEVICE, "device error bits 0x%x", out->err);
  }
  return BASS_OK;
}

int bass_gate(const uint64_t* x, const double* a, int64_t n, int nq, int q0, int q1, const double* G,
              int mode, int dedup, int segmented, uint64_t* xo, double* ao, int64_t cap_out,
              int64_t* hist_tiles, void* ws, size_t ws_bytes, void* stream) {
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_gate: nq must be 1 or 2");
  if (q0 < 0 || q0 > 62 || (nq == 2 && (q1 < 0 || q1 > 62 || q1 == q0)))
    return fail(BASS_E_ARG, "bass_gate: bad qubits %d %d", q0, q1);
  if (n < 0 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_gate: bad n %lld", (long long)n);
  if (mode == MODE_ROT && !dedup) return fail(BASS_E_ARG, "bass_gate: rotation requires dedup");
  const int M = 1 << nq;
  const int64_t tiles = n > 0 ? (n + emit_tile_entries(nq) - 1) / emit_tile_entries(nq) : 0;
  const int64_t need_out = segmented ? tiles * SEG_ENTRIES : (int64_t)M * n;
  if (segmented && cap_out < need_out)
    return fail(BASS_E_OVERFLOW, "bass_gate: segmented output needs %lld slots", (long long)need_out);
  WS_REQUIRE(need_out > n ? need_out : n);
  hipStream_t s = (hipStream_t)stream;
  if (hist_tiles) *hist_tiles = 0;
  OpStats* st = at<OpStats>(ws, L.st);
  static const char* names[2][2][2] = {{{"emit1q", "emit2q"}, {"gate1q", "gate2q"}},
                                       {{"rot1q?", "rot2q?"}, {"rot1q", "rot2q"}}};
  const int tk = ktimer_begin(names[mode == MODE_ROT][dedup != 0][nq - 1], s);
  // zero header A and (dense output) the gate's look-back status; stats nnz/sums
  fill(ws, L.hdrA_end, 0ull, s);
  if (!segmented) fill(at<void>(ws, L.lb_gate), (size_t)(tiles + 1) * 8, 0ull, s);
  fill(st, 48, 0ull, s);  // nnz, total, kept, norm, sum2, sum4 (err untouched)
  GateMat gm;
  memset(&gm, 0, sizeof(gm));
  for (int j = 0; j < M * M; ++j) gm.m[j] = make_double2(G[2 * j], G[2 * j + 1]);
  int log2cap = 0;
  uint64_t* gslot = nullptr;
  if (dedup && n > 0) {
    log2cap = log2_ceil(2 * n);
    if (log2cap < 6) log2cap = 6;
    const size_t cap = (size_t)1 << log2cap;
    const size_t bytes = cap * (size_t)group_slot_bytes(nq);
    if (bytes > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_gate: group table does not fit");
    gslot = at<uint64_t>(ws, L.big);
    fill(gslot, bytes, ~0ull, s);
    launch_group_build(nq, x, n, q0, q1, gslot, log2cap, &st->err, s);
  }
  int* seg = segmented ? at<int>(ws, L.seg_cnt) : nullptr;
  if (n > 0) {
    launch_emit(nq, mode, dedup != 0, x, (const double2*)a, n, q0, nq == 2 ? q1 : 0, gslot, log2cap, gm, xo,
                (double2*)ao, cap_out, at<unsigned long long>(ws, L.lb_gate), at<unsigned int>(ws, L.tick_gate),
                at<unsigned int>(ws, L.hist1), at<double>(ws, L.tile_psum), &st->nnz, seg, &st->err, s);
    if (segmented) launch_seg_scan(seg, tiles, at<int64_t>(ws, L.soff), st, s);
    if (hist_tiles) *hist_tiles = tiles;
  }
  ktimer_end(tk, s);
  return hip_check("bass_gate");
}

int bass_lift(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
              uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws, size_t ws_bytes,
              void* stream) {
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_lift: nq must be 1 or 2");
  if (mode != MODE_GATE && mode != MODE_ROT && mode != MODE_MOVE) return fail(BASS_E_ARG, "bass_lift: bad mode");
  if (mode == MODE_MOVE && nq != 1) return fail(BASS_E_ARG, "bass_lift: a move lifts one bit");
  if (j < nq - 1 || j > 62) return fail(BASS_E_ARG, "bass_lift: bad bit %d for %d qubit(s)", j, nq);
  if (n < 0 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_lift: bad n %lld", (long long)n);
  const int tm = lift_tile_entries(nq, mode);
  const int64_t tiles = n > 0 ? (n + tm - 1) / tm : 0;
  const int64_t need_out = mode == MODE_MOVE ? n : tiles * SEG_ENTRIES;
  if (cap_out < need_out)
    return fail(BASS_E_OVERFLOW, "bass_lift: output needs %lld slots", (long long)need_out);
  WS_REQUIRE(need_out > n ? need_out : n);
  hipStream_t s = (hipStream_t)stream;
  if (seg_tiles) *seg_tiles = mode == MODE_MOVE ? 0 : tiles;
  OpStats* st = at<OpStats>(ws, L.st);
  static const char* names[3][2] = {{"gate1q", "gate2q"}, {"rot1q", "rot2q"}, {"move", "move"}};
  const int tk = ktimer_begin(names[mode][nq - 1], s);
  GateMat gm;
  memset(&gm, 0, sizeof(gm));
  if (mode != MODE_MOVE) {
    const int M = 1 << nq;
    for (int e = 0; e < M * M; ++e) gm.m[e] = make_double2(G[2 * e], G[2 * e + 1]);
    fill(at<void>(ws, L.hist1), HIST_BINS * 4, 0ull, s);
  }
  fill(st, 48, 0ull, s);
  if (n > 0) {
    int* seg = at<int>(ws, L.seg_cnt);
    launch_lift(nq, mode, x, (const double2*)a, n, j, gm, xo, (double2*)ao, at<LiftBound>(ws, L.lift_b),
                at<unsigned int>(ws, L.hist1), at<double>(ws, L.tile_psum), seg, s);
    if (mode != MODE_MOVE) launch_seg_scan(seg, tiles, at<int64_t>(ws, L.soff), st, s);
  }
  ktimer_end(tk, s);
  return hip_check("bass_lift");
}

int bass_densify(const uint64_t* x, const double* a, int64_t seg_tiles, uint64_t* xo, double* ao,
                 int64_t cap_out, void* ws, size_t ws_bytes, void* stream) {
  if (seg_tiles < 0) return fail(BASS_E_ARG, "bass_densify: bad segment count");
  WS_REQUIRE(seg_tiles * SEG_ENTRIES);
  (void)cap_out;
  launch_densify(x, (const double2*)a, at<int>(ws, L.seg_cnt), at<int64_t>(ws, L.soff), seg_tiles, xo,
                 (double2*)ao, (hipStream_t)stream);
  return hip_check("bass_densify");
}

int bass_truncate(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo, double* ao,
                  int64_t cap_out, int renorm, int64_t hist_tiles, int in_flags, const int32_t* key_map,
                  int key_bits, void* ws, size_t ws_bytes, void* stream) {
  const bool segmented = (in_flags & 1) != 0;
  const bool wide = (in_flags & 2) != 0;
  if (k < 1 || (!segmented && n <= k))
    return fail(BASS_E_ARG, "bass_truncate: need n > k >= 1 (n=%lld k=%lld)", (long long)n, (long long)k);
  if (cap_out < k) return fail(BASS_E_OVERFLOW, "bass_truncate: cap_out < k");
  if (n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_truncate: n too large");
  if (segmented && hist_tiles <= 0) return fail(BASS_E_ARG, "bass_truncate: segmented input needs hist_tiles");
  KeyMap km;
  memset(&km, 0, sizeof(km));
  if (key_map) {
    if (key_bits < 1 || key_bits > 63) return fail(BASS_E_ARG, "bass_truncate: key_bits out of range");
    uint64_t seen = 0;
    for (int p = 0; p < key_bits; ++p) {
      const int q = key_map[p];
      if (q < 0 || q >= key_bits || ((seen >> q) & 1ull))
        return fail(BASS_E_ARG, "bass_truncate: key_map is not a permutation");
      seen |= 1ull << q;
      km.q[p] = (unsigned char)q;
    }
    km.nbits = key_bits;
  }
  const int64_t tiles = segmented ? hist_tiles : sel_tiles(n);
  const int64_t slots = tiles * SEG_ENTRIES;
  WS_REQUIRE(slots > n ? slots : n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  fill(at<void>(ws, L.hist2), L.hdrB_end - L.hist2, 0ull, s);
  if (hist_tiles <= 0) fill(at<void>(ws, L.hist1), HIST_BINS * 4, 0ull, s);
  char* big = at<char>(ws, L.big);
  const size_t p8 = al((size_t)8 * slots), p4 = al((size_t)4 * slots);
  if (3 * p8 + 2 * p4 > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_truncate: candidates do not fit");
  SelProblem P;
  memset(&P, 0, sizeof(P));
  P.x = x;
  P.a = (const double2*)a;
  P.seg_cnt = segmented ? at<int>(ws, L.seg_cnt) : nullptr;
  P.n_dense = n;
  P.tile0 = 0;
  P.ntiles = tiles;
  P.row0 = 0;
  P.nrows = select_rows(tiles, false);
  P.npsum = hist_tiles > 0 ? hist_tiles : tiles;
  P.k = (unsigned long long)k;
  P.xo = xo;
  P.ao = (double2*)ao;
  P.cap_out = cap_out;
  P.renorm = renorm;
  P.hist1 = at<unsigned int>(ws, L.hist1);
  P.hist2 = at<unsigned int>(ws, L.hist2);
  P.hist_w = at<unsigned int>(ws, L.hist_w);
  P.sel = at<SelState>(ws, L.sel);
  P.st = st;
  P.c2p = (unsigned long long*)(big + p8 + p4);
  P.c2i = (unsigned int*)(big + 2 * p8 + p4);
  P.tk = (uint64_t*)(big + 2 * p8 + 2 * p4);
  P.km = km;
  const SelTiles T = tiles_view(ws, L, (unsigned long long*)big, (unsigned int*)(big + p8));
  const int tk = ktimer_begin("select", s);
  launch_select(P, T, hist_tiles > 0, wide, s);
  ktimer_end(tk, s);
  return hip_check("bass_truncate");
}

int bass_diag2q(const uint64_t* x, double* a, int64_t n, int q1, int q2, const double* diag, void* stream) {
  if (q1 < 0 || q1 > 62 || q2 < 0 || q2 > 62 || q1 == q2) return fail(BASS_E_ARG, "bass_diag2q: bad qubits");
  double2 d[4];
  for (int j = 0; j < 4; ++j) d[j] = make_double2(diag[2 * j], diag[2 * j + 1]);
  const int tk = ktimer_begin("diag2q", (hipStream_t)stream);
  launch_diag2q(x, (double2*)a, n, q1, q2, d, (hipStream_t)stream);
  ktimer_end(tk, (hipStream_t)stream);
  return hip_check("bass_diag2q");
}

int bass_moments(const double* a, int64_t n, void* ws, size_t ws_bytes, void* stream) {
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  launch_moments((const double2*)a, n, at<double>(ws, L.tile_s2), at<double>(ws, L.tile_s4), st, s);
  return hip_check("bass_moments");
}

int bass_scale(double* a, int64_t n, double div, void* stream) {
  launch_scale((double2*)a, n, div, (hipStream_t)stream);
  return hip_check("bass_scale");
}

int bass_gather(const uint64_t* x, const double* a, const int64_t* idx, int64_t m, uint64_t* xo, double* ao,
                void* stream) {
  launch_gather(x, (const double2*)a, idx, m, xo, (double2*)ao, (hipStream_t)stream);
  return hip_check("bass_gather");
}

int bass_rdm1q(const uint64_t* x, const double* a, int64_t n, int N, double* rho_out, void* ws,
               size_t ws_bytes, void* stream) {
  if (N < 1 || N > 63) return fail(BASS_E_ARG, "bass_rdm1q: N out of range");
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  double* rho = at<double>(ws, L.rho);
  if (n <= 0) {
    fill(rho, (size_t)N * 4 * 8, 0ull, s);
  } else {
    int log2cap = log2_ceil(2 * n);
    if (log2cap < 6) log2cap = 6;
    const size_t cap = (size_t)1 << log2cap;
    if (cap * 12 > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_rdm1q: table does not fit");
    uint64_t* tk = at<uint64_t>(ws, L.big);
    int32_t* tv = at<int32_t>(ws, L.big + cap * 8);
    const int tm = ktimer_begin("rdm", s);
    fill(tk, cap * 8, ~0ull, s);
    launch_table_build(x, n, tk, tv, log2cap, 0, &st->err, s);
    launch_rdm(x, (const double2*)a, n, N, tk, tv, log2cap, at<double>(ws, L.rdm_part), rho, s);
    ktimer_end(tm, s);
  }
  int rc = hip_check("bass_rdm1q");
  if (rc) return rc;
  if (hipMemcpyAsync(rho_out, rho, (size_t)N * 4 * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_rdm1q copy: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

int bass_rdm1q_ordered(const uint64_t* x, const double* a, int64_t n, int N, double* rho_out, void* ws,
                       size_t ws_bytes, void* stream) {
  if (N < 1 || N > 63) return fail(BASS_E_ARG, "bass_rdm1q_ordered: N out of range");
  if (n < 0 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_rdm1q_ordered: bad n");
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  double* rho = at<double>(ws, L.rho);
  if (n <= 0) {
    fill(rho, (size_t)N * 4 * 8, 0ull, s);
  } else {
    if (rdm_ordered_scratch(n, N) > L.total - L.big)
      return fail(BASS_E_WORKSPACE, "bass_rdm1q_ordered: scratch does not fit");
    const int tm = ktimer_begin("rdm", s);
    launch_rdm_ordered(x, (const double2*)a, n, N, at<void>(ws, L.big), rho, s);
    ktimer_end(tm, s);
  }
  int rc = hip_check("bass_rdm1q_ordered");
  if (rc) return rc;
  if (hipMemcpyAsync(rho_out, rho, (size_t)N * 4 * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_rdm1q_ordered copy: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

int bass_ht_clear(uint64_t* ht_keys, int32_t* ht_vals, int log2cap, void* stream) {
  if (log2cap < 1 || log2cap > 40) return fail(BASS_E_ARG, "bass_ht_clear: bad capacity");
  hipStream_t s = (hipStream_t)stream;
  fill(ht_keys, ((size_t)1 << log2cap) * 8, ~0ull, s);
  (void)ht_vals;
  return hip_check("bass_ht_clear");
}

int bass_ht_insert(uint64_t* ht_keys, int32_t* ht_vals, int log2cap, const uint64_t* keys, int64_t n,
                   int64_t index_base, void* ws, size_t ws_bytes, void* stream) {
  if (log2cap < 1 || log2cap > 40) return fail(BASS_E_ARG, "bass_ht_insert: bad capacity");
  if (index_base != 0) return fail(BASS_E_ARG, "bass_ht_insert: index_base must be 0");
  WS_REQUIRE(0);
  OpStats* st = at<OpStats>(ws, L.st);
  launch_table_build(keys, n, ht_keys, ht_vals, log2cap, 1, &st->err, (hipStream_t)stream);
  return hip_check("bass_ht_insert");
}

int bass_ht_lookup(const uint64_t* ht_keys, const int32_t* ht_vals, int log2cap, const uint64_t* keys,
                   int64_t n, int64_t* out, void* stream) {
  if (log2cap < 1 || log2cap > 40) return fail(BASS_E_ARG, "bass_ht_lookup: bad capacity");
  launch_table_lookup(ht_keys, ht_vals, log2cap, keys, n, out, (hipStream_t)stream);
  return hip_check("bass_ht_lookup");
}

int bass_radix_sort(uint64_t* keys, double* vals, int64_t n, int key_bits, void* ws, size_t ws_bytes,
                    void* stream) {
  if (key_bits < 1 || key_bits > 64) return fail(BASS_E_ARG, "bass_radix_sort: key_bits out of range");
  if (n <= 1) return BASS_OK;
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  char* big = at<char>(ws, L.big);
  uint64_t* tk = (uint64_t*)big;
  double2* tv = (double2*)(big + al((size_t)8 * n));
  const int tm = ktimer_begin("sort", s);
  launch_radix_sort(keys, (double2*)vals, n, key_bits, tk, tv, at<unsigned int>(ws, L.sort_hist),
                    at<unsigned int>(ws, L.sort_status), at<unsigned int>(ws, L.sort_tick), &st->err, s);
  ktimer_end(tm, s);
  return hip_check("bass_radix_sort");
}

int bass_merge_sorted(const uint64_t* x, const double* a, int64_t n, uint64_t* xo, double* ao, void* ws,
                      size_t ws_bytes, void* stream) {
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  fill(st, 8, 0ull, s);
  if (n <= 0) return hip_check("bass_merge_sorted");
  const int64_t tiles = sel_tiles(n);
  fill(at<void>(ws, L.lb1), (size_t)(tiles + 1) * 8, 0ull, s);
  fill(at<void>(ws, L.tick_sel), 64, 0ull, s);
  launch_merge(x, (const double2*)a, n, xo, (double2*)ao, at<unsigned long long>(ws, L.lb1),
               at<unsigned int>(ws, L.tick_sel), &st->nnz, &st->err, s);
  return hip_check("bass_merge_sorted");
}

}  // extern "C"
