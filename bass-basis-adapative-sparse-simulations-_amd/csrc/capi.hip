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

// Zero two regions (8-byte multiples) in one launch.
__global__ void k_zero2(unsigned long long* p, int64_t wp, unsigned long long* q, int64_t wq) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < wp + wq; i += stride) {
    if (i < wp)
      p[i] = 0ull;
    else
      q[i - wp] = 0ull;
  }
}

static void zero2(void* p, size_t pbytes, void* q, size_t qbytes, hipStream_t s) {
  const int64_t wp = (int64_t)(pbytes / 8), wq = (int64_t)(qbytes / 8);
  int64_t blocks = (wp + wq + NT - 1) / NT;
  if (blocks > 64) blocks = 64;
  if (blocks < 1) return;
  hipLaunchKernelGGL(k_zero2, dim3((unsigned)blocks), dim3(NT), 0, s, (unsigned long long*)p, wp,
                     (unsigned long long*)q, wq);
}

static void fill(void* p, size_t bytes, unsigned long long v, hipStream_t s) {
  const int64_t words = (int64_t)(bytes / 8);
  if (words <= 0) return;
  int64_t blocks = (words + NT - 1) / NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(NT), 0, s, (unsigned long long*)p, words, v);
}

// Counted mode (bass_lift_counted / bass_truncate flag 4): entries of the op
// log, one per truncation, read back by bass_counted_read.
constexpr int64_t OPLOG_CAP = 4096;

// Append the stats block to the op log and set the state count slot to the
// operation's output count (one thread; stream order serialises appends).
__global__ void k_oplog_append(const OpStats* st, unsigned long long* log_n, OpStats* log, int64_t* count) {
  const unsigned long long i = *log_n;
  if (i < (unsigned long long)OPLOG_CAP) log[i] = *st;
  *log_n = i + 1;
  *count = st->nnz;
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
  size_t seg_cnt, soff, cnt1, tile_dcnt, ksum, kcnt, koff, lift_b, hist2_rows, chunks;
  size_t count_slot, oplog;  // counted mode: state entry count; op log (counter + entries)
  size_t ktab;               // key-map byte tables of the select (KTAB_WORDS words)
  size_t spec;               // digit-1 bucket of the last select: a speculative lift's prediction
  size_t big;
  size_t total;
};

// Entries of each per-chunk array: chunks of the largest tile count a call
// reduces (emission tiles, >= 512 output slots each).
static size_t chunk_slots(int64_t S) { return (size_t)sel_chunks(S / 512 + 2) + 1; }

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
  L.sort_hist = o; o += al(SORT_MAX_PASSES * 1024 * 4);
  L.sort_tick = o; o += al(64 * 4);
  L.sort_status = o; o += al(sort_status_bytes(S));
  L.lb_gate = o; o += al(etiles * 8);
  L.lb1 = o; o += al(stiles * 8);
  L.seg_cnt = o; o += al(etiles * 4);
  L.soff = o; o += al(etiles * 8);
  L.cnt1 = o; o += al(stiles * 4);
  L.tile_dcnt = o; o += al(stiles * 4);
  L.ksum = o; o += al(stiles * 8);
  L.kcnt = o; o += al(stiles * 4);
  L.koff = o; o += al(stiles * 8);
  L.lift_b = o; o += al((size_t)lift_bounds_entries(etiles + 1) * sizeof(LiftBound));
  L.hist2_rows = o; o += al((size_t)stiles * HIST_BINS * 2);
  // per-chunk values of the chunked reductions: 6 arrays for a select, 2 for
  // a gate's segment scan (read later by bass_densify); a select reading a
  // dense gate's per-tile sums spans up to etiles tiles
  L.chunks = o; o += al((size_t)8 * 8 * chunk_slots(S));
  L.count_slot = o; o += al(8);
  L.oplog = o; o += al(8 + (size_t)OPLOG_CAP * sizeof(OpStats));
  L.ktab = o; o += al((size_t)KTAB_WORDS * 8);
  L.spec = o; o += al(8);
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

// chunk counts / offsets of a gate's segment scan (after the select's six arrays)
static int64_t* seg_chunk_cnt(void* ws, const Layout& L) {
  return at<int64_t>(ws, L.chunks) + 6 * chunk_slots(L.S);
}
static int64_t* seg_chunk_off(void* ws, const Layout& L) { return seg_chunk_cnt(ws, L) + chunk_slots(L.S); }

static SelTiles tiles_view(void* ws, const Layout& L, unsigned long long* c1p, unsigned int* c1i) {
  SelTiles T;
  memset(&T, 0, sizeof(T));
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
  const size_t nc = chunk_slots(L.S);
  int64_t* ch = at<int64_t>(ws, L.chunks);
  T.ch_cnt = ch;
  T.ch_off = ch + nc;
  T.ch_sum = (double*)(ch + 2 * nc);
  T.ch_psum = (double*)(ch + 3 * nc);
  T.ch_s2 = (double*)(ch + 4 * nc);
  T.ch_s4 = (double*)(ch + 5 * nc);
  return T;
}

// Candidate arrays of a select over `slots` input slots, in the workspace's
// big region: c1p, c1i, c2p, c2i, tk, and (fused gate selects) c1k.
struct CandRegion {
  unsigned long long* c1p;
  unsigned int* c1i;
  unsigned long long* c2p;
  unsigned int* c2i;
  uint64_t* tk;
  unsigned long long* c1k;
  size_t bytes;
};
static CandRegion cand_region(void* ws, const Layout& L, int64_t slots) {
  char* big = at<char>(ws, L.big);
  const size_t p8 = al((size_t)8 * slots), p4 = al((size_t)4 * slots);
  CandRegion c;
  c.c1p = (unsigned long long*)big;
  c.c1i = (unsigned int*)(big + p8);
  c.c2p = (unsigned long long*)(big + p8 + p4);
  c.c2i = (unsigned int*)(big + 2 * p8 + p4);
  c.tk = (uint64_t*)(big + 2 * p8 + 2 * p4);
  c.c1k = (unsigned long long*)(big + 3 * p8 + 2 * p4);
  c.bytes = 4 * p8 + 2 * p4;
  return c;
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

// ------------------------------------------------------- size-class tuning ----
// Per size class of a call's input count n (class 0: n < 2^14, 1: < 2^18,
// 2: < 2^22, 3: larger), set with bass_tune (typically from the autotuner's
// table): the hash/group-table capacity factor (capacity = next power of two
// >= factor * n: factor 2 keeps the load factor <= 1/2, 4 <= 1/4), the items
// per thread of the emission kernels (the block's tile = 256 x items entries)
// and the radix digit width.  0 = the built-in choice.  Keys, amplitudes and
// kept sums do not depend on them; the emission tile size groups a dense gate
// output's input total (per-tile sums, then tiles in order), so the total may
// differ in its last bits.
enum { TUNE_TABLE_FACTOR = 0, TUNE_EMIT_ITEMS_1Q = 1, TUNE_EMIT_ITEMS_2Q = 2, TUNE_SORT_DIGITS = 3, TUNE_KNOBS = 4 };
constexpr int TUNE_CLASSES = 4;
static int g_tune[TUNE_KNOBS][TUNE_CLASSES] = {};
static std::mutex g_tune_mu;

static int tune_class(int64_t n) { return n < (1 << 14) ? 0 : n < (1 << 18) ? 1 : n < (1 << 22) ? 2 : 3; }
static int tuned(int knob, int64_t n) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tune[knob][tune_class(n)];
}
// Table capacity (log2) for n keys of slot_bytes each: the tuned factor when
// the table fits in `room` bytes of workspace, else the default factor 2.
static int table_log2cap(int64_t n, size_t slot_bytes, size_t room) {
  auto lg = [n](int f) {
    int l = 0;
    while ((1ll << l) < f * n) ++l;
    return l < 6 ? 6 : l;
  };
  const int l = lg(tuned(TUNE_TABLE_FACTOR, n) == 4 ? 4 : 2);
  return ((size_t)1 << l) * slot_bytes <= room ? l : lg(2);
}

// ---------------------------------------------------- no-allocation call ----
// bass_truncate_noalloc works in `scratch`, a static device array of this
// library (one copy per GPU, held until the process ends) laid out like a
// workspace of NOALLOC_MAX_ENTRIES entries, instead of a caller workspace.
// Host code uses the array only through the address hipGetSymbolAddress
// returns on each device (the array's name on the host is a stand-in, not a
// device address).
constexpr int64_t NOALLOC_MAX_ENTRIES = (int64_t)1 << 20;
constexpr size_t NOALLOC_SCRATCH_BYTES = bass::WS_FIXED + bass::WS_PER * (size_t)NOALLOC_MAX_ENTRIES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace bass {
__global__ void k_copy_stats(const OpStats* src, OpStats* dst) { *dst = *src; }
}  // namespace bass

// The device address of `scratch` on the current device (looked up once per
// device, then cached).
static int noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return fail(BASS_E_HIP, "bass_truncate_noalloc: no current device");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BASS_E_HIP, "bass_truncate_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    if (bass::ws_capacity(NOALLOC_SCRATCH_BYTES) < NOALLOC_MAX_ENTRIES)
      return fail(BASS_E_WORKSPACE, "bass_truncate_noalloc: scratch layout does not hold %lld entries",
                  (long long)NOALLOC_MAX_ENTRIES);
    addr[dev] = p;
  }
  *out = addr[dev];
  return BASS_OK;
}

static int parse_key_map(const int32_t* key_map, int key_bits, KeyMap& km, const char* who) {
  memset(&km, 0, sizeof(km));
  if (!key_map) return BASS_OK;
  if (key_bits < 1 || key_bits > 63) return fail(BASS_E_ARG, "%s: key_bits out of range", who);
  uint64_t seen = 0;
  for (int p = 0; p < key_bits; ++p) {
    const int q = key_map[p];
    if (q < 0 || q >= key_bits || ((seen >> q) & 1ull)) return fail(BASS_E_ARG, "%s: key_map is not a permutation", who);
    seen |= 1ull << q;
    km.q[p] = (unsigned char)q;
  }
  km.nbits = key_bits;
  return BASS_OK;
}

#define WS_REQUIRE(need)                                                                       \
  int64_t S = ws_capacity(ws_bytes);                                                           \
  if (!ws || S < (need))                                                                       \
    return fail(BASS_E_WORKSPACE, "workspace holds %lld entries, call needs %lld", (long long)S, \
                (long long)(need));                                                            \
  const Layout L = layout(S)

extern "C" {

int bass_abi_version(void) { return 3; }

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
  // copy through a pinned per-thread buffer: a device-to-pageable copy is
  // staged by the runtime and costs a second hop before the host sees it
  static thread_local bass_stats* pinned = nullptr;
  if (!pinned && hipHostMalloc((void**)&pinned, sizeof(bass_stats), hipHostMallocDefault) != hipSuccess) {
    pinned = nullptr;
    return fail(BASS_E_HIP, "bass_read_stats: hipHostMalloc: %s", hipGetErrorString(hipGetLastError()));
  }
  if (hipMemcpyAsync(pinned, st, sizeof(bass_stats), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_read_stats: %s", hipGetErrorString(hipGetLastError()));
  memcpy(out, pinned, sizeof(bass_stats));
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
  // tuned tile size for dense outputs; a segmented output keeps the default
  const int items = segmented ? 0 : tuned(nq == 1 ? TUNE_EMIT_ITEMS_1Q : TUNE_EMIT_ITEMS_2Q, n);
  const int64_t tiles = n > 0 ? (n + emit_tile_entries(nq, items) - 1) / emit_tile_entries(nq, items) : 0;
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
    log2cap = table_log2cap(n, (size_t)group_slot_bytes(nq), L.total - L.big);
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
                at<unsigned int>(ws, L.hist1), at<double>(ws, L.tile_psum), &st->nnz, seg, &st->err, s, items);
    if (segmented)
      launch_seg_scan(seg, tiles, at<int64_t>(ws, L.soff), seg_chunk_cnt(ws, L), seg_chunk_off(ws, L), st, s);
    if (hist_tiles) *hist_tiles = tiles;
  }
  ktimer_end(tk, s);
  return hip_check("bass_gate");
}

static int lift_impl(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                     uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws, size_t ws_bytes,
                     void* stream, bool counted);

int bass_lift(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
              uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws, size_t ws_bytes,
              void* stream) {
  return lift_impl(x, a, n, nq, j, G, mode, xo, ao, cap_out, seg_tiles, ws, ws_bytes, stream, false);
}

int bass_lift_counted(const uint64_t* x, const double* a, int64_t n_max, int nq, int j, const double* G, int mode,
                      uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws, size_t ws_bytes,
                      void* stream) {
  return lift_impl(x, a, n_max, nq, j, G, mode, xo, ao, cap_out, seg_tiles, ws, ws_bytes, stream, true);
}

int bass_diag2q_counted(const uint64_t* x, double* a, int64_t n_max, int q1, int q2, const double* diag, void* ws,
                        size_t ws_bytes, void* stream) {
  if (q1 < 0 || q1 > 62 || q2 < 0 || q2 > 62 || q1 == q2) return fail(BASS_E_ARG, "bass_diag2q_counted: bad qubits");
  if (n_max < 0 || n_max > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_diag2q_counted: bad n %lld", (long long)n_max);
  WS_REQUIRE(0);
  double2 d[4];
  for (int e = 0; e < 4; ++e) d[e] = make_double2(diag[2 * e], diag[2 * e + 1]);
  const int tk = ktimer_begin("diag2q", (hipStream_t)stream);
  launch_diag2q(x, (double2*)a, n_max, q1, q2, d, (hipStream_t)stream, at<int64_t>(ws, L.count_slot));
  ktimer_end(tk, (hipStream_t)stream);
  return hip_check("bass_diag2q_counted");
}

int bass_diag2q_run(const uint64_t* x, double* a, int64_t n, int ngates, const int* qubits, const double* diags,
                    int counted, void* ws, size_t ws_bytes, void* stream) {
  if (n < 0 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_diag2q_run: bad n %lld", (long long)n);
  if (ngates < 0 || (ngates > 0 && (!qubits || !diags))) return fail(BASS_E_ARG, "bass_diag2q_run: bad gate list");
  for (int g = 0; g < ngates; ++g) {
    const int q1 = qubits[2 * g], q2 = qubits[2 * g + 1];
    if (q1 < 0 || q1 > 62 || q2 < 0 || q2 > 62 || q1 == q2)
      return fail(BASS_E_ARG, "bass_diag2q_run: gate %d: bad qubits %d %d", g, q1, q2);
  }
  const int64_t* n_dev = nullptr;
  if (counted) {
    WS_REQUIRE(0);
    n_dev = at<int64_t>(ws, L.count_slot);
  }
  hipStream_t s = (hipStream_t)stream;
  const int tk = ktimer_begin("diag2q", s);
  for (int g0 = 0; g0 < ngates; g0 += DIAG_RUN_MAX) {
    DiagRun R;
    memset(&R, 0, sizeof(R));
    R.ngates = ngates - g0 < DIAG_RUN_MAX ? ngates - g0 : DIAG_RUN_MAX;
    for (int g = 0; g < R.ngates; ++g) {
      R.q[g][0] = (unsigned char)qubits[2 * (g0 + g)];
      R.q[g][1] = (unsigned char)qubits[2 * (g0 + g) + 1];
      for (int e = 0; e < 4; ++e)
        R.d[g][e] = make_double2(diags[8 * (g0 + g) + 2 * e], diags[8 * (g0 + g) + 2 * e + 1]);
    }
    launch_diag2q_run(x, (double2*)a, n, R, s, n_dev);
  }
  ktimer_end(tk, s);
  return hip_check("bass_diag2q_run");
}

int bass_set_count(void* ws, size_t ws_bytes, int64_t n, void* stream) {
  if (n < 0) return fail(BASS_E_ARG, "bass_set_count: bad n %lld", (long long)n);
  WS_REQUIRE(0);
  fill(at<void>(ws, L.count_slot), 8, (unsigned long long)n, (hipStream_t)stream);
  return hip_check("bass_set_count");
}

int bass_counted_read(void* ws, size_t ws_bytes, bass_stats* out, int64_t max_entries, int64_t* n_entries,
                      void* stream) {
  WS_REQUIRE(0);
  hipStream_t s = (hipStream_t)stream;
  unsigned long long n = 0;
  if (hipMemcpyAsync(&n, at<void>(ws, L.oplog), 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_counted_read: %s", hipGetErrorString(hipGetLastError()));
  *n_entries = (int64_t)n;
  if ((int64_t)n > OPLOG_CAP) return fail(BASS_E_OVERFLOW, "bass_counted_read: op log overflowed (%llu entries)", n);
  if ((int64_t)n > max_entries) return fail(BASS_E_OVERFLOW, "bass_counted_read: %llu entries to read", n);
  if (n > 0 && (hipMemcpyAsync(out, at<char>(ws, L.oplog) + 8, n * sizeof(OpStats), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
    return fail(BASS_E_HIP, "bass_counted_read: %s", hipGetErrorString(hipGetLastError()));
  fill(at<void>(ws, L.oplog), 8, 0ull, s);
  return BASS_OK;
}

static int lift_impl(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                     uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws, size_t ws_bytes,
                     void* stream, bool counted) {
  // mode bit 4: no scan of the output's tile counts (a segmented truncate reads them);
  // bit 5: speculative digit-1 pass of the select that reads the output next
  const bool no_scan = (mode & 16) != 0;
  const bool spec = (mode & 32) != 0;
  mode &= ~48;
  if (spec && mode == MODE_MOVE) return fail(BASS_E_ARG, "bass_lift: the speculative pass needs a gate or rotation");
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
  // cleared before the gate kernel accumulates: the digit-1 histogram (gates)
  // and the first 48 bytes of the stats block (err untouched); by the lift's
  // first boundary kernel when it runs, else here
  LiftZero zero;
  if (mode != MODE_MOVE) {
    const int M = 1 << nq;
    for (int e = 0; e < M * M; ++e) gm.m[e] = make_double2(G[2 * e], G[2 * e + 1]);
    zero.p = at<unsigned long long>(ws, L.hist1);
    zero.wp = HIST_BINS * 4 / 8;
    zero.q = (unsigned long long*)st;
    zero.wq = 48 / 8;
  } else {
    zero.p = (unsigned long long*)st;
    zero.wp = 48 / 8;
  }
  if (n <= 0) {
    if (zero.q)
      zero2(zero.p, (size_t)zero.wp * 8, zero.q, (size_t)zero.wq * 8, s);
    else
      fill(zero.p, (size_t)zero.wp * 8, 0ull, s);
  }
  LiftSpec sp;
  memset(&sp, 0, sizeof(sp));
  if (spec) {
    // the candidate arrays where bass_truncate places them for an input of
    // `tiles` segments
    const int64_t slots = tiles * SEG_ENTRIES;
    const CandRegion C = cand_region(ws, L, slots);
    if (C.bytes - al((size_t)8 * slots) > L.total - L.big)
      return fail(BASS_E_WORKSPACE, "bass_lift: candidates do not fit");
    sp.pred = at<unsigned int>(ws, L.spec);
    sp.out.c1p = C.c1p;
    sp.out.c1i = C.c1i;
    sp.out.cnt1 = at<int>(ws, L.cnt1);
    sp.out.tile_dsum = at<double>(ws, L.tile_dsum);
    sp.out.tile_dcnt = at<int>(ws, L.tile_dcnt);
  }
  if (n > 0) {
    int* seg = at<int>(ws, L.seg_cnt);
    launch_lift(nq, mode, x, (const double2*)a, n, j, gm, xo, (double2*)ao, at<LiftBound>(ws, L.lift_b),
                at<unsigned int>(ws, L.hist1), at<double>(ws, L.tile_psum), seg, s,
                counted ? at<int64_t>(ws, L.count_slot) : nullptr, zero, spec ? &sp : nullptr);
    if (mode != MODE_MOVE && !no_scan)
      launch_seg_scan(seg, tiles, at<int64_t>(ws, L.soff), seg_chunk_cnt(ws, L), seg_chunk_off(ws, L), st, s);
  }
  ktimer_end(tk, s);
  return hip_check("bass_lift");
}

int bass_densify(const uint64_t* x, const double* a, int64_t seg_tiles, uint64_t* xo, double* ao,
                 int64_t cap_out, void* ws, size_t ws_bytes, void* stream) {
  if (seg_tiles < 0) return fail(BASS_E_ARG, "bass_densify: bad segment count");
  WS_REQUIRE(seg_tiles * SEG_ENTRIES);
  (void)cap_out;
  launch_densify(x, (const double2*)a, at<int>(ws, L.seg_cnt), at<int64_t>(ws, L.soff), seg_chunk_off(ws, L),
                 seg_tiles, xo,
                 (double2*)ao, (hipStream_t)stream);
  return hip_check("bass_densify");
}

static int truncate_impl(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo, double* ao,
                         int64_t cap_out, int renorm, int64_t hist_tiles, int in_flags, const int32_t* key_map,
                         int key_bits, void* ws, size_t ws_bytes, void* stream, bool timed = true,
                         const LiftRecompute* fused = nullptr, const LiftRecompute* relift = nullptr);

int bass_truncate(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo, double* ao,
                  int64_t cap_out, int renorm, int64_t hist_tiles, int in_flags, const int32_t* key_map,
                  int key_bits, void* ws, size_t ws_bytes, void* stream) {
  return truncate_impl(x, a, n, k, xo, ao, cap_out, renorm, hist_tiles, in_flags, key_map, key_bits, ws, ws_bytes,
                       stream);
}

// fused: the input is the unstored output of a fused lift (k_lift_gate with
// FUSE) on this workspace; x is ignored (the candidates' keys are read from
// the candidate region), and the select recomputes the lift's tiles.
static int truncate_impl(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo, double* ao,
                         int64_t cap_out, int renorm, int64_t hist_tiles, int in_flags, const int32_t* key_map,
                         int key_bits, void* ws, size_t ws_bytes, void* stream, bool timed,
                         const LiftRecompute* fused, const LiftRecompute* relift) {
  const bool segmented = (in_flags & 1) != 0;
  const bool wide = (in_flags & 2) != 0;
  const bool counted = (in_flags & 4) != 0;
  const bool spec_in = (in_flags & 8) != 0;
  if (spec_in && !segmented) return fail(BASS_E_ARG, "bass_truncate: a speculative lift's output is segmented");
  if (k < 1 || (!segmented && n <= k))
    return fail(BASS_E_ARG, "bass_truncate: need n > k >= 1 (n=%lld k=%lld)", (long long)n, (long long)k);
  if (cap_out < k) return fail(BASS_E_OVERFLOW, "bass_truncate: cap_out < k");
  if (n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_truncate: n too large");
  if (segmented && hist_tiles <= 0) return fail(BASS_E_ARG, "bass_truncate: segmented input needs hist_tiles");
  KeyMap km;
  const int krc = parse_key_map(key_map, key_bits, km, "bass_truncate");
  if (krc != BASS_OK) return krc;
  const int64_t tiles = segmented ? hist_tiles : sel_tiles(n);
  const int64_t slots = tiles * SEG_ENTRIES;
  // per-tile sums read: a segmented input's tiles hold SEG_ENTRIES slots, a
  // dense gate output's emission tiles >= 512 (the layout keeps S / 512 + 2)
  const int64_t psum_slots = hist_tiles <= 0 ? 0 : segmented ? hist_tiles * SEG_ENTRIES
                                                             : (hist_tiles > 2 ? (hist_tiles - 2) * 512 : 0);
  const int64_t need_ws = slots > n ? slots : n;
  WS_REQUIRE(need_ws > psum_slots ? need_ws : psum_slots);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  // header B (digit-2 and wide histograms, select state, tickets) is cleared by
  // k_find1, the select's first kernel after the histogram pass
  if (hist_tiles <= 0) fill(at<void>(ws, L.hist1), HIST_BINS * 4, 0ull, s);
  const CandRegion C = cand_region(ws, L, slots);
  const size_t cbytes = fused ? C.bytes : C.bytes - al((size_t)8 * slots);
  if (cbytes > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_truncate: candidates do not fit");
  SelProblem P;
  memset(&P, 0, sizeof(P));
  P.x = fused ? (const uint64_t*)C.c1k : x;
  P.a = fused ? nullptr : (const double2*)a;
  P.fused = fused ? 1 : 0;
  P.pruned = relift ? 1 + relift->margin : 0;
  P.seg_cnt = segmented ? at<int>(ws, L.seg_cnt) : nullptr;
  P.n_dense = n;
  P.tile0 = 0;
  P.ntiles = tiles;
  P.row0 = 0;
  P.nrows = select_rows(tiles, false);
  P.npsum = hist_tiles > 0 ? hist_tiles : tiles;
  P.chunked = sel_chunks(P.npsum > tiles ? P.npsum : tiles) > 1 ? 1 : 0;
  P.finish_any = (counted || fused) ? 1 : 0;
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
  P.c2p = C.c2p;
  P.c2i = C.c2i;
  P.tk = C.tk;
  P.zero_hdr = at<unsigned long long>(ws, L.hist2);
  P.zero_words = (int64_t)((L.hdrB_end - L.hist2) / 8);
  P.km = km;
  P.ktab = km.nbits ? at<uint64_t>(ws, L.ktab) : nullptr;
  P.spec_word = at<unsigned int>(ws, L.spec);
  P.spec_in = spec_in && !wide ? 1 : 0;
  SelTiles T = tiles_view(ws, L, C.c1p, C.c1i);
  if (fused) T.c1k = C.c1k;
  const int tk = timed ? ktimer_begin("select", s) : -1;
  launch_select(P, T, hist_tiles > 0, wide, s, fused, relift);
  ktimer_end(tk, s);
  if (counted)
    hipLaunchKernelGGL(k_oplog_append, dim3(1), dim3(1), 0, s, (const OpStats*)st,
                       at<unsigned long long>(ws, L.oplog), (OpStats*)(at<char>(ws, L.oplog) + 8),
                       at<int64_t>(ws, L.count_slot));
  return hip_check("bass_truncate");
}

int bass_gate_truncate(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                       int64_t k, uint64_t* xo, double* ao, int64_t cap_out, int renorm, int flags,
                       const int32_t* key_map, int key_bits, void* ws, size_t ws_bytes, void* stream) {
  const bool counted = (flags & 1) != 0;
  const bool spec = (flags & 2) != 0;
  if (flags & ~3) return fail(BASS_E_ARG, "bass_gate_truncate: unknown flags 0x%x", flags);
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_gate_truncate: nq must be 1 or 2");
  if (mode != MODE_GATE && mode != MODE_ROT)
    return fail(BASS_E_ARG, "bass_gate_truncate: mode must be 0 (gate) or 1 (rotation)");
  if (j < nq - 1 || j > 62) return fail(BASS_E_ARG, "bass_gate_truncate: bad bit %d for %d qubit(s)", j, nq);
  if (n < 1 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_gate_truncate: bad n %lld", (long long)n);
  if (k < 1) return fail(BASS_E_ARG, "bass_gate_truncate: need k >= 1");
  if (cap_out < k) return fail(BASS_E_OVERFLOW, "bass_gate_truncate: cap_out < k");
  if (!x || !a || !xo || !ao || !G) return fail(BASS_E_ARG, "bass_gate_truncate: null pointer");
  {
    // the compaction reads the input while it writes the output
    const char *xb = (const char*)x, *ab = (const char*)a, *xob = (const char*)xo, *aob = (const char*)ao;
    const size_t ni = (size_t)n, no = (size_t)cap_out;
    auto overlap = [](const char* p, size_t pn, const char* q, size_t qn) { return p < q + qn && q < p + pn; };
    if (overlap(xb, 8 * ni, xob, 8 * no) || overlap(xb, 8 * ni, aob, 16 * no) || overlap(ab, 16 * ni, xob, 8 * no) ||
        overlap(ab, 16 * ni, aob, 16 * no))
      return fail(BASS_E_ARG, "bass_gate_truncate: the output overlaps the input");
  }
  KeyMap km;
  const int krc = parse_key_map(key_map, key_bits, km, "bass_gate_truncate");
  if (krc != BASS_OK) return krc;
  const int tm = SEG_ENTRIES >> nq;
  const int64_t tiles = (n + tm - 1) / tm;
  const int64_t slots = tiles * SEG_ENTRIES;
  WS_REQUIRE(slots > n ? slots : n);
  const CandRegion C = cand_region(ws, L, slots);
  if (C.bytes > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_gate_truncate: candidates do not fit");
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  GateMat gm;
  memset(&gm, 0, sizeof(gm));
  const int M = 1 << nq;
  for (int e = 0; e < M * M; ++e) gm.m[e] = make_double2(G[2 * e], G[2 * e + 1]);
  // cleared by the lift's first boundary kernel: the digit-1 histogram and
  // the first 48 bytes of the stats block (err untouched)
  LiftZero zero;
  zero.p = at<unsigned long long>(ws, L.hist1);
  zero.wp = HIST_BINS * 4 / 8;
  zero.q = (unsigned long long*)st;
  zero.wq = 48 / 8;
  LiftSpec sp;
  memset(&sp, 0, sizeof(sp));
  if (spec) {
    sp.pred = at<unsigned int>(ws, L.spec);
    sp.out.c1p = C.c1p;
    sp.out.c1i = C.c1i;
    sp.out.cnt1 = at<int>(ws, L.cnt1);
    sp.out.tile_dsum = at<double>(ws, L.tile_dsum);
    sp.out.tile_dcnt = at<int>(ws, L.tile_dcnt);
    sp.out.c1k = C.c1k;
  }
  const int64_t* n_dev = counted ? at<int64_t>(ws, L.count_slot) : nullptr;
  LiftBound* bounds = at<LiftBound>(ws, L.lift_b);
  static const char* names[2][2] = {{"gate1q", "gate2q"}, {"rot1q", "rot2q"}};
  const int tk = ktimer_begin(names[mode][nq - 1], s);
  launch_lift(nq, mode, x, (const double2*)a, n, j, gm, nullptr, nullptr, bounds, at<unsigned int>(ws, L.hist1),
              at<double>(ws, L.tile_psum), at<int>(ws, L.seg_cnt), s, n_dev, zero, spec ? &sp : nullptr, true);
  ktimer_end(tk, s);
  const int hrc = hip_check("bass_gate_truncate");
  if (hrc != BASS_OK) return hrc;
  LiftRecompute R;
  R.qb = nq;
  R.mode = mode;
  R.x = x;
  R.a = (const double2*)a;
  R.n = n;
  R.n_dev = n_dev;
  R.j = j;
  R.bounds = bounds;
  R.G = gm;
  return truncate_impl(nullptr, nullptr, 0, k, xo, ao, cap_out, renorm, tiles, 1 | (counted ? 4 : 0) | (spec ? 8 : 0),
                       key_map, key_bits, ws, ws_bytes, stream, true, &R);
}

int bass_lift_truncate(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                       int64_t k, uint64_t* seg_x, double* seg_a, int64_t seg_cap, uint64_t* xo, double* ao,
                       int64_t cap_out, int renorm, int flags, const int32_t* key_map, int key_bits, void* ws,
                       size_t ws_bytes, void* stream) {
  const bool counted = (flags & 1) != 0;
  const bool spec = (flags & 2) != 0;
  // bit 2: prune to the predicted bucket itself (margin 0) instead of one
  // binade below it
  const unsigned int margin = (flags & 4) ? 0u : PRUNE_MARGIN;
  if (flags & ~7) return fail(BASS_E_ARG, "bass_lift_truncate: unknown flags 0x%x", flags);
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_lift_truncate: nq must be 1 or 2");
  if (mode != MODE_GATE && mode != MODE_ROT)
    return fail(BASS_E_ARG, "bass_lift_truncate: mode must be 0 (gate) or 1 (rotation)");
  if (j < nq - 1 || j > 62) return fail(BASS_E_ARG, "bass_lift_truncate: bad bit %d for %d qubit(s)", j, nq);
  if (n < 1 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_lift_truncate: bad n %lld", (long long)n);
  if (k < 1) return fail(BASS_E_ARG, "bass_lift_truncate: need k >= 1");
  if (cap_out < k) return fail(BASS_E_OVERFLOW, "bass_lift_truncate: cap_out < k");
  if (!x || !a || !xo || !ao || !G || !seg_x || !seg_a) return fail(BASS_E_ARG, "bass_lift_truncate: null pointer");
  const int tm = SEG_ENTRIES >> nq;
  const int64_t tiles = (n + tm - 1) / tm;
  const int64_t slots = tiles * SEG_ENTRIES;
  if (seg_cap < slots) return fail(BASS_E_OVERFLOW, "bass_lift_truncate: segments need %lld slots", (long long)slots);
  {
    // the segments may share memory with neither the input (a relift reads
    // it after the lift wrote them) nor the output (the compaction reads
    // them while it writes); the output may be the input
    const char *xb = (const char*)x, *ab = (const char*)a, *xob = (const char*)xo, *aob = (const char*)ao;
    const char *sxb = (const char*)seg_x, *sab = (const char*)seg_a;
    const size_t ni = (size_t)n, no = (size_t)cap_out, ns = (size_t)slots;
    auto overlap = [](const char* p, size_t pn, const char* q, size_t qn) { return p < q + qn && q < p + pn; };
    const char* seg[2] = {sxb, sab};
    const size_t segn[2] = {8 * ns, 16 * ns};
    const char* oth[4] = {xb, ab, xob, aob};
    const size_t othn[4] = {8 * ni, 16 * ni, 8 * no, 16 * no};
    for (int u = 0; u < 2; ++u)
      for (int v = 0; v < 4; ++v)
        if (overlap(seg[u], segn[u], oth[v], othn[v]))
          return fail(BASS_E_ARG, "bass_lift_truncate: the segment buffers overlap the input or the output");
    if (overlap(sxb, 8 * ns, sab, 16 * ns)) return fail(BASS_E_ARG, "bass_lift_truncate: segment buffers overlap");
  }
  KeyMap km;
  const int krc = parse_key_map(key_map, key_bits, km, "bass_lift_truncate");
  if (krc != BASS_OK) return krc;
  WS_REQUIRE(slots > n ? slots : n);
  const CandRegion C = cand_region(ws, L, slots);
  if (C.bytes - al((size_t)8 * slots) > L.total - L.big)
    return fail(BASS_E_WORKSPACE, "bass_lift_truncate: candidates do not fit");
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  GateMat gm;
  memset(&gm, 0, sizeof(gm));
  const int M = 1 << nq;
  for (int e = 0; e < M * M; ++e) gm.m[e] = make_double2(G[2 * e], G[2 * e + 1]);
  LiftZero zero;
  zero.p = at<unsigned long long>(ws, L.hist1);
  zero.wp = HIST_BINS * 4 / 8;
  zero.q = (unsigned long long*)st;
  zero.wq = 48 / 8;
  LiftSpec sp;
  memset(&sp, 0, sizeof(sp));
  if (spec) {
    sp.pred = at<unsigned int>(ws, L.spec);
    sp.out.c1p = C.c1p;
    sp.out.c1i = C.c1i;
    sp.out.cnt1 = at<int>(ws, L.cnt1);
    sp.out.tile_dsum = at<double>(ws, L.tile_dsum);
    sp.out.tile_dcnt = at<int>(ws, L.tile_dcnt);
    sp.margin = margin;
  }
  const int64_t* n_dev = counted ? at<int64_t>(ws, L.count_slot) : nullptr;
  LiftBound* bounds = at<LiftBound>(ws, L.lift_b);
  static const char* names[2][2] = {{"gate1q", "gate2q"}, {"rot1q", "rot2q"}};
  const int tk = ktimer_begin(names[mode][nq - 1], s);
  // with the prediction the lift also prunes (stores only outputs that can
  // matter to a threshold within PRUNE_MARGIN buckets below it)
  launch_lift(nq, mode, x, (const double2*)a, n, j, gm, seg_x, (double2*)seg_a, bounds, at<unsigned int>(ws, L.hist1),
              at<double>(ws, L.tile_psum), at<int>(ws, L.seg_cnt), s, n_dev, zero, spec ? &sp : nullptr, false, spec);
  ktimer_end(tk, s);
  const int hrc = hip_check("bass_lift_truncate");
  if (hrc != BASS_OK) return hrc;
  LiftRecompute R;
  R.qb = nq;
  R.mode = mode;
  R.x = x;
  R.a = (const double2*)a;
  R.n = n;
  R.n_dev = n_dev;
  R.j = j;
  R.bounds = bounds;
  R.G = gm;
  R.margin = (int)margin;
  return truncate_impl(seg_x, seg_a, 0, k, xo, ao, cap_out, renorm, tiles, 1 | (counted ? 4 : 0) | (spec ? 8 : 0),
                       key_map, key_bits, ws, ws_bytes, stream, true, nullptr, spec ? &R : nullptr);
}

int bass_truncate_noalloc(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo, double* ao,
                          int64_t cap_out, int renorm, const int32_t* key_map, int key_bits, bass_stats* stats_out,
                          void* stream) {
  if (n > NOALLOC_MAX_ENTRIES)
    return fail(BASS_E_ARG, "bass_truncate_noalloc: n = %lld exceeds the maximum of %lld entries", (long long)n,
                (long long)NOALLOC_MAX_ENTRIES);
  if (!stats_out) return fail(BASS_E_ARG, "bass_truncate_noalloc: stats_out is null");
  void* ws = nullptr;
  const int rc0 = noalloc_scratch(&ws);
  if (rc0) return rc0;
  hipStream_t s = (hipStream_t)stream;
  const Layout L = layout(ws_capacity(NOALLOC_SCRATCH_BYTES));
  // the array keeps the previous call's contents: a fresh stats block
  fill(at<void>(ws, L.st), sizeof(OpStats), 0ull, s);
  const int rc = truncate_impl(x, a, n, k, xo, ao, cap_out, renorm, 0, 0, key_map, key_bits, ws,
                               NOALLOC_SCRATCH_BYTES, stream, false);
  if (rc) return rc;
  hipLaunchKernelGGL(k_copy_stats, dim3(1), dim3(1), 0, s, at<const OpStats>(ws, L.st), (OpStats*)stats_out);
  return hip_check("bass_truncate_noalloc");
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
    int log2cap = table_log2cap(n, 12, L.total - L.big);
    
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

int bass_rdm2q_ordered(const uint64_t* x, const double* a, int64_t n, int j, double* rho_out, void* ws,
                       size_t ws_bytes, void* stream) {
  if (j < 1 || j > 62) return fail(BASS_E_ARG, "bass_rdm2q_ordered: bit j must be in [1, 62]");
  if (n < 0 || n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_rdm2q_ordered: bad n");
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  double* v = at<double>(ws, L.rho);  // 16 doubles
  if (n <= 0) {
    fill(v, 16 * 8, 0ull, s);
  } else {
    if (rdm2_scratch(n) > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_rdm2q_ordered: scratch does not fit");
    const int tm = ktimer_begin("rdm2", s);
    launch_rdm2(x, (const double2*)a, n, j, at<void>(ws, L.big), v, s);
    ktimer_end(tm, s);
  }
  int rc = hip_check("bass_rdm2q_ordered");
  if (rc) return rc;
  double h[16];
  if (hipMemcpyAsync(h, v, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_rdm2q_ordered copy: %s", hipGetErrorString(hipGetLastError()));
  // full 4x4 complex matrix, row-major (re, im): diagonal real, lower triangle
  // the conjugate of the upper
  for (int e = 0; e < 32; ++e) rho_out[e] = 0.0;
  for (int p = 0; p < 4; ++p) rho_out[2 * (p * 4 + p)] = h[p];
  int k = 4;
  for (int p = 0; p < 4; ++p)
    for (int q = p + 1; q < 4; ++q) {
      rho_out[2 * (p * 4 + q)] = h[k];
      rho_out[2 * (p * 4 + q) + 1] = h[k + 1];
      rho_out[2 * (q * 4 + p)] = h[k];
      rho_out[2 * (q * 4 + p) + 1] = -h[k + 1];
      k += 2;
    }
  return BASS_OK;
}

// ------------------------------------------------------------ dense ----
int bass_dense_scatter(const uint64_t* x, const double* a, int64_t n, const int32_t* key_map, int key_bits,
                       double* psi, int N, void* stream) {
  if (N < 1 || N > 40) return fail(BASS_E_ARG, "bass_dense_scatter: N out of range");
  if (n < 0 || n > (1ll << N)) return fail(BASS_E_ARG, "bass_dense_scatter: bad n");
  KeyMap km;
  memset(&km, 0, sizeof(km));
  if (key_bits) {
    if (key_bits != N || !key_map) return fail(BASS_E_ARG, "bass_dense_scatter: key_map must have N entries");
    uint64_t seen = 0;
    for (int p = 0; p < key_bits; ++p) {
      const int q = key_map[p];
      if (q < 0 || q >= N || ((seen >> q) & 1ull)) return fail(BASS_E_ARG, "bass_dense_scatter: key_map is not a permutation");
      seen |= 1ull << q;
      km.q[p] = (unsigned char)q;
    }
    km.nbits = key_bits;
  }
  hipStream_t s = (hipStream_t)stream;
  launch_dense_scatter(x, (const double2*)a, n, km, (double2*)psi, 1ll << N, s);
  return hip_check("bass_dense_scatter");
}

int bass_dense_gate(double* psi, int N, int nq, int q0, int q1, const double* G, void* stream) {
  if (N < 1 || N > 40) return fail(BASS_E_ARG, "bass_dense_gate: N out of range");
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_dense_gate: nq must be 1 or 2");
  if (q0 < 0 || q0 >= N || (nq == 2 && (q1 < 0 || q1 >= N || q1 == q0)))
    return fail(BASS_E_ARG, "bass_dense_gate: bad qubits");
  if (nq == 2 && N < 2) return fail(BASS_E_ARG, "bass_dense_gate: 2-qubit gate needs N >= 2");
  GateMat g;
  memset(&g, 0, sizeof(g));
  const int M = 1 << nq;
  for (int e = 0; e < M * M; ++e) g.m[e] = make_double2(G[2 * e], G[2 * e + 1]);
  launch_dense_gate((double2*)psi, N, nq, q0, q1, g, (hipStream_t)stream);
  return hip_check("bass_dense_gate");
}

int bass_dense_vdot(const double* u, const double* v, int64_t len, double* out, void* ws, size_t ws_bytes,
                    void* stream) {
  if (len < 0) return fail(BASS_E_ARG, "bass_dense_vdot: bad length");
  WS_REQUIRE(0);
  if (dense_vdot_scratch(len) + 16 > L.total - L.big) return fail(BASS_E_WORKSPACE, "bass_dense_vdot: scratch does not fit");
  hipStream_t s = (hipStream_t)stream;
  double2* res = at<double2>(ws, L.rho);
  if (len == 0) {
    fill(res, 16, 0ull, s);
  } else {
    launch_dense_vdot((const double2*)u, (const double2*)v, len, at<void>(ws, L.big), res, s);
  }
  int rc = hip_check("bass_dense_vdot");
  if (rc) return rc;
  if (hipMemcpyAsync(out, res, 16, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_dense_vdot copy: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

// ----------------------------------------------------------- batches ----
// Regions of a batch launch sequence, from a base offset, for jc jobs and tc
// output tiles: per-job stats first (bass_batch_stats reads them at the base
// whatever the rest), select states, histograms, RDM results, the staged
// tables (each header directly followed by its jobs: one copy), per-tile
// scratch, lift boundaries, and the select's per-slot candidate arrays (also
// the scratch of the RDM and moment calls).
struct BLayout {
  size_t st, sel, hist1, hist2, rho, hdr_l, ljobs, probs, hdr_e, ewjobs, rjobs;
  size_t tile_psum, seg_cnt, cnt1, tile_dsum, tile_dcnt, ksum, kcnt, koff, tile_s2, tile_s4, hist2_rows, bounds;
  size_t c1p, c1i, c2p, c2i, tk, end;
  int jc;
  int64_t tc;
};

static BLayout blayout(size_t base, int jc, int64_t tc) {
  BLayout B;
  B.jc = jc;
  B.tc = tc;
  const size_t t = (size_t)(tc > 0 ? tc : 0);
  const size_t slots = t * SEG_ENTRIES;
  size_t o = base;
  B.st = o; o += al((size_t)jc * sizeof(OpStats));
  B.sel = o; o += al((size_t)jc * sizeof(SelState));
  B.hist1 = o; o += al((size_t)jc * HIST_BINS * 4);
  B.hist2 = o; o += al((size_t)jc * HIST_BINS * 4);
  B.rho = o; o += al((size_t)jc * 64 * 4 * 8);
  B.hdr_l = o; B.ljobs = o + sizeof(BatchHdr); o = al(B.ljobs + (size_t)jc * sizeof(LiftJob));
  B.probs = o; o += al((size_t)jc * sizeof(SelProblem));
  B.hdr_e = o; B.ewjobs = o + sizeof(BatchHdr); o = al(B.ewjobs + (size_t)jc * sizeof(EwJob));
  B.rjobs = o; o += al((size_t)jc * sizeof(RdmJob));
  B.tile_psum = o; o += al(t * 8);
  B.seg_cnt = o; o += al(t * 4);
  B.cnt1 = o; o += al(t * 4);
  B.tile_dsum = o; o += al(t * 8);
  B.tile_dcnt = o; o += al(t * 4);
  B.ksum = o; o += al(t * 8);
  B.kcnt = o; o += al(t * 4);
  B.koff = o; o += al(t * 8);
  B.tile_s2 = o; o += al(t * 8);
  B.tile_s4 = o; o += al(t * 8);
  B.hist2_rows = o; o += al(t * HIST_BINS * 2);
  B.bounds = o; o += al((t + (size_t)jc + 1) * sizeof(LiftBound));
  B.c1p = o; o += al(slots * 8);
  B.c1i = o; o += al(slots * 4);
  B.c2p = o; o += al(slots * 8);
  B.c2i = o; o += al(slots * 4);
  B.tk = o; o += al(slots * 8);
  B.end = o;
  return B;
}

static SelTiles btiles(void* ws, const BLayout& B) {
  SelTiles T;
  memset(&T, 0, sizeof(T));
  T.tile_psum = at<double>(ws, B.tile_psum);
  T.c1p = at<unsigned long long>(ws, B.c1p);
  T.c1i = at<unsigned int>(ws, B.c1i);
  T.cnt1 = at<int>(ws, B.cnt1);
  T.tile_dsum = at<double>(ws, B.tile_dsum);
  T.tile_dcnt = at<int>(ws, B.tile_dcnt);
  T.ksum = at<double>(ws, B.ksum);
  T.kcnt = at<int>(ws, B.kcnt);
  T.koff = at<int64_t>(ws, B.koff);
  T.tile_s2 = at<double>(ws, B.tile_s2);
  T.tile_s4 = at<double>(ws, B.tile_s4);
  T.hist2_rows = at<unsigned short>(ws, B.hist2_rows);
  return T;
}

// Pinned staging ring for host -> device job tables: a slot is reused only
// after the copy that last read it has completed.  32 slots: a pipelined
// engine stages up to ~12 tables per tick and runs a tick ahead of the device,
// so reusing a slot must not wait on the tick in flight.
constexpr int STAGE_SLOTS = 32;
struct StageSlot {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
};
static std::mutex g_stage_mu;
static StageSlot g_stage[STAGE_SLOTS];
static int g_stage_next = 0;

static int stage_copy(void* dev, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  StageSlot& S = g_stage[g_stage_next];
  g_stage_next = (g_stage_next + 1) % STAGE_SLOTS;
  if (S.done) (void)hipEventSynchronize(S.done);
  if (S.cap < bytes) {
    if (S.host) (void)hipHostFree(S.host);
    S.host = nullptr;
    S.cap = 0;
    // grow geometrically from 256 KiB so a slot is reallocated a few times at most
    size_t want = (size_t)1 << 18;
    while (want < bytes) want <<= 1;
    const hipError_t e = hipHostMalloc(&S.host, want);
    if (e != hipSuccess) {
      S.host = nullptr;
      return fail(BASS_E_HIP, "pinned staging allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
    }
    S.cap = want;
  }
  if (!S.done && hipEventCreateWithFlags(&S.done, hipEventDisableTiming) != hipSuccess)
    return fail(BASS_E_HIP, "staging event creation failed");
  memcpy(S.host, src, bytes);
  if (hipMemcpyAsync(dev, S.host, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipEventRecord(S.done, s) != hipSuccess)
    return fail(BASS_E_HIP, "job table copy: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

static int check_job(const bass_job& J, int c, bool need_k, int jmin, const char* who, KeyMap* km) {
  if (J.n < 0 || J.n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "%s: job %d: bad n %lld", who, c, (long long)J.n);
  if (J.j < jmin || J.j > 62) return fail(BASS_E_ARG, "%s: job %d: bad bit %d", who, c, J.j);
  if (need_k && (J.k < 1 || J.cap_out < J.k))
    return fail(BASS_E_OVERFLOW, "%s: job %d: need 1 <= k <= cap_out", who, c);
  if (km) {
    memset(km, 0, sizeof(*km));
    if (J.key_bits) {
      if (J.key_bits < 1 || J.key_bits > 63) return fail(BASS_E_ARG, "%s: job %d: key_bits out of range", who, c);
      uint64_t seen = 0;
      for (int p = 0; p < J.key_bits; ++p) {
        const int q = J.key_map[p];
        if (q < 0 || q >= J.key_bits || ((seen >> q) & 1ull))
          return fail(BASS_E_ARG, "%s: job %d: key_map is not a permutation", who, c);
        seen |= 1ull << q;
        km->q[p] = (unsigned char)q;
      }
      km->nbits = J.key_bits;
    }
  }
  return BASS_OK;
}

// Host tables of a batch lift (step: with select problems; move: without).
// The staged blob is the header followed by the lift jobs.  *tbound: output
// tiles for the jobs' n (a job with a count cell gives its count's bound).
struct BTables {
  std::vector<char> hl;  // BatchHdr + LiftJob[njobs]
  std::vector<SelProblem> pr;
};

// nq: gate qubits of every job, or 0 for each job's own (bass_job.nq); the
// kinds present are returned in *qmask (bit nq - 1).
static int build_lift_tables(const bass_job* jobs, int njobs, int nq, int mode, void* ws, const BLayout& B,
                             uint64_t* tiles_x, double* tiles_a, BTables& X, int64_t* tbound, int* qmask,
                             const char* who) {
  const bool step = mode != MODE_MOVE;
  *qmask = 0;
  X.hl.assign(sizeof(BatchHdr) + (size_t)njobs * sizeof(LiftJob), 0);
  BatchHdr* h = (BatchHdr*)X.hl.data();
  LiftJob* lj = (LiftJob*)(X.hl.data() + sizeof(BatchHdr));
  if (step) X.pr.assign(njobs, SelProblem());
  int64_t T = 0;
  for (int c = 0; c < njobs; ++c) {
    const bass_job& J = jobs[c];
    const int qb = step ? (nq ? nq : J.nq) : 0;
    if (step && qb != 1 && qb != 2) return fail(BASS_E_ARG, "%s: job %d: gate qubits must be 1 or 2", who, c);
    if (step && mode == MODE_ROT && nq == 0) return fail(BASS_E_ARG, "%s: rotations take one kind per call", who);
    KeyMap km;
    const int rc = check_job(J, c, step, step ? qb - 1 : 0, who, step ? &km : nullptr);
    if (rc) return rc;
    if (step) *qmask |= 1 << (qb - 1);
    if (!step && J.cap_out < J.n) return fail(BASS_E_OVERFLOW, "%s: job %d: cap_out < n", who, c);
    LiftJob& L = lj[c];
    L.x = J.x;
    L.a = (const double2*)J.a;
    L.xo = J.xo;
    L.ao = (double2*)J.ao;
    L.n = J.n;
    L.cnt_in = J.cnt_in;
    L.cnt_out = step ? nullptr : J.cnt_out;
    L.j = J.j;
    L.qb = qb;
    if (step) {
      const int M = 1 << qb;
      for (int e = 0; e < M * M; ++e) L.g.m[e] = make_double2(J.g[2 * e], J.g[2 * e + 1]);
      SelProblem& P = X.pr[c];
      memset(&P, 0, sizeof(P));
      P.x = tiles_x;
      P.a = (const double2*)tiles_a;
      P.seg_cnt = at<int>(ws, B.seg_cnt);
      P.k = (unsigned long long)J.k;
      P.xo = J.xo;
      P.ao = (double2*)J.ao;
      P.cap_out = J.cap_out;
      P.renorm = J.renorm;
      P.finish_any = 1;
      P.hist1 = at<unsigned int>(ws, B.hist1) + (size_t)c * HIST_BINS;
      P.hist2 = at<unsigned int>(ws, B.hist2) + (size_t)c * HIST_BINS;
      P.sel = at<SelState>(ws, B.sel) + c;
      P.st = at<OpStats>(ws, B.st) + c;
      P.km = km;
      P.cnt_out = J.cnt_out;
    }
    const int tm = lift_job_tm(qb);
    T += (J.n + tm - 1) / tm;
  }
  h->njobs = njobs;
  h->cap_tiles = B.tc;
  *tbound = T;
  if (T > B.tc) return fail(BASS_E_OVERFLOW, "%s: jobs need %lld tiles, the buffers hold %lld", who, (long long)T,
                            (long long)B.tc);
  return BASS_OK;
}

// Launch sequences.  No argument depends on a count: tiles / jobs are bounds
// that size the persistent grids.
// qmask: gate kinds of the jobs (bit 0 one qubit, bit 1 two), 0 for a move.
static void seq_lift(void* ws, const BLayout& B, int qmask, int mode, uint64_t* tiles_x, double* tiles_a,
                     int64_t tiles, int jobs, hipStream_t s) {
  LiftJob* lj = at<LiftJob>(ws, B.ljobs);
  BatchHdr* hdr = at<BatchHdr>(ws, B.hdr_l);
  if (mode == MODE_MOVE) {
    launch_batch_prep(lj, nullptr, hdr, SelBase{nullptr, nullptr, nullptr},
                      BatchZero{nullptr, nullptr, nullptr, nullptr}, s);
    launch_lift_batch(0, MODE_MOVE, lj, hdr, tiles, jobs, at<LiftBound>(ws, B.bounds), nullptr, nullptr, nullptr,
                      nullptr, nullptr, s);
    return;
  }
  SelProblem* pr = at<SelProblem>(ws, B.probs);
  const SelBase sb{at<unsigned long long>(ws, B.c2p), at<unsigned int>(ws, B.c2i), at<uint64_t>(ws, B.tk)};
  const BatchZero z{at<OpStats>(ws, B.st), at<SelState>(ws, B.sel), at<unsigned int>(ws, B.hist1),
                    at<unsigned int>(ws, B.hist2)};
  launch_batch_prep(lj, pr, hdr, sb, z, s);
  launch_lift_batch(qmask, mode, lj, hdr, tiles, jobs, at<LiftBound>(ws, B.bounds), tiles_x, (double2*)tiles_a,
                    at<unsigned int>(ws, B.hist1), at<double>(ws, B.tile_psum), at<int>(ws, B.seg_cnt), s);
  launch_select_batch(pr, hdr, btiles(ws, B), tiles, jobs, s);
}

// Elementwise tables: header + EwJob[njobs]; *ptiles: moment tiles.
static int build_ew_tables(int op, const bass_job* jobs, int njobs, std::vector<char>& he, int64_t* ptiles,
                           const char* who) {
  he.assign(sizeof(BatchHdr) + (size_t)njobs * sizeof(EwJob), 0);
  BatchHdr* h = (BatchHdr*)he.data();
  EwJob* ej = (EwJob*)(he.data() + sizeof(BatchHdr));
  int64_t ntiles = 0;
  for (int c = 0; c < njobs; ++c) {
    const bass_job& J = jobs[c];
    if (J.n < 0) return fail(BASS_E_ARG, "%s: job %d: bad n", who, c);
    if (op == 0 && (J.j < 0 || J.j > 62 || J.j2 < 0 || J.j2 > 62 || J.j == J.j2))
      return fail(BASS_E_ARG, "%s: job %d: bad bits", who, c);
    if (op == 2 && J.cnt_in) return fail(BASS_E_ARG, "%s: job %d: moments take the staged count", who, c);
    EwJob& E = ej[c];
    E.x = J.x;
    E.a = (double2*)J.a;
    E.n = J.n;
    E.cnt_in = J.cnt_in;
    E.j1 = J.j;
    E.j2 = J.j2;
    for (int e = 0; e < 4; ++e) E.d[e] = make_double2(J.g[2 * e], J.g[2 * e + 1]);
    E.div = J.scale;
    E.toff = ntiles;
    ntiles += (J.n + SEG_ENTRIES - 1) / SEG_ENTRIES;
  }
  h->njobs = njobs;
  *ptiles = ntiles;
  return BASS_OK;
}

// RDM tables (staged counts) and scratch needs.
static int build_rdm_tables(const bass_job* jobs, int njobs, std::vector<RdmJob>& rj, int64_t* nb_, int64_t* np_,
                            int64_t* nd_) {
  const int tm = rdm_tile_entries();
  rj.assign(njobs, RdmJob());
  int64_t nb = 0, np = 0, nd = 0;
  for (int c = 0; c < njobs; ++c) {
    const bass_job& J = jobs[c];
    if (J.n < 0 || J.n > 0x7FFFFFFFll) return fail(BASS_E_ARG, "bass_batch_rdm1q: job %d: bad n", c);
    if (J.key_bits < 1 || J.key_bits > 63) return fail(BASS_E_ARG, "bass_batch_rdm1q: job %d: bad key_bits", c);
    if (J.cnt_in) return fail(BASS_E_ARG, "bass_batch_rdm1q: job %d: RDMs take the staged count", c);
    RdmJob& R = rj[c];
    memset(&R, 0, sizeof(R));
    R.x = J.x;
    R.a = (const double2*)J.a;
    R.n = J.n;
    R.nbits = J.key_bits;
    R.ptiles = (J.n + tm - 1) / tm;
    R.dtiles = (J.n + SEG_ENTRIES - 1) / SEG_ENTRIES;
    R.bound_off = nb;
    R.ptile_off = np;
    R.dtile_off = nd;
    nb += (int64_t)J.key_bits * (R.ptiles + 1);
    np += (int64_t)J.key_bits * R.ptiles;
    nd += R.dtiles;
  }
  *nb_ = nb;
  *np_ = np;
  *nd_ = nd;
  return BASS_OK;
}

static int d2h_async(void* host, const void* dev, size_t bytes, hipStream_t s, const char* who) {
  if (!host || !bytes) return BASS_OK;
  if (hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s) != hipSuccess)
    return fail(BASS_E_HIP, "%s: result copy: %s", who, hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

// ------------------------------------------------------------ batch plan ----
// A plan fixes the workspace regions for jc jobs and tc tiles, so every call's
// launch sequence has one shape; with graphs on, each sequence is captured
// once into a HIP graph (on a private stream) and replayed on the caller's
// stream after the call's tables are staged.
enum { BOP_GATE1 = 0, BOP_GATE2, BOP_ROT1, BOP_ROT2, BOP_MOVE, BOP_DIAG2Q, BOP_SCALE, BOP_MOMENTS, BOP_RDM1Q,
       BOP_GATE12, BOP_COUNT };
// Graphs per size class of an operation: tile and job bounds rounded up to
// powers of 4 (grids near the work of the call; blocks past it exit).
constexpr int BCLS_T = 12, BCLS_J = 8;
static int bclass(int64_t v, int ncls) {
  int c = 0;
  int64_t b = 1;
  while (b < v && c < ncls - 1) {
    b <<= 2;
    ++c;
  }
  return c;
}
static int64_t bclass_bound(int c) { return (int64_t)1 << (2 * c); }

struct BPlan {
  int jc = 0;
  int64_t tc = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  uint64_t* tiles_x = nullptr;
  double* tiles_a = nullptr;
  BLayout B;
  bool graphs = false;
  int dev = 0;
  hipStream_t cap = nullptr;
  hipGraphExec_t exec[BOP_COUNT][BCLS_T][BCLS_J] = {};
  BTables X;
  std::vector<char> he;
  std::vector<RdmJob> rj;
};

// The fixed sequence of op (tiles / jobs: bounds sizing the persistent grids).
static void plan_sequence(BPlan* P, int op, int64_t tiles, int jobs, hipStream_t s) {
  switch (op) {
    case BOP_GATE1: seq_lift(P->ws, P->B, 1, MODE_GATE, P->tiles_x, P->tiles_a, tiles, jobs, s); break;
    case BOP_GATE2: seq_lift(P->ws, P->B, 2, MODE_GATE, P->tiles_x, P->tiles_a, tiles, jobs, s); break;
    case BOP_GATE12: seq_lift(P->ws, P->B, 3, MODE_GATE, P->tiles_x, P->tiles_a, tiles, jobs, s); break;
    case BOP_ROT1: seq_lift(P->ws, P->B, 1, MODE_ROT, P->tiles_x, P->tiles_a, tiles, jobs, s); break;
    case BOP_ROT2: seq_lift(P->ws, P->B, 2, MODE_ROT, P->tiles_x, P->tiles_a, tiles, jobs, s); break;
    case BOP_MOVE: seq_lift(P->ws, P->B, 0, MODE_MOVE, nullptr, nullptr, tiles, jobs, s); break;
    case BOP_DIAG2Q:
    case BOP_SCALE:
      launch_ew_batch(op == BOP_DIAG2Q ? 0 : 1, at<EwJob>(P->ws, P->B.ewjobs), at<BatchHdr>(P->ws, P->B.hdr_e), jobs,
                      0, nullptr, at<OpStats>(P->ws, P->B.st), s);
      break;
    default: break;
  }
}

static int plan_launch(BPlan* P, int op, int64_t tiles, int jobs, hipStream_t s) {
  if (!P->graphs) {
    plan_sequence(P, op, tiles, jobs, s);
    return hip_check("bass_bplan_run");
  }
  // the size class of the call's bounds (the last class holds any size: its
  // grids are the plan's capacity)
  const int ct = bclass(tiles, BCLS_T), cj = bclass(jobs, BCLS_J);
  hipGraphExec_t& ex = P->exec[op][ct][cj];
  if (!ex) {
    // capture once per class, with the persistent grids of its bounds
    const int64_t bt = ct == BCLS_T - 1 ? P->tc : bclass_bound(ct);
    const int bj = cj == BCLS_J - 1 ? P->jc : (int)bclass_bound(cj);
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail(BASS_E_HIP, "bass_bplan_run: begin capture: %s", hipGetErrorString(hipGetLastError()));
    plan_sequence(P, op, bt > P->tc ? P->tc : bt, bj > P->jc ? P->jc : bj, P->cap);
    const hipError_t e1 = hipStreamEndCapture(P->cap, &g);
    if (e1 != hipSuccess || !g) return fail(BASS_E_HIP, "bass_bplan_run: end capture: %s", hipGetErrorString(e1));
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      ex = nullptr;
      return fail(BASS_E_HIP, "bass_bplan_run: graph instantiation: %s", hipGetErrorString(e2));
    }
  }
  if (hipGraphLaunch(ex, s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_bplan_run: graph launch: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

size_t bass_bplan_ws_bytes(int jobs_cap, int64_t tiles_cap) {
  if (jobs_cap < 1 || tiles_cap < 1) return 0;
  return blayout(0, jobs_cap, tiles_cap).end;
}

int bass_bplan_create(int jobs_cap, int64_t tiles_cap, uint64_t* tiles_x, double* tiles_a, void* ws, size_t ws_bytes,
                      int flags, void** plan) {
  if (!plan) return fail(BASS_E_ARG, "bass_bplan_create: null plan");
  *plan = nullptr;
  if (jobs_cap < 1 || jobs_cap > (1 << 20) || tiles_cap < 1)
    return fail(BASS_E_ARG, "bass_bplan_create: bad capacities");
  if (!ws || !tiles_x || !tiles_a) return fail(BASS_E_ARG, "bass_bplan_create: null buffer");
  const BLayout B = blayout(0, jobs_cap, tiles_cap);
  if (B.end > ws_bytes)
    return fail(BASS_E_WORKSPACE, "bass_bplan_create: workspace of %zu bytes, plan needs %zu", ws_bytes, B.end);
  BPlan* P = new BPlan();
  P->jc = jobs_cap;
  P->tc = tiles_cap;
  P->ws = ws;
  P->ws_bytes = ws_bytes;
  P->tiles_x = tiles_x;
  P->tiles_a = tiles_a;
  P->B = B;
  P->graphs = (flags & 1) != 0;
  (void)hipGetDevice(&P->dev);
  if (P->graphs && hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking) != hipSuccess) {
    delete P;
    return fail(BASS_E_HIP, "bass_bplan_create: capture stream: %s", hipGetErrorString(hipGetLastError()));
  }
  *plan = P;
  return BASS_OK;
}

int bass_bplan_destroy(void* plan) {
  BPlan* P = (BPlan*)plan;
  if (!P) return BASS_OK;
  for (int i = 0; i < BOP_COUNT; ++i)
    for (int a = 0; a < BCLS_T; ++a)
      for (int b = 0; b < BCLS_J; ++b)
        if (P->exec[i][a][b]) (void)hipGraphExecDestroy(P->exec[i][a][b]);
  if (P->cap) (void)hipStreamDestroy(P->cap);
  delete P;
  return BASS_OK;
}

int bass_bplan_run(void* plan, int op, const bass_job* jobs, int njobs, void* out_host, void* stream) {
  BPlan* P = (BPlan*)plan;
  if (!P) return fail(BASS_E_ARG, "bass_bplan_run: null plan");
  if (op < 0 || op >= BOP_COUNT) return fail(BASS_E_ARG, "bass_bplan_run: bad op %d", op);
  if (njobs < 1 || njobs > P->jc) return fail(BASS_E_ARG, "bass_bplan_run: %d jobs, plan holds %d", njobs, P->jc);
  hipStream_t s = (hipStream_t)stream;
  void* ws = P->ws;
  const BLayout& B = P->B;
  static const char* names[BOP_COUNT] = {"batch_gate1q", "batch_gate2q", "batch_rot1q", "batch_rot2q", "batch_move",
                                         "batch_diag2q", "batch_scale", "batch_moments", "batch_rdm", "batch_gate12"};
  int rc = BASS_OK;
  if (op <= BOP_MOVE || op == BOP_GATE12) {
    const int nq = op == BOP_GATE12 ? 0 : ((op == BOP_GATE2 || op == BOP_ROT2) ? 2 : 1);
    const int mode = op == BOP_MOVE ? MODE_MOVE : ((op == BOP_ROT1 || op == BOP_ROT2) ? MODE_ROT : MODE_GATE);
    int64_t T = 0;
    int qmask = 0;
    rc = build_lift_tables(jobs, njobs, nq, mode, ws, B, P->tiles_x, P->tiles_a, P->X, &T, &qmask, "bass_bplan_run");
    if (rc) return rc;
    const int tk = ktimer_begin(names[op], s);
    rc = stage_copy(at<void>(ws, B.hdr_l), P->X.hl.data(), P->X.hl.size(), s);
    if (!rc && mode != MODE_MOVE)
      rc = stage_copy(at<void>(ws, B.probs), P->X.pr.data(), sizeof(SelProblem) * njobs, s);
    if (!rc) rc = plan_launch(P, op, T, njobs, s);
    ktimer_end(tk, s);
    if (!rc && mode != MODE_MOVE) rc = d2h_async(out_host, at<void>(ws, B.st), sizeof(bass_stats) * njobs, s, names[op]);
    return rc;
  }
  if (op == BOP_DIAG2Q || op == BOP_SCALE || op == BOP_MOMENTS) {
    const int eop = op == BOP_DIAG2Q ? 0 : (op == BOP_SCALE ? 1 : 2);
    int64_t pt = 0;
    rc = build_ew_tables(eop, jobs, njobs, P->he, &pt, "bass_bplan_run");
    if (rc) return rc;
    if (eop == 2 && (size_t)pt * 16 + 256 > B.end - B.c1p)
      return fail(BASS_E_WORKSPACE, "bass_bplan_run: moments of %lld tiles do not fit", (long long)pt);
    const int tk = ktimer_begin(names[op], s);
    rc = stage_copy(at<void>(ws, B.hdr_e), P->he.data(), P->he.size(), s);
    if (!rc && eop < 2) {
      rc = plan_launch(P, op, 1, njobs, s);
    } else if (!rc) {
      fill(at<void>(ws, B.st), sizeof(OpStats) * njobs, 0ull, s);
      launch_ew_batch(2, at<EwJob>(ws, B.ewjobs), at<BatchHdr>(ws, B.hdr_e), njobs, pt, at<double>(ws, B.c1p),
                      at<OpStats>(ws, B.st), s);
      rc = hip_check("bass_bplan_run");
      if (!rc) rc = d2h_async(out_host, at<void>(ws, B.st), sizeof(bass_stats) * njobs, s, names[op]);
    }
    ktimer_end(tk, s);
    return rc;
  }
  // RDM1Q: staged counts, result rho (njobs x 64 x 4 doubles) copied to out_host
  int64_t nb = 0, np = 0, nd = 0;
  rc = build_rdm_tables(jobs, njobs, P->rj, &nb, &np, &nd);
  if (rc) return rc;
  if (rdm_batch_scratch(nb, np, nd) > B.end - B.c1p)
    return fail(BASS_E_WORKSPACE, "bass_bplan_run: RDM scratch does not fit");
  const size_t rho_bytes = (size_t)njobs * 64 * 4 * 8;
  double* rho = at<double>(ws, B.rho);
  const int tk = ktimer_begin(names[op], s);
  fill(rho, rho_bytes, 0ull, s);
  rc = stage_copy(at<void>(ws, B.rjobs), P->rj.data(), sizeof(RdmJob) * njobs, s);
  if (!rc) {
    launch_rdm_batch(at<RdmJob>(ws, B.rjobs), njobs, nb, np, nd, at<void>(ws, B.c1p), rho, s);
    rc = hip_check("bass_bplan_run");
  }
  ktimer_end(tk, s);
  if (!rc) rc = d2h_async(out_host, rho, rho_bytes, s, names[op]);
  return rc;
}

// ---------------------------------------------- batch calls without a plan --
// Laid out per call from the workspace's big area for the call's own counts
// (direct launches, exact grids).
int bass_batch_step(const bass_job* jobs, int njobs, int nq, int mode, int wide, uint64_t* tiles_x, double* tiles_a,
                    int64_t tiles_cap, void* ws, size_t ws_bytes, void* stream) {
  (void)wide;  // every threshold is resolved in the call
  if (njobs < 1) return fail(BASS_E_ARG, "bass_batch_step: no jobs");
  if (nq != 1 && nq != 2) return fail(BASS_E_ARG, "bass_batch_step: nq must be 1 or 2");
  if (mode != MODE_GATE && mode != MODE_ROT) return fail(BASS_E_ARG, "bass_batch_step: bad mode");
  const int tm = lift_tile_entries(nq, mode);
  int64_t T = 0;
  for (int c = 0; c < njobs; ++c) T += (jobs[c].n > 0 ? (jobs[c].n + tm - 1) / tm : 0);
  const int64_t slots = T * SEG_ENTRIES;
  if (tiles_cap < slots)
    return fail(BASS_E_OVERFLOW, "bass_batch_step: tiles need %lld slots", (long long)slots);
  WS_REQUIRE(slots > njobs ? slots : njobs);
  const BLayout B = blayout(L.big, njobs, T > 0 ? T : 1);
  if (B.end > L.total) return fail(BASS_E_WORKSPACE, "bass_batch_step: batch does not fit");
  thread_local BTables X;
  int64_t Tb = 0;
  int qmask = 0;
  int rc = build_lift_tables(jobs, njobs, nq, mode, ws, B, tiles_x, tiles_a, X, &Tb, &qmask, "bass_batch_step");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  static const char* names[2][2] = {{"batch_gate1q", "batch_gate2q"}, {"batch_rot1q", "batch_rot2q"}};
  const int tk = ktimer_begin(names[mode][nq - 1], s);
  rc = stage_copy(at<void>(ws, B.hdr_l), X.hl.data(), X.hl.size(), s);
  if (!rc) rc = stage_copy(at<void>(ws, B.probs), X.pr.data(), sizeof(SelProblem) * njobs, s);
  if (rc) return rc;
  seq_lift(ws, B, qmask, mode, tiles_x, tiles_a, T, njobs, s);
  ktimer_end(tk, s);
  return hip_check("bass_batch_step");
}

int bass_batch_move(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream) {
  if (njobs < 1) return fail(BASS_E_ARG, "bass_batch_move: no jobs");
  int64_t T = 0;
  for (int c = 0; c < njobs; ++c) T += (jobs[c].n > 0 ? (jobs[c].n + 1023) / 1024 : 0);
  WS_REQUIRE(T * 1024 + njobs);
  const BLayout B = blayout(L.big, njobs, T > 0 ? T : 1);
  if (B.end > L.total) return fail(BASS_E_WORKSPACE, "bass_batch_move: batch does not fit");
  thread_local BTables X;
  int64_t Tb = 0;
  int qmask = 0;
  int rc = build_lift_tables(jobs, njobs, 1, MODE_MOVE, ws, B, nullptr, nullptr, X, &Tb, &qmask, "bass_batch_move");
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int tk = ktimer_begin("batch_move", s);
  rc = stage_copy(at<void>(ws, B.hdr_l), X.hl.data(), X.hl.size(), s);
  if (rc) return rc;
  seq_lift(ws, B, 0, MODE_MOVE, nullptr, nullptr, T, njobs, s);
  ktimer_end(tk, s);
  return hip_check("bass_batch_move");
}

static int batch_ew(int op, const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream,
                    const char* who) {
  if (njobs < 1) return fail(BASS_E_ARG, "%s: no jobs", who);
  thread_local std::vector<char> he;
  int64_t pt = 0;
  int rc = build_ew_tables(op, jobs, njobs, he, &pt, who);
  if (rc) return rc;
  WS_REQUIRE(0);
  const BLayout B = blayout(L.big, njobs, 0);
  if (B.end + ew_batch_scratch(pt) > L.total) return fail(BASS_E_WORKSPACE, "%s: batch does not fit", who);
  hipStream_t s = (hipStream_t)stream;
  if (op == 2) fill(at<void>(ws, B.st), (size_t)njobs * sizeof(OpStats), 0ull, s);
  rc = stage_copy(at<void>(ws, B.hdr_e), he.data(), he.size(), s);
  if (rc) return rc;
  launch_ew_batch(op, at<EwJob>(ws, B.ewjobs), at<BatchHdr>(ws, B.hdr_e), njobs, pt, at<double>(ws, B.end),
                  at<OpStats>(ws, B.st), s);
  return hip_check(who);
}

int bass_batch_diag2q(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream) {
  return batch_ew(0, jobs, njobs, ws, ws_bytes, stream, "bass_batch_diag2q");
}
int bass_batch_scale(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream) {
  return batch_ew(1, jobs, njobs, ws, ws_bytes, stream, "bass_batch_scale");
}
int bass_batch_moments(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream) {
  return batch_ew(2, jobs, njobs, ws, ws_bytes, stream, "bass_batch_moments");
}

int bass_batch_rdm1q(const bass_job* jobs, int njobs, double* rho_out, void* ws, size_t ws_bytes, void* stream) {
  if (njobs < 1) return fail(BASS_E_ARG, "bass_batch_rdm1q: no jobs");
  thread_local std::vector<RdmJob> rj;
  int64_t nb = 0, np = 0, nd = 0;
  int rc = build_rdm_tables(jobs, njobs, rj, &nb, &np, &nd);
  if (rc) return rc;
  WS_REQUIRE(0);
  const BLayout B = blayout(L.big, njobs, 0);
  const size_t rho_bytes = (size_t)njobs * 64 * 4 * 8;
  if (B.end + rdm_batch_scratch(nb, np, nd) > L.total)
    return fail(BASS_E_WORKSPACE, "bass_batch_rdm1q: batch does not fit");
  hipStream_t s = (hipStream_t)stream;
  double* rho = at<double>(ws, B.rho);
  const int tm2 = ktimer_begin("batch_rdm", s);
  fill(rho, rho_bytes, 0ull, s);
  rc = stage_copy(at<void>(ws, B.rjobs), rj.data(), sizeof(RdmJob) * njobs, s);
  if (rc) return rc;
  launch_rdm_batch(at<RdmJob>(ws, B.rjobs), njobs, nb, np, nd, at<void>(ws, B.end), rho, s);
  ktimer_end(tm2, s);
  rc = hip_check("bass_batch_rdm1q");
  if (rc) return rc;
  if (hipMemcpyAsync(rho_out, rho, rho_bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_batch_rdm1q copy: %s", hipGetErrorString(hipGetLastError()));
  return BASS_OK;
}

int bass_batch_stats(void* ws, size_t ws_bytes, int njobs, bass_stats* out, void* stream) {
  if (njobs < 1) return fail(BASS_E_ARG, "bass_batch_stats: no jobs");
  WS_REQUIRE(0);
  if (L.big + (size_t)njobs * sizeof(OpStats) > L.total) return fail(BASS_E_ARG, "bass_batch_stats: too many jobs");
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.big);  // the stats region of every batch layout starts at the big area
  if (hipMemcpyAsync(out, st, sizeof(bass_stats) * njobs, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(BASS_E_HIP, "bass_batch_stats: %s", hipGetErrorString(hipGetLastError()));
  for (int c = 0; c < njobs; ++c)
    if (out[c].err) {
      fill(st, sizeof(OpStats) * njobs, 0ull, s);
      return fail(BASS_E_DEVICE, "job %d: device error bits 0x%x", c, out[c].err);
    }
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
  if (n > SORT_MAX_N) return fail(BASS_E_ARG, "bass_radix_sort: n = %lld exceeds 2^30 - 1", (long long)n);
  if (n <= 1) return BASS_OK;
  WS_REQUIRE(n);
  hipStream_t s = (hipStream_t)stream;
  OpStats* st = at<OpStats>(ws, L.st);
  char* big = at<char>(ws, L.big);
  // two record buffers of n entries: 2 x (8 + 16) n bytes of the 48 S big region
  uint64_t* tk = (uint64_t*)big;
  double2* tv = (double2*)(big + al((size_t)8 * n));
  uint64_t* tk2 = (uint64_t*)(big + al((size_t)8 * n) + al((size_t)16 * n));
  double2* tv2 = (double2*)((char*)tk2 + al((size_t)8 * n));
  const int tm = ktimer_begin("sort", s);
  // histograms, tickets and the first pass's look-back words start at zero
  // (contiguous in the layout)
  fill(at<void>(ws, L.sort_hist), (L.sort_status - L.sort_hist) + (size_t)sweep_tiles(n) * 1024 * 4, 0ull, s);
  // an explicit process-wide width (bass_sort_digit_bits) wins over the table
  const int db = sort_digit_bits_get();
  launch_radix_sort(keys, (double2*)vals, n, key_bits, tk, tv, tk2, tv2, at<unsigned int>(ws, L.sort_hist),
                    at<unsigned int>(ws, L.sort_status), at<unsigned int>(ws, L.sort_tick), &st->err, s,
                    db ? 0 : tuned(TUNE_SORT_DIGITS, n));
  ktimer_end(tm, s);
  return hip_check("bass_radix_sort");
}

int bass_tune(int knob, int size_class, int value) {
  if (knob < 0 || knob >= TUNE_KNOBS) return fail(BASS_E_ARG, "bass_tune: unknown knob %d", knob);
  if (size_class < 0 || size_class >= TUNE_CLASSES) return fail(BASS_E_ARG, "bass_tune: bad size class %d", size_class);
  bool ok = value == 0;
  if (knob == TUNE_TABLE_FACTOR) ok |= value == 2 || value == 4;
  if (knob == TUNE_EMIT_ITEMS_1Q) ok |= emit_items_ok(1, value);
  if (knob == TUNE_EMIT_ITEMS_2Q) ok |= emit_items_ok(2, value);
  if (knob == TUNE_SORT_DIGITS) ok |= value == 8 || value == 10;
  if (!ok) return fail(BASS_E_ARG, "bass_tune: value %d not allowed for knob %d", value, knob);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const int old = g_tune[knob][size_class];
  g_tune[knob][size_class] = value;
  return old;
}

int bass_sort_digit_bits(int bits) {
  if (bits != 0 && bits != 8 && bits != 10) return fail(BASS_E_ARG, "bass_sort_digit_bits: %d is not 0, 8 or 10", bits);
  return sort_digit_bits_set(bits);
}

int bass_sort_plan(int key_bits, int* passes) {
  if (key_bits < 1 || key_bits > 64 || !passes) return fail(BASS_E_ARG, "bass_sort_plan: key_bits out of range");
  return sort_plan(key_bits, passes);
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
