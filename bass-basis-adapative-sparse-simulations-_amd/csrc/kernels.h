// Host-side launch interface shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bass {

constexpr int HIST_BINS = 2048;  // 11-bit digits of the |a|^2 bit pattern
enum { MODE_GATE = 0, MODE_ROT = 1, MODE_MOVE = 2 };

struct GateMat {
  double2 m[16];  // row-major (2^QB x 2^QB), index (out, in) with in = (b_q0 << 1) | b_q1
};

// Two word ranges a lift clears in the first kernel it launches (block 0),
// before its gate kernel accumulates into them: the digit-1 histogram and the
// stats block.  p == nullptr: nothing to clear.
struct LiftZero {
  unsigned long long* p = nullptr;
  unsigned long long* q = nullptr;
  int wp = 0, wq = 0;
};

// Top-k selection bookkeeping, device resident.
struct SelState {
  unsigned long long need;       // elements still to take inside the current bucket
  unsigned long long prefix;     // resolved high bits of the threshold pattern
  unsigned long long pstar;      // |a|^2 bit pattern of the k-th element
  unsigned long long keystar;    // largest key taken among elements equal to pstar
  unsigned long long cand1;      // candidates after digit 1
  unsigned int cand2;            // candidates after digit 2 (atomic append)
  unsigned int b1, b2;           // threshold buckets of digits 1 and 2
  unsigned int keepall;          // the input holds <= k entries: keep all, no renormalisation
  unsigned long long def_count;  // elements above the digit-1 bucket
  unsigned long long cnt_eq;     // wide path: entries equal to the resolved prefix
  unsigned long long kp;         // wide path: resolved high bits of the key threshold
  unsigned int tcount;           // wide path: tied candidates collected
  unsigned int tie;              // wide path: the key tie-break is needed
  // staged finish (k_finish_staged): block tickets, finished blocks and the
  // published flag of every stage, and an abort word for error exits
  unsigned int stg_ticket;
  unsigned int stg_abort;
  unsigned int stg_done[12];
  unsigned int stg_flag[12];
  // the speculative lift's digit-1 pass used this select's bucket: k_extract1
  // only builds the digit-2 rows from its candidates
  unsigned int spec_hit;
  // a pruned lift dropped outputs this select needs (its bucket fell more than
  // PRUNE_MARGIN below the prediction, or it keeps everything): k_lift_relift
  // stores the whole output again before the select reads it
  unsigned int relift;
};

// A pruned lift (bass_lift_truncate with the speculative pass) stores only
// the outputs whose |a|^2 exponent bucket is at least the predicted threshold
// bucket minus PRUNE_MARGIN; its histogram and sums still cover every output.
constexpr unsigned int PRUNE_MARGIN = 1;

// One block resolves at most FIN_MAX digit-2 candidates.  Beyond that a single
// problem's select runs the staged multi-block stages of k_finish_staged in
// the same launch; the problems of a batch (at most 4 x 2^20 inputs each)
// are resolved by their own block whatever the count.
constexpr unsigned int FIN_MAX = 1u << 16;
constexpr int WIDE_LEVELS = 10;  // 4 |a|^2 digits + 6 key digits

// Scalars an operation reports back to the host.
struct OpStats {
  int64_t nnz;    // output entry count
  double total;   // sum |a|^2 of the operation input (truncate) or output (gate)
  double kept;    // sum |a|^2 of the kept entries (before renormalisation)
  double norm;    // sqrt(kept) used for renormalisation (1 if not renormalised)
  double sum2;    // sum |a|^2 of the output as written
  double sum4;    // sum |a|^4 of the output as written
  int64_t n_in;   // truncate: input entry count
  int32_t err;    // error bits (ERR_*)
  int32_t flags;  // truncate: bit 0 = entries dropped
};

int group_slot_bytes(int qb);
int launch_group_build(int qb, const uint64_t* x, int64_t n, int q0, int q1, uint64_t* gslot, int log2cap,
                       int* err, hipStream_t s);
int emit_default_items(int qb);
bool emit_items_ok(int qb, int items);
int emit_tile_entries(int qb, int items = 0);
int launch_emit(int qb, int mode, bool dedup, const uint64_t* x, const double2* a, int64_t n, int q0,
                int q1, const uint64_t* gslot, int log2cap, const GateMat& G,
                uint64_t* xo, double2* ao, int64_t cap_out, unsigned long long* lb_status,
                unsigned int* ticket, unsigned int* hist, double* tile_psum, int64_t* n_out, int* seg_cnt,
                int* err, hipStream_t s, int items = 0);
// Every emission tile covers SEG_ENTRIES output slots (1q: 1024 inputs x 2, 2q: 512 x 4).
constexpr int SEG_ENTRIES = 2048;
int launch_diag2q(const uint64_t* x, double2* a, int64_t n, int q1, int q2, const double2* d,
                  hipStream_t s, const int64_t* n_dev = nullptr);
// A run of up to DIAG_RUN_MAX diagonal 2-qubit gates (kernel argument: 2 KB
// of tables), applied in order in one pass.
constexpr int DIAG_RUN_MAX = 32;
struct DiagRun {
  double2 d[DIAG_RUN_MAX][4];
  unsigned char q[DIAG_RUN_MAX][2];
  int ngates;
};
int launch_diag2q_run(const uint64_t* x, double2* a, int64_t n, const DiagRun& R, hipStream_t s,
                      const int64_t* n_dev = nullptr);

// Key-ordered state: merge split of the run around one tile boundary.
struct LiftBound {
  int s0, sm, s1;  // run [s0, s1); its entries with the lifted bit set start at sm
  int a, b;        // A / B cursors of the merge at the boundary
  int pad0, pad1, pad2;
};
// Header of a batch launch sequence (device memory).  The host stages njobs
// and cap_tiles; the sequence's prep kernel fills the rest from the jobs'
// input counts, so no launch argument depends on a count: every batch kernel
// runs a persistent grid over the totals here, and one sequence can be
// captured in a graph and replayed whatever the counts.
struct BatchHdr {
  int njobs;          // staged: jobs in the table
  int nrun;           // prep: jobs the kernels process (0 after an overflow)
  int64_t cap_tiles;  // staged: output tiles the tile buffers hold
  int64_t ntiles;     // prep: output tiles over all jobs
  int64_t nbound;     // prep: lift boundaries, ntiles + njobs
  int64_t max_rows;   // prep: largest tile count of one job
  int64_t pad[3];
};
// Persistent grids of the batch kernels: at most the blocks of the kernel the
// device holds at once (so every block starts at once and the work is shared
// evenly), at most `bound` (virtual blocks of work, when known), at least 1.
int resident_blocks(const void* kernel, int threads);
inline unsigned persist_grid(const void* kernel, int threads, int64_t bound) {
  const int64_t r = resident_blocks(kernel, threads);
  const int64_t g = bound < r ? bound : r;
  return (unsigned)(g < 1 ? 1 : g);
}

// One state of a batched lift (see launch_lift_batch).
struct LiftJob {
  const uint64_t* x;  // input (sorted by stored word)
  const double2* a;
  uint64_t* xo;       // move: dense output of this job
  double2* ao;
  int64_t n;          // input entries (the prep kernel's copy of *cnt_in when set)
  int64_t tile0;      // first global output tile (prep kernel)
  const int64_t* cnt_in;  // device count cell of the input, or null (n as staged)
  int64_t* cnt_out;       // move: count cell of the output (set to n), or null
  int j;              // bit to lift
  int qb;             // gate qubits (1, 2), or 0 for a move: tile size and lifted-to bit
  GateMat g;
};
// Tile entries and the bit lifted to, per job kind (LiftJob::qb).
__host__ __device__ inline int lift_job_tm(int qb) { return qb == 0 ? 1024 : SEG_ENTRIES >> qb; }
__host__ __device__ inline int lift_job_t(int qb) { return qb == 2 ? 1 : 0; }
// Tile-slot bases of the select's per-problem candidate arrays (the prep
// kernel points each problem at its own tiles).
struct SelBase {
  unsigned long long* c2p;
  unsigned int* c2i;
  uint64_t* tk;
};
struct SelProblem;
// Zeroed per-job regions of a batch step (cleared by its prep kernel).
struct BatchZero {
  OpStats* st;
  SelState* sel;
  unsigned int* hist1;
  unsigned int* hist2;
};
// Prep of a lift / step sequence: counts, tile geometry (tiles of each job's
// kind), header, zeroing (probs == nullptr: a move).
int launch_batch_prep(LiftJob* d_jobs, SelProblem* d_probs, BatchHdr* hdr, const SelBase& sb, const BatchZero& z,
                      hipStream_t s);
// qmask: gate kinds present (bit 0: 1-qubit jobs, bit 1: 2-qubit jobs; 0 = a
// move); tiles / jobs: bounds of the tile and job counts (size the grids)
int launch_lift_batch(int qmask, int mode, const LiftJob* d_jobs, const BatchHdr* hdr, int64_t tiles, int jobs,
                      LiftBound* bounds, uint64_t* xo, double2* ao, unsigned int* hist, double* tile_psum,
                      int* seg_cnt, hipStream_t s);
int lift_tile_entries(int qb, int mode);
// Outputs of the select's digit-1 pass per tile (ex1_tile_core).
struct Ex1Out {
  unsigned long long* c1p;
  unsigned int* c1i;
  int* cnt1;
  double* tile_dsum;
  int* tile_dcnt;
  // keys of the candidates (a lift whose output is not stored): c1k[i] is the
  // stored word of candidate i and c1i[i] = i, so a select reading keys as
  // x[c1i[i]] reads them from x = c1k
  unsigned long long* c1k;
};

// A gate lift that also runs the select's digit-1 pass on each output tile,
// against the bucket *pred (the previous select's), into the select's
// candidate arrays; the select uses them when its bucket equals *pred.
struct LiftSpec {
  const unsigned int* pred;
  Ex1Out out;
  unsigned int margin;  // a pruning lift keeps buckets >= *pred - margin (0 or PRUNE_MARGIN)
};

int launch_lift(int qb, int mode, const uint64_t* x, const double2* a, int64_t n, int j, const GateMat& G,
                uint64_t* xo, double2* ao, LiftBound* bounds, unsigned int* hist, double* tile_psum, int* seg_cnt,
                hipStream_t s, const int64_t* n_dev = nullptr, const LiftZero& zero = LiftZero(),
                const LiftSpec* spec = nullptr, bool fuse = false, bool prune = false);

// One state of a batched elementwise operation (see launch_ew_batch):
// op 0 diagonal 2-qubit gate a *= d[(b_j1 << 1) | b_j2], op 1 a /= div,
// op 2 moments into st[job].sum2 / sum4.
struct EwJob {
  const uint64_t* x;
  double2* a;
  int64_t n;
  const int64_t* cnt_in;  // device count cell of the state, or null (n as staged)
  int j1, j2;
  double2 d[4];
  double div;
  int64_t toff;  // moments: first of the job's SEG_ENTRIES tiles in the scratch arrays
};
// hdr->njobs jobs (staged); jobs: a bound of that count (sizes the grid);
// part_tiles: moment tiles of all jobs (part holds 2 x part_tiles doubles)
int launch_ew_batch(int op, const EwJob* d_jobs, const BatchHdr* hdr, int jobs, int64_t part_tiles,
                    double* part, OpStats* st, hipStream_t s);
size_t ew_batch_scratch(int64_t ntiles_total);

// One state of a batched RDM computation (see launch_rdm_batch).
struct RdmJob {
  const uint64_t* x;
  const double2* a;
  int64_t n;
  int64_t bound_off, ptile_off, dtile_off;  // offsets in the flattened index spaces
  int64_t ptiles, dtiles;                   // pair tiles per bit, diagonal tiles
  int nbits;
  int pad;
};
int launch_rdm_batch(const RdmJob* d_jobs, int njobs, int64_t nbound, int64_t nptiles, int64_t ndtiles, void* scratch,
                     double* rho, hipStream_t s);
size_t rdm_batch_scratch(int64_t nbound, int64_t nptiles, int64_t ndtiles);
int rdm_tile_entries();
size_t rdm_ordered_scratch(int64_t n, int nbits);
size_t rdm2_scratch(int64_t n);
int64_t lift_bounds_entries(int64_t nbound);
int launch_rdm2(const uint64_t* x, const double2* a, int64_t n, int j, void* scratch, double* out, hipStream_t s);
int launch_rdm_ordered(const uint64_t* x, const double2* a, int64_t n, int nbits, void* scratch, double* rho,
                       hipStream_t s);

// Stored-bit -> qubit map of a key-ordered state (nbits = 0: keys are stored as is).
struct KeyMap {
  int nbits;
  unsigned char q[64];
};

// One top-k truncation problem: its input tiles [tile0, tile0 + ntiles) of the
// per-tile scratch (SelTiles), its output and its own bookkeeping.
struct SelProblem {
  const uint64_t* x;     // input keys (segmented: the gate's tile buffer, indexed by global slot)
  const double2* a;
  const int* seg_cnt;    // per global tile counts of a segmented input, or null (dense)
  int64_t n_dense;       // dense input entries
  int64_t tile0, ntiles; // global tiles of the input
  int64_t row0, nrows;   // digit-2 histogram rows of the input
  int64_t npsum;         // tile_psum entries from tile0 holding this input's |a|^2 sums
  unsigned long long k;
  uint64_t* xo;
  double2* ao;
  int64_t cap_out;
  int renorm;
  int chunked;           // per-tile reductions run by chunk (single large problems)
  int finish_any;        // k_finish resolves any candidate count itself (no wide-path hand-off)
  int spec_in;           // the input comes from a speculative lift (LiftSpec) launched just before
  // the input is a gate lift's output that was not stored (fused gate select):
  // x is the candidate key list (SelTiles::c1k), a is unused; the digit-1 pass
  // after a missed speculation and the compaction recompute the lift's tiles
  int fused;
  // the input comes from a pruned lift issued just before, whose prediction
  // *spec_word still holds: k_find1 checks that nothing it dropped is needed.
  // 0: not pruned; else 1 + the lift's margin (it dropped buckets below
  // *spec_word - margin)
  int pruned;
  unsigned int* hist1;   // HIST_BINS, zeroed by the caller unless filled by the producing gate
  unsigned int* hist2;   // HIST_BINS, zeroed
  unsigned int* hist_w;  // HIST_BINS, zeroed (wide path)
  SelState* sel;         // zeroed
  OpStats* st;
  unsigned long long* c2p;  // >= ntiles * SEG_ENTRIES slots
  unsigned int* c2i;
  uint64_t* tk;             // >= ntiles * SEG_ENTRIES
  unsigned long long* zero_hdr;  // cleared by k_find1 before it starts (zero_words words), or nullptr
  int64_t zero_words;
  KeyMap km;             // ties are broken by the logical key
  // byte tables of km (8 x 256 words: entry 256 b + v holds the logical bits
  // of stored byte b = v), written by k_find1 when km.nbits > 0, or null
  // (the map is then applied bit by bit)
  uint64_t* ktab;
  // single problems: the digit-1 bucket a speculative lift predicts (the last
  // select's bucket, 0xFFFFFFFF after a keep-all select); k_find1 compares
  // and updates it.  Null: no speculation.
  unsigned int* spec_word;
  // batch: count cell set to the output count (k_sums_finalize), or null
  int64_t* cnt_out;
};
constexpr int KTAB_WORDS = 8 * 256;

// Per-tile scratch shared by all problems of a launch (indexed by global tile;
// c1p / c1i by global slot).
struct SelTiles {
  double* tile_psum;
  unsigned long long* c1p;
  unsigned int* c1i;
  int* cnt1;
  double* tile_dsum;
  int* tile_dcnt;
  double* ksum;
  int* kcnt;
  int64_t* koff;
  double* tile_s2;
  double* tile_s4;
  unsigned short* hist2_rows;
  unsigned long long* c1k;  // candidate keys of a fused gate select (Ex1Out::c1k), else null
  // per chunk of SEL_CHUNK tiles (chunked problems): kept count, kept-entry
  // offset, kept / input / output sums
  int64_t* ch_cnt;
  int64_t* ch_off;
  double* ch_sum;
  double* ch_psum;
  double* ch_s2;
  double* ch_s4;
};

// Tiles per chunk of the chunked per-tile reductions (one block of NTB threads
// handling 4 tiles each); inputs of more tiles than this reduce in two levels.
constexpr int64_t SEL_CHUNK = 4096;
inline int64_t sel_chunks(int64_t ntiles) { return (ntiles + SEL_CHUNK - 1) / SEL_CHUNK; }

int select_rows(int64_t ntiles, bool batch);
// The lift a fused gate select recomputes: its input state, the bit it lifts,
// its tile boundaries (still in the workspace from the lift) and matrix.
struct LiftRecompute {
  int qb, mode;
  const uint64_t* x;
  const double2* a;
  int64_t n;
  const int64_t* n_dev;  // counted input: the count, or null
  int j;
  const LiftBound* bounds;
  GateMat G;
  int margin;  // relift of a pruned lift: the margin it pruned with
};
// Fused select kernels over the lift's tiles: which = 0 the digit-1 pass when
// the speculation missed (ext tiles per digit-2 row), 1 the compaction; 2 the
// relift of a pruned lift's select (the whole output into the select's input
// tiles P.x / P.a / P.seg_cnt, when k_find1 asked for it).
int launch_lift_recompute(int which, const LiftRecompute& R, const SelProblem& P, const SelTiles& T, int ext,
                          hipStream_t s);
int extract1_tiles(int64_t ntiles);
int launch_select(const SelProblem& P, const SelTiles& T, bool hist_ready, bool wide, hipStream_t s,
                  const LiftRecompute* fused = nullptr, const LiftRecompute* relift = nullptr);
// A batch of problems with the geometry in hdr (prep kernel); tiles / jobs:
// bounds of the tile and problem counts (size the persistent grids).  Every
// threshold is resolved in the sequence (one block per problem).
int launch_select_batch(const SelProblem* d_probs, const BatchHdr* hdr, const SelTiles& T, int64_t tiles,
                        int jobs, hipStream_t s);
int64_t sel_tiles(int64_t n);
int launch_seg_scan(const int* seg_cnt, int64_t ntiles, int64_t* soff, int64_t* ch_cnt, int64_t* ch_off,
                    OpStats* st, hipStream_t s);
int launch_densify(const uint64_t* x, const double2* a, const int* seg_cnt, const int64_t* soff,
                   const int64_t* ch_off, int64_t ntiles, uint64_t* xo, double2* ao, hipStream_t s);
int launch_moments(const double2* a, int64_t n, double* tile_s2, double* tile_s4, OpStats* st,
                   hipStream_t s);
int launch_scale(double2* a, int64_t n, double div, hipStream_t s);
int launch_gather(const uint64_t* x, const double2* a, const int64_t* idx, int64_t m, uint64_t* xo,
                  double2* ao, hipStream_t s);
int rdm_blocks();
int launch_table_build(const uint64_t* x, int64_t n, uint64_t* tkeys, int32_t* tvals, int log2cap,
                       int overwrite, int* err, hipStream_t s);
int launch_table_lookup(const uint64_t* tkeys, const int32_t* tvals, int log2cap, const uint64_t* q,
                        int64_t m, int64_t* out, hipStream_t s);
int launch_rdm(const uint64_t* x, const double2* a, int64_t n, int N, const uint64_t* tkeys,
               const int32_t* tvals, int log2cap, double* part, double* rho, hipStream_t s);
constexpr int SORT_MAX_PASSES = 8;
constexpr int64_t SORT_MAX_N = (1ll << 30) - 1;  // 32-bit positions, 30-bit look-back counts
// (tk, tv), (tk2, tv2): two scratch record buffers of n entries; hist:
// SORT_MAX_PASSES x 1024 words, tickets: SORT_MAX_PASSES words, both zero on
// entry; status: sort_status_bytes(n), its first sweep_tiles(n) x 1024 words zero.
// digit_bits: 8 or 10, or 0 for the process-wide choice (sort_digit_bits_set)
int launch_radix_sort(uint64_t* keys, double2* vals, int64_t n, int key_bits, uint64_t* tk, double2* tv,
                      uint64_t* tk2, double2* tv2, unsigned int* hist, unsigned int* status,
                      unsigned int* tickets, int* err, hipStream_t s, int digit_bits = 0);
int sweep_tiles(int64_t n);
size_t sort_status_bytes(int64_t n);
int sort_plan(int key_bits, int* passes);
int sort_digit_bits_set(int bits);
int sort_digit_bits_get();
int launch_merge(const uint64_t* x, const double2* a, int64_t n, uint64_t* xo, double2* ao,
                 unsigned long long* lb, unsigned int* ticket, int64_t* n_out, int* err, hipStream_t s);

// Optional per-operation timing with HIP events recorded on the launch stream
// (enabled with bass_timing(1)).  ktimer_begin returns -1 when disabled.
int ktimer_begin(const char* name, hipStream_t s);
void ktimer_end(int token, hipStream_t s);

// dense state vectors (dense_kernels.hip)
int launch_dense_scatter(const uint64_t* x, const double2* a, int64_t n, const KeyMap& km, double2* psi, int64_t len,
                         hipStream_t s);
int launch_dense_gate(double2* psi, int N, int nq, int q0, int q1, const GateMat& G, hipStream_t s);
size_t dense_vdot_scratch(int64_t len);
int launch_dense_vdot(const double2* u, const double2* v, int64_t len, void* scratch, double2* out, hipStream_t s);

}  // namespace bass
