/* C ABI of the MI355X (gfx950) sparse-state core.
 *
 * Drop-in boundary for the reference's Numba kernels and drivers:
 *   src/core/sparse_state.py, src/core/truncation.py,
 *   src/simulation/bass_simulator.py, src/simulation/simulator.py.
 * The reference binds these kernels as @njit functions on NumPy arrays; the
 * Python host layer of this package binds the entry points below with ctypes
 * and keeps the reference's function names and argument order.
 *
 * Conventions
 *   - keys  : uint64_t[n] device array, qubit q = bit q (reference layout).
 *   - amps  : double[2 n] device array, complex128 (re, im) interleaved.
 *   - gate / unitary matrices: host double arrays, row-major complex
 *     (re, im) pairs, 2x2 (8 doubles) or 4x4 (32 doubles); 2-qubit index
 *     (b_q1 << 1) | b_q2 as in the reference.
 *   - ws    : caller-owned device workspace of bass_workspace_bytes(max_n)
 *     bytes; the library performs no device allocation of its own.
 *   - stream: hipStream_t (passed as void*); every call is asynchronous on it
 *     except those documented as synchronous.
 *   - return: 0 on success, a negative BASS_E* code otherwise;
 *     bass_last_error() describes the last failure of the calling thread.
 */
#ifndef BASS_HIP_H
#define BASS_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BASS_OK = 0,
  BASS_E_OVERFLOW = -1, /* output capacity exceeded; the reference returns -1
                           (src/core/sparse_state.py:117-118) */
  BASS_E_ARG = -2,      /* invalid argument */
  BASS_E_WORKSPACE = -3,/* workspace too small for this call */
  BASS_E_HIP = -4,      /* HIP runtime error */
  BASS_E_DEVICE = -5    /* device-side failure (probe limit, select bookkeeping) */
};

/* Scalars a device operation reports (read with bass_read_stats). */
typedef struct bass_stats {
  int64_t nnz;  /* output entry count */
  double total; /* sum |a|^2 of the truncation input (or of the gate output) */
  double kept;  /* sum |a|^2 of the kept entries before renormalisation */
  double norm;  /* divisor applied by renormalisation (1 when none) */
  double sum2;  /* sum |a|^2 of the output as written */
  double sum4;  /* sum |a|^4 of the output as written */
  int64_t n_in; /* bass_truncate: entries in its input */
  int32_t err;  /* device error bits (0 = none) */
  int32_t flags; /* bass_truncate: bit 0 = entries were dropped (0: kept all, no renormalisation) */
} bass_stats;

/* One state of a batched call (bass_batch_*).  Device pointers; the table
 * itself is host memory, staged by the library. */
typedef struct bass_job {
  const uint64_t* x;  /* input keys (stored words of a key-ordered state) */
  const double* a;    /* input amplitudes (re, im) */
  uint64_t* xo;       /* output keys: truncation / move destination */
  double* ao;
  int64_t n;          /* input entries */
  int64_t k;          /* truncation size */
  int64_t cap_out;    /* entries xo / ao hold */
  int32_t j;          /* stored bit lifted; first bit of a diagonal gate */
  int32_t j2;         /* second bit of a diagonal gate */
  int32_t renorm;     /* truncation: divide kept entries by sqrt(kept) */
  int32_t key_bits;   /* entries of key_map (0: keys stored as is) */
  double g[32];       /* gate matrix (re, im) row-major, 2x2 or 4x4; diagonal gate: 4 values */
  int32_t key_map[64];/* stored bit -> qubit */
  double scale;       /* bass_batch_scale divisor */
  const int64_t* cnt_in; /* device count cell of the input, or NULL: n entries.  When set,
                            the call reads the count on the device (n bounds it) */
  int64_t* cnt_out;   /* device count cell the call sets to its output count, or NULL */
  int32_t nq;         /* BASS_BOP_GATE12: this job's gate qubits (1 or 2) */
  int32_t pad;
} bass_job;

int bass_abi_version(void);
const char* bass_last_error(void);

/* Workspace size (bytes) for operations whose inputs and outputs hold at most
 * max_entries entries. */
size_t bass_workspace_bytes(int64_t max_entries);

/* Synchronous: waits for the stream, copies the last operation's scalars to
 * *out and clears the device error bits.  Returns BASS_E_DEVICE when err != 0. */
int bass_read_stats(void* ws, size_t ws_bytes, bass_stats* out, void* stream);

/* Gate application.  Replaces:
 *   dedup=0, nq=1: apply_1qubit_gate_kernel_opt   (src/core/sparse_state.py:15-67)
 *   dedup=0, nq=2: apply_2qubit_gate_kernel_opt   (src/core/sparse_state.py:70-124)
 *   dedup=1, nq=1, mode=0: apply_1qubit_gate_kernel_opt + radix_sort_uint64 +
 *                  merge_sorted_kernel (src/core/sparse_state.py:268-285), as a set
 *   dedup=1, nq=2, mode=0: apply_2qubit_gate_ht   (src/simulation/bass_simulator.py:292-354)
 *   dedup=1, mode=1: apply_1qubit_rotation_sparse (bass_simulator.py:156-199) /
 *                  apply_2qubit_rotation_sparse (bass_simulator.py:237-289)
 * Writes the outputs to (xo, ao) and, in the workspace, the |a|^2 digit
 * histogram consumed by a following bass_truncate(..., hist_tiles).
 * *hist_tiles receives the value to pass on (0 when no histogram was made).
 * Non-deduplicating calls keep the reference's output order (input order,
 * out index ascending); deduplicating calls order outputs by the first input
 * entry of each group.
 * segmented=0: dense output, cap_out >= 2^nq n.  segmented=1: tile t of the
 * input writes its outputs to slots [t*2048, t*2048 + count_t) with no
 * coordination between tiles (cap_out >= *hist_tiles * 2048); the result feeds
 * bass_truncate(..., segmented=1) or bass_densify.  stats.nnz is the total.
 * Returns BASS_E_OVERFLOW if cap_out is too small. */
int bass_gate(const uint64_t* x, const double* a, int64_t n, int nq, int q0, int q1,
              const double* G, int mode, int dedup, int segmented, uint64_t* xo, double* ao,
              int64_t cap_out, int64_t* hist_tiles, void* ws, size_t ws_bytes, void* stream);

/* Copies the segmented output of the preceding bass_gate (seg_tiles = its
 * *hist_tiles) to dense positions in (xo, ao), preserving order.  No reference
 * counterpart (the reference's kernels write densely). */
int bass_densify(const uint64_t* x, const double* a, int64_t seg_tiles, uint64_t* xo, double* ao,
                 int64_t cap_out, void* ws, size_t ws_bytes, void* stream);

/* Key-ordered gate application (the simulator's internal layout).  (x, a) is
 * sorted by stored word w, a bit permutation of the basis key that the caller
 * tracks.  The call moves bit j of w down to bit nq-1 -- bits nq-1..j-1 move
 * up by one -- and applies the gate to the groups this makes adjacent:
 *   mode=0 (gate) / 1 (rotation), nq=1: gate on the qubit stored at bit j;
 *   nq=2: gate on the qubits stored at bit j (first matrix index bit) and at
 *         bit 0 (second), j >= 1;
 *   mode=2, nq=1: the bit move alone (G unused), dense output of n entries.
 * Gate and rotation outputs are written per tile as in bass_gate(segmented=1)
 * (*seg_tiles receives the tile count for bass_truncate / bass_densify) and
 * come out sorted by the new stored word.  Output entries and amplitudes
 * equal those of bass_gate(dedup=1) on the same state with keys mapped back;
 * a group's products are summed in stored order.  Replaces, inside the
 * driver, the gate step of src/simulation/bass_simulator.py:861-896 and the
 * tentative rotation of :768-783 (no reference counterpart as an entry point).
 * mode | 16 (gate or rotation): skip the closing scan of the per-tile output
 * counts, so stats.nnz and the segment offsets bass_densify needs are not
 * produced; for an output that a bass_truncate(in_flags bit 0) reads next.
 * mode | 32 (gate only): also run the truncate's digit-1 pass on every output
 * tile (count and sum above one IEEE-exponent bucket of |a|^2, the candidates
 * inside it) against the bucket of the workspace's last select, into that
 * truncate's candidate arrays; a bass_truncate with in_flags bit 3 that reads
 * this output next, on the same workspace, skips its own pass over the
 * amplitudes when its bucket is the predicted one.  Results are the same
 * either way (the pass writes what the truncate's would).
 * Returns BASS_E_OVERFLOW if cap_out is below the output size. */
int bass_lift(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G,
              int mode, uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws,
              size_t ws_bytes, void* stream);

/* bass_lift, with the input count read on the device from the workspace's
 * state-count slot (written by a counted bass_truncate or bass_set_count) and
 * clamped to [0, n_max]; grids, output tiles (*seg_tiles) and buffer checks
 * follow n_max, tiles past the actual count come out empty (a move copies
 * the actual count and leaves the slot valid).  With counted truncations this
 * runs a gate loop without a host round trip per gate (the driver reads the
 * op log at its decision points).  mode | 16 as in bass_lift.  No counterpart
 * in the reference. */
int bass_lift_counted(const uint64_t* x, const double* a, int64_t n_max, int nq, int j, const double* G,
                      int mode, uint64_t* xo, double* ao, int64_t cap_out, int64_t* seg_tiles, void* ws,
                      size_t ws_bytes, void* stream);

/* bass_diag2q on the first min(count slot, n_max) entries (count from the
 * workspace's state-count slot; the count is unchanged).  No counterpart in
 * the reference. */
int bass_diag2q_counted(const uint64_t* x, double* a, int64_t n_max, int q1, int q2, const double* diag,
                        void* ws, size_t ws_bytes, void* stream);

/* ngates diagonal 2-qubit gates applied in order in one pass over the state:
 * for g = 0 .. ngates-1, alpha[i] *= diags[g][(b_{qubits[2g]} << 1) |
 * b_{qubits[2g+1]}] (diags: ngates x 4 complex128, interleaved re/im), the
 * same products in the same order as ngates calls of bass_diag2q, so the
 * result is bit-identical, from one read and one write of the state per 32
 * gates instead of one per gate.  counted != 0: the entry count is the
 * workspace's state-count slot clamped to n, as in bass_diag2q_counted
 * (ws unused otherwise).  The reference applies the gates one call each
 * (src/core/sparse_state.py:179-190 from src/simulation/bass_simulator.py:871-880);
 * this fuses a run of them and has no counterpart there. */
int bass_diag2q_run(const uint64_t* x, double* a, int64_t n, int ngates, const int* qubits, const double* diags,
                    int counted, void* ws, size_t ws_bytes, void* stream);

/* Write n into the workspace's state-count slot (asynchronous). */
int bass_set_count(void* ws, size_t ws_bytes, int64_t n, void* stream);

/* Copy the op log of counted truncations since the last read (one bass_stats
 * per truncation, in call order; at most 4096 between reads) into out
 * (max_entries), its length into *n_entries, and empty it.  Synchronous.
 * BASS_E_OVERFLOW if the log overflowed or does not fit out. */
int bass_counted_read(void* ws, size_t ws_bytes, bass_stats* out, int64_t max_entries, int64_t* n_entries,
                      void* stream);

/* Top-k truncation: keep the k entries of largest |a|^2 (ties: smaller key),
 * compacted in input order into (xo, ao); renorm=1 divides them by
 * sqrt(sum kept |a|^2).  Replaces top_k_truncate_kernel_opt
 * (src/core/truncation.py:68-113; renorm=0, gamma_sq = stats.kept) and the
 * driver truncations of src/simulation/bass_simulator.py:565-572, :899-916 and
 * src/simulation/simulator.py:169-201 (renorm=1).  Requires n > k.
 * hist_tiles: value reported by the bass_gate call that produced (x, a), or 0;
 * in_flags bit 0 (segmented): (x, a) is that call's segmented output; n is not needed
 * (the count comes from the device) and the input may hold <= k entries:
 * then all are kept, in order and not renormalised, and stats.flags bit 0 is
 * clear -- the gate / truncate pair needs no host round trip in between.
 * Any threshold bucket is resolved inside the call: up to 65536 candidates
 * by one block, more (massive |a|^2 ties) by multi-block stages of the same
 * launch; the call never asks for a repeat.
 * in_flags bit 1 (wide): resolve the threshold with the multi-launch passes
 * instead (kept for callers that request them; same result).
 * in_flags bit 2 (counted): then append the stats block to the workspace's op
 * log and set its state-count slot to the output count, for
 * bass_lift_counted and bass_counted_read.
 * in_flags bit 3 (with bit 0): (x, a) is the output of a bass_lift(mode | 32)
 * issued just before on this workspace; its digit-1 pass is used when it
 * predicted this call's bucket (stats.flags bit 2 set).  Every call records
 * its bucket as the next lift's prediction.
 * key_map (key_bits entries, or NULL): stored bit p of x holds qubit
 * key_map[p]; ties are then ordered by the key mapped back. */
int bass_truncate(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo,
                  double* ao, int64_t cap_out, int renorm, int64_t hist_tiles, int in_flags,
                  const int32_t* key_map, int key_bits, void* ws, size_t ws_bytes, void* stream);

/* bass_lift of (x, a) into the segments (seg_x, seg_a) followed by
 * bass_truncate of them into (xo, ao), in one call: the results (xo[:nnz],
 * ao, the stats block, the op log and state-count slot in counted mode, the
 * next prediction) are bit for bit those of
 *   bass_lift(x, a, n, nq, j, G, mode | 16 [| 32 with flags bit 1], seg_x, seg_a, seg_cap, ...)
 *   bass_truncate(seg_x, seg_a, 0, k, xo, ao, cap_out, renorm, seg_tiles,
 *                 1 [| 4 counted] [| 8 with flags bit 1], key_map, key_bits, ...).
 * With flags bit 1 (the speculative digit-1 pass, against the bucket of the
 * workspace's last select) the lift also prunes: it stores only the outputs
 * whose |a|^2 lies in the predicted threshold's IEEE-exponent bucket, the one
 * below it or higher -- the others cannot be kept or be candidates unless
 * the threshold falls further -- while its histogram and sums still cover
 * every output (flags bit 2: only the predicted bucket or higher -- cheaper
 * when the threshold rarely falls, a relift whenever it does); when the
 * select's bucket does fall further, or the select
 * keeps everything and something was dropped, a relift stores every output
 * before the select reads them (stats.flags bit 3).  Replaces, inside the
 * driver, a gate step of src/simulation/bass_simulator.py:861-896 with its
 * truncation at :899-916, and a rotation trial's rotate + truncate at
 * :768-790 (no reference counterpart as one entry point).  flags bit 0:
 * counted (n bounds the count read from the state-count slot).  (xo, ao) may
 * be (x, a); the segments (seg_cap >= seg_tiles * 2048 slots) may overlap
 * neither. */
int bass_lift_truncate(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                       int64_t k, uint64_t* seg_x, double* seg_a, int64_t seg_cap, uint64_t* xo, double* ao,
                       int64_t cap_out, int renorm, int flags, const int32_t* key_map, int key_bits, void* ws,
                       size_t ws_bytes, void* stream);

/* A gate or rotation on a key-ordered state followed by the top-k truncation
 * of its output, without storing that output: the results (xo[:nnz], ao,
 * the stats block, the op log and state-count slot in counted mode, the
 * next speculative prediction) are those of
 *   bass_lift(x, a, n, nq, j, G, mode | 16 [| 32 with flags bit 1], tmp, ...)
 *   bass_truncate(tmp, ..., k, xo, ao, cap_out, renorm, seg_tiles,
 *                 1 [| 4 counted] [| 8 with flags bit 1], key_map, key_bits, ...)
 * bit for bit.  The lift keeps only what the truncation reads before its
 * compaction (digit-1 histogram, per-tile sums and counts, the candidates of
 * the predicted |a|^2 bucket with their keys); the truncation recomputes the
 * lift's tiles to write the kept entries (and for its digit-1 pass when the
 * prediction missed), so the gate's 2^nq n outputs never go to memory.
 * Replaces, inside the driver, a gate step of
 * src/simulation/bass_simulator.py:861-896 with its truncation at :899-916,
 * and a rotation trial's rotate + truncate at :768-790 (no reference
 * counterpart as one entry point).
 * mode: 0 gate, 1 rotation (nq and j as in bass_lift).  flags bit 0
 * (counted): n bounds the count read from the state-count slot, as
 * bass_lift_counted + a counted bass_truncate.  flags bit 1: the speculative
 * digit-1 pass against the last select's bucket (as mode | 32 / in_flags 8).
 * (xo, ao) must not overlap (x, a).  The workspace must hold the sizes that
 * bass_lift + bass_truncate need (seg_tiles * 2048 slots). */
int bass_gate_truncate(const uint64_t* x, const double* a, int64_t n, int nq, int j, const double* G, int mode,
                       int64_t k, uint64_t* xo, double* ao, int64_t cap_out, int renorm, int flags,
                       const int32_t* key_map, int key_bits, void* ws, size_t ws_bytes, void* stream);

/* bass_truncate of a dense input (hist_tiles = 0, in_flags = 0) that
 * allocates nothing and takes no workspace: its scratch (select histograms,
 * candidate lists, per-tile sums, stats) is a static device array of the
 * library, one per GPU, sized for n <= 2^20 entries (72 MiB, held until the
 * process ends).  n above that fails with BASS_E_ARG.  stats_out: a device
 * pointer (caller-owned) that receives the call's bass_stats, written in
 * stream order; nothing is read back to the host.  Calls on one GPU share
 * the array, so calls that overlap on one GPU (different streams or
 * threads) race; each call clears what it needs zeroed.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int bass_truncate_noalloc(const uint64_t* x, const double* a, int64_t n, int64_t k, uint64_t* xo,
                          double* ao, int64_t cap_out, int renorm, const int32_t* key_map, int key_bits,
                          bass_stats* stats_out, void* stream);

/* In place: a[i] *= diag[(b_q1 << 1) | b_q2]; diag = 4 complex (8 doubles).
 * Replaces apply_2qubit_diagonal_kernel (src/core/sparse_state.py:178-190). */
int bass_diag2q(const uint64_t* x, double* a, int64_t n, int q1, int q2, const double* diag,
                void* stream);

/* stats.sum2 = sum |a|^2, stats.sum4 = sum |a|^4 (fixed-order reduction).
 * Serves _renorm / _pr (src/simulation/bass_simulator.py:551-554, :574-578). */
int bass_moments(const double* a, int64_t n, void* ws, size_t ws_bytes, void* stream);

/* In place a[i] /= div (componentwise).  Serves _renorm. */
int bass_scale(double* a, int64_t n, double div, void* stream);

/* Gather (xo[j], ao[j]) = (x[idx[j]], a[idx[j]]), idx: int64 device array. */
int bass_gather(const uint64_t* x, const double* a, const int64_t* idx, int64_t m, uint64_t* xo,
                double* ao, void* stream);

/* All single-qubit RDMs.  Synchronous: rho_out (host) receives N x 4 doubles
 * (rho00, rho11, Re rho01, Im rho01) per qubit.  Replaces ht_build +
 * compute_all_1qubit_rdms (src/simulation/bass_simulator.py:108-123, :441-470). */
int bass_rdm1q(const uint64_t* x, const double* a, int64_t n, int N, double* rho_out, void* ws,
               size_t ws_bytes, void* stream);

/* Hash table of keys -> int32 index, capacity 2^log2cap slots, EMPTY =
 * 0xFFFFFFFFFFFFFFFF.  Replaces ht_clear / ht_build / ht_insert / ht_lookup
 * (src/simulation/bass_simulator.py:100-153) with batched device calls.
 * The slot hash differs from the reference's key & (cap-1); lookups return the
 * same indices. */
/* Single-qubit RDMs of a key-ordered state (x sorted ascending by stored
 * word; see bass_lift), one per stored bit p < N, same output layout as
 * bass_rdm1q.  Pairs (x, x | 2^p) are found by the merge that lifts bit p, so
 * no hash table is built.  Replaces compute_all_1qubit_rdms
 * (src/simulation/bass_simulator.py:441-470) inside the driver.  Synchronous. */
int bass_rdm1q_ordered(const uint64_t* x, const double* a, int64_t n, int N, double* rho_out, void* ws,
                       size_t ws_bytes, void* stream);

/* Two-qubit RDM of a key-ordered state for the pair (first qubit stored at
 * bit j >= 1, second at bit 0): rho[p][q] = sum over entries equal outside the
 * pair of a_p conj(a_q), index (b_first << 1) | b_second.  rho_out: 4x4
 * complex, row-major, 32 doubles (re, im).  The pair's groups are made
 * adjacent by the merge that lifts bit j to bit 1, as for a 2-qubit gate.
 * Replaces compute_2qubit_rdm (src/simulation/bass_simulator.py:202-234) in
 * the 2-qubit pass (:679-728).  Synchronous. */
int bass_rdm2q_ordered(const uint64_t* x, const double* a, int64_t n, int j, double* rho_out, void* ws,
                       size_t ws_bytes, void* stream);

/* ---- dense state vectors (exact simulation, fidelities) ---------------
 * psi: complex128[2^N] on the device, index = basis key. */

/* psi = 0, then psi[key(x_i)] = a_i for the n entries; key_map (key_bits = N
 * entries, stored bit -> qubit) maps the stored words of a key-ordered state
 * back to keys, key_bits = 0 takes x as keys.  The scatter half of
 * BASS.to_statevector (src/simulation/bass_simulator.py:608-621). */
int bass_dense_scatter(const uint64_t* x, const double* a, int64_t n, const int32_t* key_map, int key_bits,
                       double* psi, int N, void* stream);

/* In place: 1- or 2-qubit gate G (row-major complex, index (b_q0 << 1) | b_q1)
 * on a dense vector.  Replaces ExactSimulator._apply_1qubit_dense /
 * _apply_2qubit_dense (src/simulation/exact_simulator.py:109-144) and the
 * per-qubit frame rotation of BASS.to_statevector. */
int bass_dense_gate(double* psi, int N, int nq, int q0, int q1, const double* G, void* stream);

/* out[0..1] = sum_i conj(u_i) v_i (np.vdot), fixed-order reduction.  For
 * fidelity_bass (src/experimentation/runner.py:620-626) and compute_fidelity
 * (src/benchmarking/fidelity.py:48-65).  Synchronous. */
int bass_dense_vdot(const double* u, const double* v, int64_t len, double* out, void* ws, size_t ws_bytes,
                    void* stream);

/* ---- batches: many independent key-ordered states per launch ----------
 * Each call runs one launch sequence for all njobs states (per-job scalars via
 * bass_batch_stats).  Job tables are staged through pinned host buffers.  A
 * job's count cells (cnt_in / cnt_out) keep state sizes on the device, so
 * consecutive gates need no host read; the geometry of every launch is then
 * derived on the device.
 *
 * bass_batch_step: for every job, the gate (mode 0) or rotation (mode 1) of
 * bass_lift on (x, a, n, j, g), then bass_truncate of its output into (xo, ao)
 * with k / renorm / key_map, keeping all entries when there are <= k (stats
 * flags bit 0 says whether entries were dropped).  tiles_x / tiles_a (capacity
 * tiles_cap) hold the gate outputs: sum over jobs of ceil(n / (2048 >> nq)) *
 * 2048 slots.  wide=1: multi-block threshold resolution (repeat a call whose
 * stats report flags bit 1).  Replaces, per circuit, the gate + truncation
 * step of src/simulation/bass_simulator.py:861-916 and the tentative rotation
 * of :768-783. */
int bass_batch_step(const bass_job* jobs, int njobs, int nq, int mode, int wide, uint64_t* tiles_x, double* tiles_a,
                    int64_t tiles_cap, void* ws, size_t ws_bytes, void* stream);

/* bass_lift(mode=2) per job: move stored bit j to bit 0, dense output in (xo, ao). */
int bass_batch_move(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream);

/* Per job in place: diagonal 2-qubit gate on stored bits (j, j2) with the 4
 * values g[0..7] (src/core/sparse_state.py:178-190); division by `scale`
 * (:551-554); moments sum |a|^2, sum |a|^4 into the job's stats. */
int bass_batch_diag2q(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream);
int bass_batch_scale(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream);
int bass_batch_moments(const bass_job* jobs, int njobs, void* ws, size_t ws_bytes, void* stream);

/* Single-qubit RDMs of every stored bit p < key_bits of every job (as
 * bass_rdm1q_ordered); rho_out: njobs x 64 x 4 doubles.  Synchronous. */
int bass_batch_rdm1q(const bass_job* jobs, int njobs, double* rho_out, void* ws, size_t ws_bytes, void* stream);

/* Synchronous: the per-job stats of the last batch call (njobs entries);
 * clears device error bits.  BASS_E_DEVICE if any job reports one. */
int bass_batch_stats(void* ws, size_t ws_bytes, int njobs, bass_stats* out, void* stream);

/* ---- batch plans: fixed launch shapes, captured in HIP graphs ---------
 * No counterpart in the reference, which runs circuits one after another
 * (src/experimentation/runner.py:697-874 calling
 * src/simulation/bass_simulator.py:847-929); the operations per circuit are
 * those of the bass_batch_* calls above.
 *
 * A plan fixes the workspace regions for up to jobs_cap states and tiles_cap
 * gate-output tiles (tiles_x / tiles_a hold tiles_cap * 2048 slots), so each
 * operation's launch sequence has one shape: a prep kernel reads the input
 * counts (count cells or staged n) and lays out the tiles on the device, and
 * every kernel after it runs a persistent grid over the totals it wrote.
 * flags bit 0: capture each sequence once into a HIP graph and replay it. */
enum {
  BASS_BOP_GATE1 = 0, /* bass_batch_step nq=1 mode=0 */
  BASS_BOP_GATE2 = 1, /* nq=2 mode=0 */
  BASS_BOP_ROT1 = 2,  /* nq=1 mode=1 */
  BASS_BOP_ROT2 = 3,  /* nq=2 mode=1 */
  BASS_BOP_MOVE = 4,  /* bass_batch_move */
  BASS_BOP_DIAG2Q = 5,
  BASS_BOP_SCALE = 6,
  BASS_BOP_MOMENTS = 7, /* staged counts only */
  BASS_BOP_RDM1Q = 8,   /* staged counts only */
  BASS_BOP_GATE12 = 9   /* gates of 1 and 2 qubits (bass_job.nq) in one lift pass pair and one select */
};
size_t bass_bplan_ws_bytes(int jobs_cap, int64_t tiles_cap);
int bass_bplan_create(int jobs_cap, int64_t tiles_cap, uint64_t* tiles_x, double* tiles_a, void* ws,
                      size_t ws_bytes, int flags, void** plan);
int bass_bplan_destroy(void* plan);
/* One batch operation, asynchronous on the stream.  out_host (NULL: none),
 * written when the stream reaches the end of the call: steps and moments,
 * njobs bass_stats; RDM1Q, njobs x 64 x 4 doubles as bass_batch_rdm1q.
 * Pinned host memory keeps the copy asynchronous.  The caller checks the
 * stats' err fields. */
int bass_bplan_run(void* plan, int op, const bass_job* jobs, int njobs, void* out_host, void* stream);

int bass_ht_clear(uint64_t* ht_keys, int32_t* ht_vals, int log2cap, void* stream);
int bass_ht_insert(uint64_t* ht_keys, int32_t* ht_vals, int log2cap, const uint64_t* keys,
                   int64_t n, int64_t index_base, void* ws, size_t ws_bytes, void* stream);
int bass_ht_lookup(const uint64_t* ht_keys, const int32_t* ht_vals, int log2cap,
                   const uint64_t* keys, int64_t n, int64_t* out, void* stream);

/* Stable LSD radix sort of keys[0:n) (low key_bits bits significant, the rest
 * must be zero) carrying complex128 payload.  Replaces radix_sort_uint64
 * (src/core/sparse_state.py:150-172), which takes 4 passes of 16 bits; here
 * one histogram pass plus one pass per 8-bit digit (plus a copy back after an
 * odd pass count).  Uses ws as the ping-pong buffer.  n < 2^30 (BASS_E_ARG
 * otherwise). */
int bass_radix_sort(uint64_t* keys, double* vals, int64_t n, int key_bits, void* ws,
                    size_t ws_bytes, void* stream);

/* Digit width of bass_radix_sort: 10 selects 10-bit digits (fewer passes,
 * slower ones on MI355X), 8 or 0 the default 8-bit digits.  Returns the
 * previous setting.  Process-wide tuning knob; no counterpart in the reference. */
int bass_sort_digit_bits(int bits);

/* Per-size-class launch tuning (no counterpart in the reference): knob 0 =
 * hash / group-table capacity factor (2 or 4), 1 = items per thread of 1-qubit
 * emission blocks (2, 4, 8), 2 = the same for 2-qubit gates (1, 2, 4), 3 =
 * radix digit width (8 or 10); size_class 0..3 for input counts < 2^14,
 * < 2^18, < 2^22 and larger; value 0 = the built-in choice.  Output keys,
 * amplitudes and kept sums do not depend on any knob; the emission tile size
 * sets the grouping of a dense gate output's input total (summed per tile,
 * then over tiles in order), which may move its last bits.  Returns the
 * previous value, or BASS_E_ARG.  Process-wide. */
int bass_tune(int knob, int size_class, int value);

/* The plan bass_radix_sort follows for keys of key_bits bits: returns the
 * digit width (8 or 10) and stores the number of digit passes in *passes (an
 * odd count adds one copy back into keys / vals).  No counterpart in the
 * reference. */
int bass_sort_plan(int key_bits, int* passes);

/* Run-length sum of equal adjacent keys of a sorted array into (xo, ao); zero
 * sums are kept.  Replaces merge_sorted_kernel (src/core/sparse_state.py:127-147). */
int bass_merge_sorted(const uint64_t* x, const double* a, int64_t n, uint64_t* xo, double* ao,
                      void* ws, size_t ws_bytes, void* stream);

/* Measurement: bass_timing(1) resets and enables per-operation timing with HIP
 * events recorded on each call's stream around its kernels ("gate1q", "gate2q",
 * "rot1q", "rot2q", "emit1q", "emit2q", "select", "diag2q", "rdm", "sort");
 * bass_timing_read waits for the recorded events and returns the launch count
 * and summed milliseconds of one operation.  No reference counterpart (the
 * reference times with perf_counter around simulate(),
 * src/experimentation/runner.py:833-846). */
int bass_timing(int enable);
int bass_timing_read(const char* name, int64_t* count, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* BASS_HIP_H */
