/*
 * qsim_mi355x.h - C ABI of the MI355X (gfx950) state-vector library.
 *
 * The library replaces the array arithmetic that qibo's NumpyBackend performs
 * for the state-vector hot path (qibo 0.3.5, src/qibo/backends/abstract.py).
 * Each entry point names the reference method it stands in for.
 *
 * Conventions shared by every entry point:
 *   - All state / probability / index buffers are DEVICE pointers owned by the
 *     caller (the Python host keeps them in torch tensors). Small inputs that
 *     are marked "host" (qubit lists, gate matrices, term tables) are read
 *     during the call and copied into kernel arguments.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     Every call only enqueues work on `stream` and returns; no call
 *     synchronises the device unless its comment says so.
 *   - Qubit numbering is qibo's: qubit q of an n-qubit state is bit
 *     (n-1-q) of the amplitude index (qubit 0 is the most significant bit).
 *   - Matrices are row-major complex (interleaved re, im) of the state dtype.
 *     A matrix acting on targets (t0, t1, ...) indexes its rows/columns with t0
 *     as the most significant bit, as gate.qubits orders them in qibo.
 *   - Return value: 0 on success, a negative QSIM_E* code otherwise; the
 *     message of the last failure on the calling thread is qsim_last_error().
 *   - Every entry point is re-entrant (qibo's parallel.py calls
 *     execute_circuit from threads). State shared between calls: the table of
 *     run-time specialised kernels (qsim_blocked_specialize; guarded by a
 *     mutex) and the device array of qsim_expval_pauli_noalloc (calls that
 *     overlap on one GPU race, as its comment says).
 *
 * Where this ABI departs from a plain one-call-per-reference-method list:
 *   - controlled gates are qsim_apply_dense / qsim_apply_diag / qsim_apply_swap
 *     with a control list (no separate "apply_controlled" entry);
 *   - sampling is qsim_sample_exact (draws from probabilities in one call,
 *     with np.random.choice's checks), plus qsim_cumsum_exact and
 *     qsim_search_sorted for callers that need the cumulative sum itself;
 *   - precision is the state dtype: complex64 states run FP32 (MFMA for fused
 *     blocks), complex128 FP64; qsim_compare_states measures the accuracy of
 *     one against the other (no "precision_mode" argument);
 *   - gradients are qsim_adjoint_blocked + qsim_adjoint_reduce over batched
 *     blocked passes, with qsim_apply_pauli_sum / qsim_inner_matrix for the
 *     per-set sweep;
 *   - serving is the Python executor (qibo_amd/serving.py) over these entry
 *     points (batched launches, HIP graphs of them); there is no separate
 *     serving entry point.
 */
#ifndef QSIM_MI355X_H
#define QSIM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSIM_ABI_VERSION 11

/* state dtypes */
#define QSIM_C128 0 /* complex128: double re, double im */
#define QSIM_C64 1  /* complex64:  float re, float im   */

/* probability dtypes */
#define QSIM_F64 0
#define QSIM_F32 1

/* error codes */
#define QSIM_OK 0
#define QSIM_EINVAL -1
#define QSIM_ENOMEM -2
#define QSIM_EHIP -3

/* Largest qubit count any entry point accepts. */
#define QSIM_MAX_QUBITS 36

/* ABI version of the loaded library (QSIM_ABI_VERSION it was built with). */
int qsim_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* qsim_last_error(void);

/* |0...0>: zero the 2^n amplitudes and set amplitude 0 to 1.
 * Replaces Backend.zero_state (abstract.py:2243-2275). */
int qsim_zero_state(void* state, int nqubits, int dtype, void* stream);

/* Zeros at the amplitudes whose index has a 1 outside `support` (bit b of the
 * mask = index bit b); the others are neither read nor written. Completes a
 * state that passes from |0...0> wrote only on their support (see
 * qsim_apply_blocked_support). No counterpart in the reference. */
int qsim_zero_outside_support(void* state, int nqubits, int dtype, uint64_t support, void* stream);

/* Dense 2^k x 2^k matrix on `targets` (k = ntargets, 1..3), applied only on
 * the amplitudes whose `controls` bits are all 1, in place.
 * Replaces Backend.apply_gate's transpose/matmul/transpose for dense gates
 * (abstract.py:2350-2361) and Backend._apply_gate_controlled_by
 * (abstract.py:3176-3197). `matrix` is a host pointer. */
int qsim_apply_dense(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                     const int* controls, int ncontrols, const void* matrix, void* stream);

/* Collapse onto the outcome `outcome` of measuring `qubits` (qibo order: bit
 * j of the outcome, most significant first, is qubits[j]): amplitudes whose
 * measured bits equal the outcome are divided by `norm`, all others set to 0.
 * Replaces Backend._collapse_statevector (abstract.py:3279-3304); the caller
 * passes norm = sqrt(probability of the outcome), or 1 without normalising. */
int qsim_collapse(void* state, int nqubits, int dtype, const int* qubits, int nmeasured, int64_t outcome,
                  double norm, void* stream);

/* Diagonal matrix (2^k complex entries, k = ntargets, 0..5) on `targets`,
 * restricted to amplitudes whose `controls` bits are all 1. k = 0 multiplies
 * the controlled subspace by diag[0]. Same reference as qsim_apply_dense;
 * used for gates whose matrix is diagonal (RZ, U1, CZ, CU1, ...), which
 * touch only the amplitudes they change. `diag` is a host pointer. */
int qsim_apply_diag(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                    const int* controls, int ncontrols, const void* diag, void* stream);

/* Exchange of the |..0..1..> and |..1..0..> amplitudes of qubits (q0, q1)
 * under `controls`: the SWAP matrix (npmatrices.py:264-268) applied by
 * touching only the two amplitudes it moves. */
int qsim_apply_swap(void* state, int nqubits, int dtype, int q0, int q1, const int* controls,
                    int ncontrols, void* stream);

/* Dense 2^k x 2^k block on k = ntargets (1..5) targets with the matrix in
 * DEVICE memory: the application of a FusedGate (abstract.py:2680-2717 builds
 * the matrix; apply_gate applies it). complex128 blocks of 4-5 qubits
 * (nqubits >= 11, 2048-amplitude LDS tiles): v_mfma_f64_16x16x4f64 with FP64
 * accumulation, as three real products (Re M Re X, Im M Im X,
 * (Re M + Im M)(Re X + Im X)); complex64
 * blocks of 4-5 qubits (nqubits >= 12, 4096-amplitude LDS tiles):
 * v_mfma_f32_32x32x2f32 with FP32 accumulation, on the same three real
 * products for 5 qubits and on the real embedding [[Re, -Im], [Im, Re]] for
 * 4; smaller blocks:
 * FMA per column in the state's precision. */
int qsim_apply_fused(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                     const void* matrix_dev, void* stream);

/* Many gates in one pass over the state: the gate loop of
 * Backend._execute_circuit (abstract.py:3321-3322, each step an apply_gate,
 * abstract.py:2322-2361) for a run of consecutive gates that the host has
 * grouped so that their non-diagonal targets fit in 12 "tile bits".
 * `tile_bits` = 12 ascending amplitude-index bit positions (not qubit
 * numbers). `program_dev` = DEVICE buffer (16-byte aligned) of
 * `program_bytes` bytes holding `ninstr` variable-length records: a 32-byte
 * header (u8 op, u8 arg, u8 nt, u8 register-condition mask, u32 number of
 * complex entries, u64 condition mask, u8 bits[8], u64 zero) followed by the
 * entries, padded to 16 bytes. Ops: 0 diagonal on up to 5 bits, 1 complex 2x2
 * on a register bit of the current mapping, 2 change of register mapping,
 * 3 dense 2^k block on k <= 3 local tile bits, 4 real 2x2 on a register bit,
 * 5 16-entry register phase table, 6 per-lane diagonal factor, 7 sign flip,
 * 10 unconditional rotation / reflection on a register bit as three shears
 * (arg = bit | mode << 2, one entry (t, s); mode 0 [[c,-s],[s,c]],
 * 1 [[c,s],[s,-c]], 2 [[c,-is],[-is,c]], with t = s / (1 + c), c >= 0),
 * 11 controlled phase on register bit arg under lane condition bits, held in
 * a per-lane factor, 12 application of that factor to the registers with bit
 * arg set (a program must place 12 before any non-diagonal record on that
 * bit, an op-11 record on another bit, ops 2 and 3, and its end).
 * The buffer must stay readable for program_bytes + QSIM_BLOCKED_PAD bytes
 * (the kernel reads one header ahead and reads at most 64 entries of a
 * record, so no read leaves that range); records are read while they start
 * inside program_bytes, which must be below 2^31. A program of ops 0-3 only
 * gives bitwise the state of applying the same gates one by one with
 * qsim_apply_dense / qsim_apply_diag / qsim_apply_swap; the other ops
 * reorder roundings. Requires nqubits >= 12. */
#define QSIM_BLOCKED_PAD 1056
size_t qsim_blocked_header_bytes(void);
int qsim_apply_blocked(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                       size_t program_bytes, int ninstr, void* stream);

/* qsim_apply_blocked with the input taken to be |0...0>: the buffer's old
 * contents are never read, only the result is written. The first pass of
 * Backend._execute_circuit from zero_state (abstract.py:2243-2275 then
 * 3321-3322). |0...0> is nonzero only in tile 0 and the pass acts on its
 * tile bits only, so for 22 qubits and more the call writes zeros everywhere
 * (qsim_zero_state) and runs the pass on tile 0 alone; below, it runs every
 * tile from zeros generated in registers (one launch). */
int qsim_apply_blocked_from_zero(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                                 size_t program_bytes, int ninstr, void* stream);

/* qsim_apply_blocked for an input known to vanish wherever an index bit
 * outside `support` is 1 (bit b of the mask = index bit b, i.e. qubit
 * n-1-b): the pass runs only the tiles whose bits outside the tile bits lie
 * in `support`, and leaves the others, which stay zero, untouched; inside
 * those tiles it reads only the amplitudes in the support (the others are
 * taken as zero, whatever their memory holds) and writes every amplitude.
 * A state built by passes from |0...0> has the union of their tile bits as
 * its support, and need be written only there: one amplitude set to 1, then
 * passes with the growing support, then qsim_zero_outside_support unless a
 * pass has covered every tile (no counterpart in the reference: it skips
 * amplitudes that are zero; all ones is qsim_apply_blocked). */
int qsim_apply_blocked_support(void* state, int nqubits, int dtype, const int* tile_bits, uint64_t support,
                               const void* program_dev, size_t program_bytes, int ninstr, void* stream);

/* Blocked passes specialised to their record sequence (no counterpart in the
 * reference: a faster form of qsim_apply_blocked / _from_zero / _batched and
 * qsim_adjoint_blocked, whose results it reproduces bitwise). The library
 * emits one straight-line kernel per sequence of records -- opcode, argument,
 * entry count and whether a tile condition is present -- with every other
 * field (angles, masks, entries) read from the program at run time, compiles
 * it with hiprtc for gfx950 and loads it into the current device (1-3 s per
 * pass on the host, once per sequence and process). `program` is the host
 * copy of one program (for a batch: any one of its programs, which share the
 * sequence). `variant`: QSIM_SPEC_PASS, QSIM_SPEC_PASS_FROM_ZERO (the input
 * taken to be |0...0>) or QSIM_SPEC_ADJOINT (the adjoint sweep).
 * qsim_blocked_specialize compiles (or finds) the kernel and returns its
 * handle; qsim_blocked_specialized_lookup only finds it (*handle = 0 when the
 * sequence has no kernel yet) and gives a 64-bit hash of the sequence.
 * qsim_apply_blocked_specialized launches a pass kernel with the arguments
 * and contract of qsim_apply_blocked_batched (nbatch = 1 for one state) and
 * the `support` of qsim_apply_blocked_support (all ones: every tile; a
 * from-zero kernel writes |0...0> first and ignores it, as
 * qsim_apply_blocked_from_zero does);
 * qsim_adjoint_blocked_specialized an adjoint kernel with those of
 * qsim_adjoint_blocked. The device programs must have the kernel's record
 * sequence (checked only through their record count and length: a program
 * of another sequence gives wrong amplitudes). A sequence with a conditional
 * mapping change, or whose estimated compile time exceeds 10 s (long
 * sequences of 2x2, diagonal and inner-product records), is not specialised
 * by qsim_blocked_specialize / _lookup (QSIM_EINVAL).
 * qsim_blocked_specialized_source returns the kernel's source (length + 1;
 * written to out when cap > 0) and qsim_blocked_compile_check compiles it
 * without a device (both for tests and inspection). */
#define QSIM_SPEC_PASS 0
#define QSIM_SPEC_PASS_FROM_ZERO 1
#define QSIM_SPEC_ADJOINT 2
int qsim_blocked_specialize(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                            uint64_t* handle);
int qsim_blocked_specialized_lookup(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                                    uint64_t* handle, uint64_t* key_hash);
int qsim_apply_blocked_specialized(uint64_t handle, void* states, int nbatch, int nqubits, int dtype,
                                   const int* tile_bits, uint64_t support, const void* programs_dev,
                                   size_t program_stride, size_t program_bytes, int ninstr, void* stream);
int qsim_adjoint_blocked_specialized(uint64_t handle, void* psi, void* lam, int nbatch, int nqubits, int dtype,
                                     const int* tile_bits, const void* programs_dev, size_t program_stride,
                                     size_t program_bytes, int ninstr, int nslots, double* partials,
                                     size_t partials_count, void* stream);
size_t qsim_blocked_specialized_source(const void* program, size_t program_bytes, int ninstr, int dtype,
                                       int variant, char* out, size_t cap);
int qsim_blocked_compile_check(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                               size_t* code_bytes);
size_t qsim_blocked_specialized_count(void);

/* The blocked pass for a batch of `nbatch` states stored one after the other
 * (2^nqubits amplitudes apart), state b running the program at
 * programs_dev + b * program_stride (all programs program_bytes long, same
 * padding rule). Runs the batch in one launch (grid z = batch entry): the
 * forward pass of the batched adjoint gradients. No reference counterpart
 * (the reference runs one circuit per call). */
int qsim_apply_blocked_batched(void* states, int nbatch, int nqubits, int dtype, const int* tile_bits,
                               const void* programs_dev, size_t program_stride, size_t program_bytes, int ninstr,
                               void* stream);

/* Adjoint sweep pass for a batch: the records (ops 0-2, 4, 6) are applied to
 * both psi and lam of every batch entry; op 8 (2x2 M on a register bit) and
 * op 9 (diagonal M on up to 5 bits) add Re <lam| M |psi> of each tile into
 * `partials` (qsim_adjoint_partials_count(nqubits, nbatch, nslots) doubles,
 * nslots = INNER records per program, in program order). The adjoint method
 * replaces the reference's parameter shift / finite differences
 * (derivative.py:7-226), which re-simulate per parameter. */
size_t qsim_adjoint_partials_count(int nqubits, int nbatch, int nslots);
int qsim_adjoint_blocked(void* psi, void* lam, int nbatch, int nqubits, int dtype, const int* tile_bits,
                         const void* programs_dev, size_t program_stride, size_t program_bytes, int ninstr,
                         int nslots, double* partials, size_t partials_count, void* stream);

/* grads[b * nparams + param_of[s]] = 2 * (sum of the partials of INNER
 * record s of batch entry b), summed in a fixed order. `partials` holds
 * `partials_count` doubles, at least qsim_adjoint_partials_count(nqubits,
 * nbatch, nslots) (the buffer qsim_adjoint_blocked filled). `param_of_dev` is
 * a DEVICE int array of nslots entries (-1: not stored). */
int qsim_adjoint_reduce(const double* partials, size_t partials_count, int nbatch, int nqubits, int nslots,
                        const int* param_of_dev, int nparams, double* grads, void* stream);

/* Probabilities of measuring `qubits` (m = nmeasured, in the given order, the
 * first being the most significant bit of the output index):
 *   p = abs(state)**2 with numpy's complex abs, summed over unmeasured qubits.
 * Replaces Backend.calculate_probabilities + _order_probabilities
 * (abstract.py:2734-2758, 3371-3381). Output dtype is F64 for C128 states and
 * F32 for C64 states (the state's real dtype, as `rtype` there). When all
 * qubits are measured the values are bitwise those of numpy's
 * `np.abs(state)**2`; marginals are bitwise numpy's
 * `np.sum(probs, axis=unmeasured)` too (its order: pairwise sums of the
 * trailing unmeasured runs in 8192-element buffers, then C order; elementwise
 * C order when the last qubit is measured), in the probabilities' dtype,
 * except when an output would need more than 2^24 serial additions, where a
 * parallel fixed-order FP64 reduction is used.
 * `workspace` must hold qsim_probabilities_workspace(nqubits, nmeasured) bytes. */
size_t qsim_probabilities_workspace(int nqubits, int nmeasured);
int qsim_probabilities(const void* state, int nqubits, int dtype, const int* qubits, int nmeasured,
                       void* probs, void* workspace, size_t workspace_bytes, void* stream);

/* Inclusive prefix sum of n probabilities (F64, or F32 widened to F64) that is
 * bitwise equal to the sequential left-to-right sum numpy's `cumsum` performs
 * inside np.random.choice (called from Backend.sample_shots via random_choice,
 * abstract.py:2774-2781, 1265-1309). Writes cdf[0..n-1] (float64).
 * Probabilities must be finite and >= 0.
 * `workspace` must hold qsim_cumsum_workspace(n) bytes. */
size_t qsim_cumsum_workspace(int64_t n);
/* Checks np.random.choice makes on p before drawing (mtrand choice: finite,
 * non-negative, sums to 1 within sqrt(eps)), in one pass: out[0] = sum of p
 * (FP64, fixed order), out[1] = number of negative entries, out[2] = number
 * of NaNs (device pointer, 3 doubles). Used by sample_shots /
 * sample_frequencies (abstract.py:2760-2781) before their draws.
 * `workspace` must hold qsim_prob_stats_workspace(n) bytes. */
size_t qsim_prob_stats_workspace(int64_t n);
int qsim_prob_stats(const void* probs, int prob_dtype, int64_t n, double* out, void* workspace, size_t workspace_bytes,
                    void* stream);
int qsim_cumsum_exact(const void* probs, int prob_dtype, int64_t n, double* cdf,
                      void* workspace, size_t workspace_bytes, void* stream);

/* numpy's np.sum of a 1-D array of n = 2^m probabilities (F64 or F32, the
 * result in the same dtype, one device value), bit for bit: pairwise sums of
 * consecutive 8192-element buffers accumulated in order from 0. The
 * normalisation `probabilities / sum(probabilities)` of
 * Backend.sample_frequencies (abstract.py:2762) uses it, followed by
 * qsim_divide_by (out[i] = probs[i] / divisor[0], correctly rounded).
 * `workspace` must hold qsim_np_sum_workspace(n) bytes. */
size_t qsim_np_sum_workspace(int64_t n);
int qsim_np_sum(const void* probs, int prob_dtype, int64_t n, void* out, void* workspace, size_t workspace_bytes,
                void* stream);
int qsim_divide_by(const void* probs, int prob_dtype, int64_t n, const void* divisor, void* out, void* stream);

/* Inverse-CDF draw: out[k] = number of i with fl(cdf[i] / cdf[n-1]) <= u[k],
 * i.e. numpy's `(cdf / cdf[-1]).searchsorted(u, side="right")` on the
 * normalised cumulative sum, for nshots device uniforms u in [0, 1). */
int qsim_search_sorted(const double* cdf, int64_t n, const double* uniforms, int64_t nshots,
                       int64_t* out, void* stream);

/* Backend.sample_shots (abstract.py:2774-2781: random_choice -> numpy's
 * np.random.choice(n, nshots, p=probs)) in one call, without writing the
 * cumulative sum: out[k] equals qsim_search_sorted on qsim_cumsum_exact's cdf
 * for the same uniforms (bitwise: the same exact sequential sums, read where
 * the bisection compares them). stats (device, 3 doubles) receives the
 * checks of qsim_prob_stats: sum, number of negative entries, number of
 * NaNs. Probabilities with a negative entry, a NaN or an infinite sum are not
 * sampled (out is filled with 0): the caller raises as np.random.choice does.
 * `workspace` must hold qsim_sample_workspace(n) bytes (about 9 bytes per
 * probability). */
size_t qsim_sample_workspace(int64_t n);
int qsim_sample_exact(const void* probs, int prob_dtype, int64_t n, const void* tile_stats, const double* uniforms,
                      int64_t nshots, int64_t* out, double* stats, void* workspace, size_t workspace_bytes,
                      void* stream);
/* Probabilities of all qubits in order, as qsim_probabilities with
 * qubits = 0..nqubits-1 (Backend.calculate_probabilities,
 * abstract.py:2734-2758), and in the same pass the per-4096-entry sums and
 * counts of negative/NaN entries that qsim_sample_exact computes first: pass
 * `tile_stats` (qsim_tile_stats_bytes(2^nqubits) bytes) to qsim_sample_exact
 * on these probabilities to skip that pass (null: computed there). */
size_t qsim_tile_stats_bytes(int64_t n);
int qsim_probabilities_tiles(const void* state, int nqubits, int dtype, void* probs, void* tile_stats,
                             size_t tile_stats_bytes, void* stream);

/* out[s, b] = (samples[s] >> (nbits-1-b)) & 1, as Backend.samples_to_binary
 * (abstract.py:2783-2786). out is int64 [nshots, nbits]. */
int qsim_samples_to_binary(const int64_t* samples, int64_t nshots, int nbits, int64_t* out,
                           void* stream);

/* Readout bitflips on binary samples [nshots, nbits] (int64 0/1, row-major):
 * out = s + (1 - s) [u < p0[col]] - s [u < p1[col]] with `uniforms` a device
 * array of nshots * nbits doubles in the same order, drawn on the host by
 * np.random.random_sample((nshots, nbits)); p0/p1 are host arrays of nbits
 * probabilities. Replaces Backend.apply_bitflips (abstract.py:2279-2295). */
int qsim_apply_bitflips(const int64_t* samples, int64_t nshots, int nbits, const double* uniforms,
                        const double* p0, const double* p1, int64_t* out, void* stream);

/* counts[v] += number of s with samples[s] == v, for v < nbins (int64 counts),
 * the accumulation of Backend.update_frequencies (abstract.py:2795-2801). */
int qsim_histogram(const int64_t* samples, int64_t nshots, int64_t* counts, int64_t nbins,
                   void* stream);

/* Expectation of a sum of Pauli strings, accumulated in FP64:
 *   out = sum_t coeffs[t] * Re <psi| i^{ny_t} X^{xmask_t} Z^{zmask_t} |psi>
 * where masks are over amplitude-index bits (bit n-1-q for qubit q) and ny_t is
 * the number of Y factors of term t (Y = i X Z). Replaces the per-term einsum
 * loop of Backend.exp_value_observable_symbolic (abstract.py:2946-3054).
 * `out` is one device double. Terms sharing an X mask share one state pass.
 * `workspace` must hold qsim_expval_workspace(nqubits, nterms) bytes. */
size_t qsim_expval_workspace(int nqubits, int nterms);
int qsim_expval_pauli(const void* state, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                      const uint64_t* zmasks, const int* ny, const double* coeffs, double* out,
                      void* workspace, size_t workspace_bytes, void* stream);

/* qsim_expval_pauli without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The per-block partial sums go to a file-scope device
 * array of qsim_expval_noalloc_max_groups() x 1024 doubles (64 groups,
 * 512 KiB), one copy per GPU held until the process ends. Term tables that
 * need more groups (a group = the terms of one X mask, 32 at most) are
 * rejected with QSIM_EINVAL. The stream must belong to the current device.
 * Every call on a GPU uses that GPU's one array: calls that overlap on one
 * GPU (other streams, other threads) race, so callers serialise them.
 * Same kernels and summation order as qsim_expval_pauli: the results are
 * bitwise equal. */
int qsim_expval_pauli_noalloc(const void* state, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                              const uint64_t* zmasks, const int* ny, const double* coeffs, double* out,
                              void* stream);
int qsim_expval_noalloc_max_groups(void);

/* out = sum_t coeffs[t] * i^{ny_t} X^{xmask_t} Z^{zmask_t} |psi>  (out of
 * place, out must not alias psi). The Hamiltonian-vector product that seeds the
 * adjoint sweep; no counterpart in the reference, which differentiates by
 * re-execution (derivative.py:7-226). */
int qsim_apply_pauli_sum(const void* state, int nqubits, int dtype, int nterms,
                         const uint64_t* xmasks, const uint64_t* zmasks, const int* ny,
                         const double* coeffs, void* out, void* stream);

/* out = Re <bra| M_targets |ket> for a dense 2^k x 2^k host matrix (k = 1..3)
 * under `controls`, accumulated in FP64 into one device double. The
 * derivative term 2 Re <lambda| dU/dtheta |psi> of the adjoint sweep.
 * `workspace` must hold qsim_expval_workspace(nqubits, 1) bytes. */
int qsim_inner_matrix(const void* bra, const void* ket, int nqubits, int dtype, const int* targets,
                      int ntargets, const int* controls, int ncontrols, const void* matrix,
                      double* out, void* workspace, size_t workspace_bytes, void* stream);

/* ---- density matrices ----------------------------------------------------
 * A density matrix of n qubits (1..17) is 2^(2n) complex values, row-major:
 * element (r, c) at r * 2^n + c, i.e. the state vector of 2n qubits whose
 * qubit q is row qubit q and qubit n + q is column qubit q. Unitary gates run
 * through the state-vector entry points on that 2n-qubit vector (U on the row
 * qubits, conj(U) on the column qubits), as the reference's einsum pair in
 * apply_gate (abstract.py:2322-2361) and _apply_gate_controlled_by_density_matrix
 * (abstract.py:3199-3235). Every product is rounded before the sum it enters,
 * as numpy evaluates the reference's expressions. */

/* y = a * x (accumulate = 0), y = y + a * x (accumulate = 1) or y = x / a
 * (accumulate = 2) over nelem complex values: the terms of
 * Backend.apply_channel on density matrices, (1 - sum p) rho + sum_k p_k U_k
 * rho U_k^dagger (abstract.py:2297-2311), and the mean of the final states of
 * repeated execution (abstract.py:2603-2609). y and x must not overlap. */
int qsim_dm_axpy(void* y, const void* x, uint64_t nelem, int dtype, double a, int accumulate, void* stream);

/* In place: rho <- (1 - lam) rho + lam Tr_T(rho) (x) I / 2^k on the k = 1..6
 * `targets` (depolarizing_error_density_matrix, abstract.py:2095-2131). */
int qsim_dm_depolarize(void* rho, int nqubits, int dtype, const int* targets, int ntargets, double lam,
                       void* stream);

/* In place: rho <- (1 - p0 - p1) rho + Tr_q(rho) (x) (p0 |0><0| + p1 |1><1|)
 * on `qubit` (reset_error_density_matrix, abstract.py:2212-2232). */
int qsim_dm_reset(void* rho, int nqubits, int dtype, int qubit, double p0, double p1, void* stream);

/* probs[o] = |sum over the unmeasured qubits of rho[i, i]| for the outcomes o
 * of `qubits` (requested order, first = most significant): the density-matrix
 * branch of calculate_probabilities (abstract.py:2742-2750). F64 output for
 * complex128, F32 for complex64; with every qubit measured it equals numpy's
 * abs of the diagonal bit for bit. */
int qsim_dm_probabilities(const void* rho, int nqubits, int dtype, const int* qubits, int nmeasured, void* probs,
                          void* stream);

/* out[0] = sum_t coeffs[t] * Re Tr(P_t rho) with P_t = i^ny[t] X^xmasks[t]
 * Z^zmasks[t] (the same tables as qsim_expval_pauli), FP64 accumulation: the
 * density-matrix branch of exp_value_observable_symbolic (abstract.py:2979-3005).
 * `workspace` must hold qsim_dm_expval_workspace(nterms) bytes. */
size_t qsim_dm_expval_workspace(int nterms);
int qsim_dm_expval_pauli(const void* rho, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                         const uint64_t* zmasks, const int* ny, const double* coeffs, double* out,
                         void* workspace, size_t workspace_bytes, void* stream);

/* ---- host planner ----------------------------------------------------------
 * Pass selection of the blocked execution of the circuit loop
 * (abstract.py:3321-3322; qibo_amd.blocking.partition), host code: among the
 * nitems pending gates (bit masks of their non-diagonal targets and of all
 * their bits, and whether a pass can hold them), the greedy scan from the
 * seed of the low_bits lowest bits and (unless `ordered`) from every seed of
 * those plus tile_bits - low_bits consecutive bits; the first seed taking the
 * most gates wins. Writes taken[i] = 1 for the gates it takes and the
 * resulting tile bits; returns the number taken (>= 0) or a negative error.
 * No device work. */
int qsim_plan_pass(int nitems, const uint64_t* nd_masks, const uint64_t* all_masks, const uint8_t* blockable,
                   int nqubits, int low_bits, int tile_bits, int ordered, uint8_t* taken, uint64_t* tile_mask);

/* Record compiler of one blocked pass, host code: the records that
 * qsim_apply_blocked runs for `nops` planned gates on the 12 ascending
 * `tile_bits` (the gate loop of Backend._execute_circuit, abstract.py:3321-3322,
 * for one pass; qibo_amd.blocking.compile_pass is its Python restatement).
 * Gate i is op_desc[44 i ..]: kind (0 identity, 1 diagonal, 2 SWAP, 3 dense),
 * number of targets (<= 5; dense 1..3), number of controls (<= 35), unused,
 * 5 target slots, 35 control slots (qibo qubit numbers, targets in matrix
 * order); its complex128 entries (diagonal: 2^k, dense: 4^k row-major)
 * start at op_data[2 data_offsets[i]] (-1: none). Records are written in the
 * state dtype's precision to `out` (capacity >= qsim_compile_pass_capacity
 * (nops, dtype) bytes): *out_bytes bytes, *count records, *switches mapping
 * changes. With `ordered` only the exact per-gate forms are used. No device
 * work. */
size_t qsim_compile_pass_capacity(int nops, int dtype);
int qsim_compile_pass(int nqubits, const int* tile_bits, int nops, const int* op_desc, const double* op_data,
                      const int64_t* data_offsets, int dtype, int ordered, uint8_t* out, size_t capacity,
                      size_t* out_bytes, int* count, int* switches);

/* Neighbour-fusion planner of Circuit.fuse (circuit.py:954-1003, with
 * _Queue.to_fused / from_fused :55-95 and FusedGate.can_fuse / fuse
 * gates/special.py:81-153), host code. Gate i of the queue acts on the qubits
 * of qubit_masks[i] (bit q = qibo qubit q); kind[i] = 0 for a gate, 1 for a
 * measurement, 2 for a special gate (it blocks every qubit). Writes per gate
 * status[i] (0: absorbed into another group, 1: kept as itself, 2: a fused
 * group) and, for status 2, the group's gates in application order at
 * members[member_offsets[i] .. member_offsets[i + 1]) (members holds ngates
 * ints, member_offsets ngates + 1). No device work. */
int qsim_fuse_plan(int nqubits, int ngates, const uint64_t* qubit_masks, const uint8_t* kind, int max_qubits,
                   int* status, int* members, int* member_offsets);

/* Product of a fused group's gate matrices on its `rank` sorted `qubits`
 * (Backend.matrix_fused, abstract.py:2680-2717), host code: starting from
 * the identity, each gate is applied in order to every column. Gate k is
 * gate_desc[(2 + rank) k ..]: number of controlled_by controls, number of
 * targets, the controls, the targets (qibo qubit numbers, targets in matrix
 * order); its 2^nt x 2^nt complex128 matrix starts at mats[2 mat_offsets[k]].
 * `out` receives the 2^rank x 2^rank complex128 product, row-major. */
int qsim_matrix_fused(int rank, const int* qubits, int ngates, const int* gate_desc, const double* mats,
                      const int64_t* mat_offsets, double* out);

/* ---- accuracy reports --------------------------------------------------
 * Comparison of two states of the same size (either dtype each), accumulated
 * in FP64: out[0] + i out[1] = <a|b> = sum conj(a_i) b_i, out[2] = <a|a>,
 * out[3] = <b|b>, out[4] = max_i |a_i - b_i| (`out`: 5 device doubles).
 * qibo's pure-state fidelity is |<a|b>|^2 and its infidelity 1 - fidelity
 * (src/qibo/quantum_info/metrics.py:200-291); the package reports them with
 * the max-abs amplitude error for a complex64 result against the complex128
 * execution of the same circuit. A fixed reduction order: repeated calls are
 * bitwise equal. `workspace` must hold qsim_compare_workspace(nqubits) bytes. */
size_t qsim_compare_workspace(int nqubits);
int qsim_compare_states(const void* a, int a_dtype, const void* b, int b_dtype, int nqubits, double* out,
                        void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QSIM_MI355X_H */
