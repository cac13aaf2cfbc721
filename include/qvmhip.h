/*
 * qvmhip.h -- C ABI of the MI355X (gfx950) statevector engine for the
 * pure-state path of quil-lang/qvm.
 *
 * The reference has no C plugin API: its seams are CLOS generic functions on
 * the pure-state class.  Each entry point below replaces one of them; the
 * citation names the reference definition (path:line in the qvm 1.18.0 tree)
 * whose behaviour it reproduces.  A Lisp CFFI binding calls these from the
 * methods of a new state class; the Python package qvm_amd calls them through
 * ctypes.
 *
 * Conventions
 *  - Every function returns int: 0 = QVMHIP_OK, negative = error.  The message
 *    of the last error on the calling thread is qvmhip_last_error().
 *  - Amplitudes cross the ABI as interleaved complex double (re, im),
 *    little-endian, the layout of SBCL's (complex double-float) vectors.  An
 *    FP32 state is widened on read and rounded on write.
 *  - Gate matrices are row-major 2^k x 2^k interleaved complex double.  Matrix
 *    index bit j selects qubits[j]; qubits[] is the nat-tuple order, i.e. the
 *    LAST Quil argument first (src/utilities.lisp:43-51,
 *    src/compile-gate.lisp:51-64).
 *  - Qubit q is bit q of the amplitude index (src/wavefunction.lisp:24-40).
 *  - Calls on one handle are ordered on that handle's HIP stream and are
 *    asynchronous unless they return host data; a handle is not re-entrant,
 *    distinct handles may be used from distinct threads.  Every call runs on
 *    the device its handle was created on and leaves the calling thread's
 *    current device as it found it.
 */
#ifndef QVMHIP_H
#define QVMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QVMHIP_API __attribute__((visibility("default")))

enum {
    QVMHIP_OK = 0,
    QVMHIP_EINVAL = -1,      /* bad argument (qubit out of range, arity, ...) */
    QVMHIP_ENOMEM = -2,      /* device allocation failed or memory cap hit   */
    QVMHIP_EDEVICE = -3,     /* HIP runtime error                             */
    QVMHIP_ERANGE = -4,      /* probability / norm argument out of range      */
    QVMHIP_EUNSUPPORTED = -5 /* valid request this build does not implement   */
};

enum { QVMHIP_F64 = 0, QVMHIP_F32 = 1 };

/* Sampling rules: GT = min{b : C(b) > u} (measure-all, bisection sampler,
 * src/measurement.lisp:179-225); GE = min{i : cdf[i] >= u} (multi-shot,
 * src/measurement.lisp:246-288).  Both clamp to len-1. */
enum { QVMHIP_RULE_GT = 0, QVMHIP_RULE_GE = 1 };

/* Tile-pass op kinds (gates applied inside an LDS-resident block). */
enum {
    QVMHIP_TILE_DENSE1 = 1, /* 2x2 on local qubit q0                          */
    QVMHIP_TILE_DENSE2 = 2, /* 4x4 on local qubits (q0 = matrix bit 0, q1)    */
    QVMHIP_TILE_DIAG1 = 3,  /* diag(m[0], m[1]) on q0                          */
    QVMHIP_TILE_DIAG2 = 4,  /* diag(m[0..3]) on (q0, q1)                       */
    QVMHIP_TILE_CPERM = 5,  /* swap |..c=1,t=0..> <-> |..c=1,t=1..>: q0 = t, q1 = c (CNOT) */
    QVMHIP_TILE_SWAP = 6,   /* swap local qubits q0, q1                        */
    QVMHIP_TILE_CPHASE = 7, /* multiply amplitudes with q0 = q1 = 1 by m[0]   */
    QVMHIP_TILE_GDIAG = 8   /* diagonal that also reads qubits outside the tile (below) */
};

/* QVMHIP_TILE_GDIAG: a diagonal gate on local qubit q0 and one or two state
 * qubits g0, g1 that are NOT tile qubits.  Outside the tile those qubits are
 * constant over a block, so the block's base index selects the entries.
 * q1 packs them: g0 = q1 & 0xff, g1 = (q1 >> 8) & 0xff, ng = q1 >> 16
 * (1 or 2).  m holds 2^(1+ng) complex entries, index = b(q0) + 2 b(g0) +
 * 4 b(g1).  ctrl_mask must be 0.  A 1- or 2-qubit diagonal whose qubits are
 * all outside the tile uses any local q0 with each entry repeated.  It lets
 * a pass apply diagonal gates (RZ, CZ, CPHASE, RZZ) on any qubits. */
#define QVMHIP_GDIAG_PACK(g0, g1, ng) ((uint32_t)(g0) | ((uint32_t)(g1) << 8) | ((uint32_t)(ng) << 16))

/* One gate of a tile pass.  q0/q1 are LOCAL qubit positions in the tile
 * (0 .. m-1), ctrl_mask is a mask of local positions that must be 1 for the
 * op to act (0 = none).  m holds the matrix entries needed by the kind
 * (interleaved complex double, row-major). */
typedef struct {
    uint32_t kind;
    uint32_t q0;
    uint32_t q1;
    uint32_t ctrl_mask;
    double m[32];
} qvmhip_tile_op;

/* qvmhip_config.flags: tile passes on this state contract their complex
 * products with fused multiply-adds (fewer roundings and instructions; the
 * results are then no longer bit-identical to gate-by-gate application).
 * Such passes may also apply commuting ops of a pass in another order and
 * combine consecutive diagonals (a run of ZZ phases diag(a, b, b, a) of one
 * angle becomes one multiply by a^(E - c) b^c, c = the number of the run's
 * qubit pairs whose bits differ); only the rounding changes.
 * The plain gate kernels always keep the reference's rounding. */
#define QVMHIP_FLAG_FMA_TILES 1u

typedef struct {
    uint32_t struct_size;      /* sizeof(qvmhip_config)                       */
    int32_t device;            /* HIP device ordinal; -1 = current device      */
    uint64_t memory_cap_bytes; /* refuse states larger than this; 0 = no cap   */
    uint32_t flags;            /* QVMHIP_FLAG_* bits, 0 for the defaults       */
} qvmhip_config;

typedef struct qvmhip_state qvmhip_state;
typedef struct qvmhip_program qvmhip_program;

/* --- library ----------------------------------------------------------- */
QVMHIP_API const char *qvmhip_last_error(void);
QVMHIP_API const char *qvmhip_version(void);
QVMHIP_API int qvmhip_device_count(int *out);
/* Bytes a state of n qubits and dtype needs, including scratch. */
QVMHIP_API int qvmhip_state_bytes(uint32_t n, int dtype, uint64_t *out);

/* --- state lifetime (make-qvm / make-pure-state / pure-state protocol) ---- */
/* make-pure-state + bring-to-zero-state: src/state-representation.lisp:76-102,
 * src/qvm.lisp:124-164.  The new state is |0...0>. */
QVMHIP_API int qvmhip_create(uint32_t n, int dtype, const qvmhip_config *cfg, qvmhip_state **out);
QVMHIP_API int qvmhip_destroy(qvmhip_state *s);
/* num-qubits: src/state-representation.lisp:25-47. */
QVMHIP_API int qvmhip_num_qubits(const qvmhip_state *s, uint32_t *n);
QVMHIP_API int qvmhip_dtype(const qvmhip_state *s, int *dtype);
/* set-to-zero-state / bring-to-zero-state: src/wavefunction.lisp:79-91. */
QVMHIP_API int qvmhip_reset_zero(qvmhip_state *s);
/* Wait for all work queued on the handle. */
QVMHIP_API int qvmhip_sync(qvmhip_state *s);

/* --- gate application (apply-gate-to-state / apply-matrix-operator) ----- */
/* apply-matrix-operator: src/wavefunction.lisp:308-330 and the kernels it
 * dispatches to (src/serial-kernels.lisp:19-271).  The engine detects control
 * qubits, diagonal and permutation structure and touches only the amplitudes
 * the matrix can change; the FP64 result equals the serial kernels'
 * multiply-then-add arithmetic. 1 <= k <= 6. */
QVMHIP_API int qvmhip_apply_matrix(qvmhip_state *s, const double *U, uint32_t k,
                                   const uint32_t *qubits);
/* Same, with structure detection off: every amplitude is swept through the
 * dense 2^k x 2^k product (the generic apply-operator,
 * src/wavefunction.lisp:271-306). */
QVMHIP_API int qvmhip_apply_matrix_dense(qvmhip_state *s, const double *U, uint32_t k,
                                         const uint32_t *qubits);
/* Diagonal gate diag(d[0..2^k-1]) (a diagonal quantum-operator applied by
 * apply-matrix-operator, src/wavefunction.lisp:308-330). */
QVMHIP_API int qvmhip_apply_diagonal(qvmhip_state *s, const double *d, uint32_t k,
                                     const uint32_t *qubits);
/* Permutation gate: psi'[j] = psi[perm[j]] on gate-local indices, the
 * transposition swaps of src/compile-gate.lisp:259-309. */
QVMHIP_API int qvmhip_apply_permutation(qvmhip_state *s, const uint32_t *perm, uint32_t k,
                                        const uint32_t *qubits);
/* Dense fused k-qubit block (3 <= k <= 6) on the matrix cores (MFMA).  The
 * fused blocks of compile-loaded-program with *fuse-gates-during-compilation*
 * (src/qvm.lisp:166-175) land here. */
QVMHIP_API int qvmhip_apply_fused(qvmhip_state *s, const double *U, uint32_t k,
                                  const uint32_t *qubits);
/* Apply a sequence of small gates to every 2^m-amplitude block spanned by
 * tile_qubits (ascending, 8 <= m <= 12 for F64, 13 for F32) with the block
 * resident on chip (registers, re-laid out through LDS when a target needs
 * it): one HBM sweep for the whole sequence.  The gate-by-gate arithmetic is
 * the same as qvmhip_apply_matrix, so an in-order pass is bit-identical. */
QVMHIP_API int qvmhip_apply_tile_pass(qvmhip_state *s, uint32_t m, const uint32_t *tile_qubits,
                                      uint32_t nops, const qvmhip_tile_op *ops);
/* The slot-space program a tile pass of width m runs for dtype and config
 * flags (the op list
 * with register relayouts inserted; no counterpart in the reference -- it
 * exposes the tile scheduler for inspection and tests).  Writes up to cap ops
 * to out and the full count to *nout; QVMHIP_ERANGE when cap is too small. */
QVMHIP_API int qvmhip_tile_schedule(int dtype, uint32_t flags, uint32_t m, uint32_t nops,
                                    const qvmhip_tile_op *ops, qvmhip_tile_op *out, uint32_t cap,
                                    uint32_t *nout);

/* --- reductions --------------------------------------------------------- */
/* wavefunction-excited-state-probability: src/wavefunction.lisp:62-70. */
QVMHIP_API int qvmhip_prob_excited(qvmhip_state *s, uint32_t q, double *p1);
/* wavefunction-ground-state-probability: src/wavefunction.lisp:52-60. */
QVMHIP_API int qvmhip_prob_ground(qvmhip_state *s, uint32_t q, double *p0);
/* norm: src/wavefunction.lisp:332-346. */
QVMHIP_API int qvmhip_norm(qvmhip_state *s, double *out);
/* (psum #'probability state): the sum of |psi_i|^2 before the square root of
 * norm, as the stochastic Kraus evolution accumulates it
 * (src/apply-gate.lisp:16-39). */
QVMHIP_API int qvmhip_total_probability(qvmhip_state *s, double *out);
/* normalize-wavefunction: src/wavefunction.lisp:349-364 (length <= 0:
 * compute the norm first). */
QVMHIP_API int qvmhip_normalize(qvmhip_state *s, double length);
/* |psi_i|^2 for i in [start, start+count) into host doubles (the
 * probabilities endpoint, app/src/handle-request.lisp:154-175). */
QVMHIP_API int qvmhip_probabilities(qvmhip_state *s, uint64_t start, uint64_t count, double *out);
/* <psi| sum_e Z_a Z_b |psi> over edges (a_e, b_e): the QAOA MaxCut cost
 * expectation (examples/qaoa.lisp:47-75).  edges = 2*nedges qubit indices. */
QVMHIP_API int qvmhip_expectation_zz(qvmhip_state *s, const uint32_t *edges, uint32_t nedges,
                                     double *out);
/* <a|b> (the inner product of pure-state-expectation,
 * app/src/api/expectation.lisp:78-89). Both states: same n and dtype. */
QVMHIP_API int qvmhip_inner_product(qvmhip_state *a, qvmhip_state *b, double *re, double *im);

/* --- measurement -------------------------------------------------------- */
/* force-measurement, pure state: src/measurement.lisp:10-41.  Forcing an
 * outcome whose probability is 0 (bit 1 with p1 = 0, bit 0 with p1 = 1) has
 * no normalisation and returns QVMHIP_ERANGE with the state unchanged. */
QVMHIP_API int qvmhip_force_measure(qvmhip_state *s, uint32_t q, int bit, double p1);
/* measure (src/measurement.lisp:87-105) with the caller's uniform u.
 * compiled = 0: interpreted rule -- p1 == 0 gives 0 and consumes no draw,
 *               else bit = (u <= p1).
 * compiled = 1: compiled projector (src/compile-gate.lisp:221-254) -- always
 *               consumes u, bit = (u < p0) ? 0 : 1 (0 also when p0 >= 1:
 *               the reference's draw is below 1, this one may equal 1).
 * *consumed = number of uniforms used (0 or 1). */
QVMHIP_API int qvmhip_measure(qvmhip_state *s, uint32_t q, double u, int compiled, int *bit,
                              int *consumed);
/* measure-all-state, pure (src/measurement.lisp:128-143): index =
 * min{b : C(b) > u} (len-1 if none), state collapses to |index> with
 * amplitude exactly 1. */
QVMHIP_API int qvmhip_measure_all(qvmhip_state *s, double u, uint64_t *index);
/* nshots measure-all draws of the same prepared state without collapse: the
 * reload-per-shot loop of app/src/api/multishot-measure.lisp:48-110, rule
 * C(b) > u. */
QVMHIP_API int qvmhip_sample_measure_all(qvmhip_state *s, const double *u, uint64_t nshots,
                                         uint64_t *out);
/* sample-wavefunction-multiple-times (src/measurement.lisp:246-288): per u,
 * min{i : cdf[i] >= u}, clamped to len-1.  No collapse. */
QVMHIP_API int qvmhip_sample(qvmhip_state *s, const double *u, uint64_t nshots, uint64_t *out);
/* Either sampler above (rule QVMHIP_RULE_GT / _GE) with boundary reporting,
 * the sample call of the pure-state path as a whole (measure-all-state
 * src/measurement.lisp:125-143 + 179-225, sample-wavefunction-multiple-times
 * :246-288).  A draw u that selects b is near a boundary when |u - C(b)| < eps
 * or |u - C(b-1)| < eps, with C the engine's own FP64 prefix values that
 * decided the draw (C(b-1) of a 256-amplitude chunk's first index is the
 * chunk's exclusive prefix; a clamped draw compares C(len-1) only).
 * near_flags (optional, nshots bytes) receives 1 per near draw, n_near
 * (optional) their count; eps = 0 flags nothing.  No collapse. */
QVMHIP_API int qvmhip_sample_ex(qvmhip_state *s, const double *u, uint64_t nshots, int rule, double eps,
                                uint64_t *out, uint8_t *near_flags, uint64_t *n_near);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc):
 * qvmhip_sample_ex on a statevector the caller owns, with every pointer on
 * the current device -- amps (2^n interleaved complex, FP64 or FP32 per
 * dtype), the nshots draws u, the nshots indices out and the near-boundary
 * count n_near (optional; cleared, then counted) -- enqueued on `stream` (a
 * hipStream_t; null = the default stream).  The chunk sums and prefix live in
 * a static device array of 4 MiB + 2 KiB that each GPU holds until the
 * process ends, so n is at most 26 (larger n: QVMHIP_EINVAL).  It allocates
 * nothing and does not synchronise.  Calls that overlap on one GPU share that
 * array and race.  Draws are not range-checked (below 0 selects 0, above 1
 * len-1). */
QVMHIP_API int qvmhip_sample_noalloc(const void *amps, uint32_t n, int dtype, const double *u, uint64_t nshots,
                                     int rule, double eps, uint64_t *out, uint64_t *n_near, void *stream);

/* --- amplitudes (amplitudes / state-elements / copy-wavefunction) -------- */
/* amplitudes, nth-amplitude, state-elements: src/qvm.lisp:63-82,
 * src/state-representation.lisp:138-167.  The only D2H point of the path. */
QVMHIP_API int qvmhip_read_amplitudes(qvmhip_state *s, uint64_t start, uint64_t count, double *out);
/* (setf state-elements) / (setf amplitudes). */
QVMHIP_API int qvmhip_write_amplitudes(qvmhip_state *s, uint64_t start, uint64_t count,
                                       const double *in);
/* copy-wavefunction (src/wavefunction.lisp:93-109): dst <- src, same n/dtype.
 * Queued on dst's stream after src's queued work; src's later calls wait for
 * the copy (the same holds for qvmhip_accumulate). */
QVMHIP_API int qvmhip_copy(qvmhip_state *dst, const qvmhip_state *src);
/* dst += src elementwise, same n/dtype: the running sum of the Kraus-list
 * evolution of a density matrix (map-into sum #'+, src/apply-gate.lisp:76-96). */
QVMHIP_API int qvmhip_accumulate(qvmhip_state *dst, const qvmhip_state *src);
/* For a 2*nq-qubit state holding vec(rho) row-major (rho[i][j] at i*2^nq + j):
 * out[i] = Re rho[i][i], i < 2^nq.  density-matrix-state-measurement-
 * probabilities (src/state-representation.lisp:268-286) and the diagonal of
 * get-excited-state-probability on density states (src/measurement.lisp:77-85). */
QVMHIP_API int qvmhip_density_diagonal(qvmhip_state *s, uint32_t nq, double *out);
/* Wire formats of the server's octet-stream replies: amplitudes
 * [start, start+count) as big-endian IEEE doubles re, im (16 octets each,
 * write-complex-double-float-as-binary, app/src/utilities.lisp:60-75, sent by
 * the wavefunction request, app/src/handle-request.lisp:132-152), and
 * probabilities re*re + im*im as big-endian doubles (8 octets each,
 * write-double-float-as-binary, app/src/utilities.lisp:77-85, sent by the
 * probabilities request, app/src/handle-request.lisp:154-175).  FP32 states
 * are widened to double first.  out holds 16*count or 8*count octets. */
QVMHIP_API int qvmhip_read_wavefunction_be(qvmhip_state *s, uint64_t start, uint64_t count, uint8_t *out);
QVMHIP_API int qvmhip_read_probabilities_be(qvmhip_state *s, uint64_t start, uint64_t count, uint8_t *out);

/* --- compiled programs (compile-loaded-program / compile-instruction) ---- */
/* A program records gates, tile passes and fused blocks against one state,
 * uploads their operands once, and replays them -- optionally as one captured
 * hipGraph (src/qvm.lisp:166-175, src/compile-gate.lisp:461-526).
 * Destroying the state first detaches its programs: they then refuse to
 * run (QVMHIP_EINVAL) and may still be destroyed. */
QVMHIP_API int qvmhip_program_create(qvmhip_state *s, qvmhip_program **out);
QVMHIP_API int qvmhip_program_add_matrix(qvmhip_program *p, const double *U, uint32_t k,
                                         const uint32_t *qubits);
QVMHIP_API int qvmhip_program_add_fused(qvmhip_program *p, const double *U, uint32_t k,
                                        const uint32_t *qubits);
QVMHIP_API int qvmhip_program_add_tile_pass(qvmhip_program *p, uint32_t m,
                                            const uint32_t *tile_qubits, uint32_t nops,
                                            const qvmhip_tile_op *ops);
/* use_graph = 1: capture the launches into one hipGraph. */
QVMHIP_API int qvmhip_program_finalize(qvmhip_program *p, int use_graph);
QVMHIP_API int qvmhip_program_run(qvmhip_program *p);
QVMHIP_API int qvmhip_program_destroy(qvmhip_program *p);

/* max_i |a_i - b_i| over both components (FP64 accumulate); states may
 * differ in dtype.  The parity check of a plan against the plain path. */
QVMHIP_API int qvmhip_max_abs_diff(qvmhip_state *a, qvmhip_state *b, double *out);
/* Total-variation distance of the two outcome distributions,
 * 1/2 sum_i | |a_i|^2 - |b_i|^2 | (FP64 accumulate); dtypes may differ. */
QVMHIP_API int qvmhip_tv_distance(qvmhip_state *a, qvmhip_state *b, double *out);

/* --- timing (engine-recorded kernel timings for the fusion cost model) --- */
/* Kernel families reported by qvmhip_profile_summary. */
enum {
    QVMHIP_K_DENSE1 = 1,    /* dense 1-qubit (reduced) operator          */
    QVMHIP_K_DENSE2 = 2,
    QVMHIP_K_DENSE3 = 3,
    QVMHIP_K_DENSE_BIG = 4, /* dense 4..6 qubits, LDS-staged matrix       */
    QVMHIP_K_DIAG_FULL = 5,
    QVMHIP_K_DIAG_SPARSE = 6,
    QVMHIP_K_PERM = 7,
    QVMHIP_K_TILE = 8,
    QVMHIP_K_FUSED = 9,
    QVMHIP_K_NKINDS = 10
};
/* on != 0: start recording an event pair around every kernel launch on the
 * handle (previous records are dropped); on == 0: stop recording. */
QVMHIP_API int qvmhip_profile_enable(qvmhip_state *s, int on);
/* Waits for the handle, then reports the number of recorded launches of
 * `kind` and their summed duration in milliseconds. */
QVMHIP_API int qvmhip_profile_summary(qvmhip_state *s, int kind, uint64_t *count, double *total_ms);
/* Time `reps` launches of a finalized program with HIP events on its
 * state's stream; *ms = mean per launch. */
QVMHIP_API int qvmhip_time_program(qvmhip_program *p, int reps, double *ms);

/* --- batches of small states (no single counterpart in the reference) ---
 * Many independent states (the parameter sweeps of examples/qaoa.lisp, run
 * there one QVM at a time) kept in one HBM buffer.  Each instance gets a
 * program of tile passes and small gates; finalize schedules the programs
 * level by level (all tile passes of a level and width in ONE batched kernel)
 * followed by each instance's ZZ expectation and samples (cdf >= u rule), and
 * captures the whole run as one hipGraph.  Every run starts from |0...0>. */
typedef struct qvmhip_batch qvmhip_batch;
QVMHIP_API int qvmhip_batch_create(uint32_t count, const uint32_t *nqubits, int dtype, qvmhip_batch **out);
QVMHIP_API int qvmhip_batch_destroy(qvmhip_batch *b);
/* QVMHIP_FLAG_FMA_TILES for the batch's tile passes (before any item is
 * added); no counterpart in the reference. */
QVMHIP_API int qvmhip_batch_set_flags(qvmhip_batch *b, uint32_t flags);
QVMHIP_API int qvmhip_batch_add_tile_pass(qvmhip_batch *b, uint32_t inst, uint32_t m, const uint32_t *tile_qubits,
                                          uint32_t nops, const qvmhip_tile_op *ops);
/* 1..3-qubit gate, same convention as qvmhip_apply_matrix. */
QVMHIP_API int qvmhip_batch_add_matrix(qvmhip_batch *b, uint32_t inst, const double *U, uint32_t k,
                                       const uint32_t *qubits);
/* Cost operator sum_e Z_a Z_b of instance `inst` (edges = 2*nedges qubits). */
QVMHIP_API int qvmhip_batch_set_zz(qvmhip_batch *b, uint32_t inst, const uint32_t *edges, uint32_t nedges);
/* `shots` draws per instance; u holds count*shots uniforms, instance-major. */
QVMHIP_API int qvmhip_batch_set_shots(qvmhip_batch *b, uint64_t shots, const double *u);
QVMHIP_API int qvmhip_batch_finalize(qvmhip_batch *b, int use_graph);
QVMHIP_API int qvmhip_batch_run(qvmhip_batch *b);
QVMHIP_API int qvmhip_batch_sync(qvmhip_batch *b);
/* expectations[count], shots[count*shots] (either may be NULL). */
QVMHIP_API int qvmhip_batch_results(qvmhip_batch *b, double *expectations, uint64_t *shots);
QVMHIP_API int qvmhip_batch_read_amplitudes(qvmhip_batch *b, uint32_t inst, uint64_t start, uint64_t count,
                                            double *out);

/* --- fusion cost model (no counterpart in the reference) ----------------
 * A small MLP (din -> hidden -> hidden -> 1, ReLU, FP32) that predicts the
 * log launch time of a planned launch from its features; trained and run by
 * the engine's own kernels on timings the engine records.  The planner uses it
 * to choose fusion widths and tile passes.  Parameters are one flat vector:
 * W1[hidden][din], b1, W2[hidden][hidden], b2, w3[hidden], b3. */
typedef struct qvmhip_mlp qvmhip_mlp;
QVMHIP_API int qvmhip_mlp_create(int din, int hidden, const float *init_params, qvmhip_mlp **out);
QVMHIP_API int qvmhip_mlp_destroy(qvmhip_mlp *m);
QVMHIP_API int qvmhip_mlp_num_params(const qvmhip_mlp *m, int *np);
QVMHIP_API int qvmhip_mlp_get_params(qvmhip_mlp *m, float *out);
QVMHIP_API int qvmhip_mlp_set_params(qvmhip_mlp *m, const float *in);
/* out[b] = f(X[b, :]) for b < nb (X row-major, nb x din). */
QVMHIP_API int qvmhip_mlp_predict(qvmhip_mlp *m, const float *X, int nb, float *out);
/* Gradient of mean_b 0.5 (f(X_b) - T_b)^2 with respect to the parameters. */
QVMHIP_API int qvmhip_mlp_gradient(qvmhip_mlp *m, const float *X, const float *T, int nb, float *grad,
                                   float *loss);
/* `epochs` full-batch Adam steps on the same loss. */
QVMHIP_API int qvmhip_mlp_train(qvmhip_mlp *m, const float *X, const float *T, int nb, int epochs, float lr,
                                float *final_loss);

#ifdef __cplusplus
}
#endif
#endif /* QVMHIP_H */
