/*
 * qvm_oracle.c -- CPU restatement of quil-lang/qvm 1.18.0's pure-state path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * qvm_amd/ and the timed CPU baseline; it is never linked into, loaded by or
 * called from the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * The reference is Common Lisp (SBCL + Quicklisp); there is no Lisp here, so
 * the reference itself cannot be built or run.  Every function below restates
 * one reference function and cites it as path:line inside the reference tree.
 *
 * Layout: amplitudes are interleaved complex double (re, im), 16 bytes each,
 * index bit q <-> qubit q (src/wavefunction.lisp:24-40).  Gate matrices are
 * row-major 2^k x 2^k interleaved complex double, matrix index bit j <->
 * qubits[j], where qubits[] is the nat-tuple order (last Quil argument first,
 * src/utilities.lisp:43-51).
 *
 * Rounding.  The reference has two roundings for the same gate: the serial
 * kernels and the generic path use a separate multiply and add per complex
 * product (SBCL's complex arithmetic), the AVX kernels use FMA
 * (src/impl/sbcl-avx-vops.lisp:210-245).  This file is built with
 * -ffp-contract=off so the "serial", "generic" and "compiled" modes reproduce
 * the multiply-then-add rounding; the "avx" mode calls fma() in the order of
 * the VOP.
 *
 * Parity pinning (what checks each group of functions; tests/test_oracle.py):
 *   index math        golden values of tests/utilities-tests.lisp:83-97 and
 *                     tests/measurement-tests.lisp:10-20; brute force.
 *   gate application  closed forms (H^n, GHZ, 2q QFT table of
 *                     tests/gate-tests.lisp:163-183), brute-force Kronecker
 *                     product unitary for n <= 10, and cross-mode agreement
 *                     (serial vs generic vs compiled vs avx) with asymmetric
 *                     gates on non-adjacent, reversed-order qubits.
 *   reductions        long-double sums.
 *   sampling          golden values of tests/measurement-tests.lisp:184-211,
 *                     CDF golden tests/linear-algebra-tests.lisp:66-75, brute
 *                     force linear search.
 *   measurement       closed forms on basis and Bell states; the RNG
 *                     consumption rules are read from the source text
 *                     (src/measurement.lisp:99-101, src/compile-gate.lisp:236).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Configuration specials (src/config.lisp:9-26).                           */

/* *qubits-required-for-parallelization* = 19 (src/config.lisp:14). */
static int g_par_log2 = 19;
/* Worker count: lparallel kernel size; default = logical cores
 * (src/utilities.lisp:202-239). */
static int g_workers = 0;

EXPORT void orc_set_parallelization(int log2_threshold, int workers) {
    g_par_log2 = log2_threshold;
    g_workers = workers;
}

EXPORT int orc_count_logical_cores(void) {
    /* src/utilities.lisp:205-211 */
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static int workers(void) {
    return g_workers > 0 ? g_workers : orc_count_logical_cores();
}

/* ------------------------------------------------------------------------ */
/* Index math.                                                              */

/* inject-bit: insert a 0 at bit position n (src/utilities.lisp:243-255). */
EXPORT uint64_t orc_inject_bit(uint64_t x, unsigned n) {
    uint64_t right = x & ((UINT64_C(1) << n) - 1);
    uint64_t left = (x >> n) << (n + 1);
    return left | right;
}

/* eject-bit: remove bit n (src/utilities.lisp:267-279). */
EXPORT uint64_t orc_eject_bit(uint64_t x, unsigned n) {
    uint64_t right = x & ((UINT64_C(1) << n) - 1);
    uint64_t left = (x >> 1) & ~((UINT64_C(1) << n) - 1);
    return left | right;
}

/* index-to-address (src/wavefunction.lisp:24-40). */
EXPORT uint64_t orc_index_to_address(uint64_t index, unsigned qubit, unsigned state) {
    uint64_t a = orc_inject_bit(index, qubit);
    return state ? (a | (UINT64_C(1) << qubit)) : a;
}

/* nat-tuple: store the arguments reversed (src/utilities.lisp:43-51). */
EXPORT void orc_nat_tuple(const uint32_t *args, uint32_t k, uint32_t *out) {
    for (uint32_t i = 0; i < k; ++i) out[k - 1 - i] = args[i];
}

/* map-subdivisions / subdivide (src/utilities.lisp:157-182): ranges of size
 * floor(total/num); the last range takes the slack; one range if size==0.
 * Returns the number of ranges written. */
EXPORT int orc_subdivide(uint64_t total, int num_ranges, uint64_t *starts, uint64_t *ends) {
    uint64_t size = total / (uint64_t)num_ranges;
    if (size == 0) {
        starts[0] = 0;
        ends[0] = total;
        return 1;
    }
    uint64_t start = 0;
    for (int w = 1; w <= num_ranges; ++w, start += size) {
        starts[w - 1] = start;
        ends[w - 1] = (w == num_ranges) ? total : start + size;
    }
    return num_ranges;
}

/* ------------------------------------------------------------------------ */
/* Thread runtime: with-parallel-subdivisions / pdotimes / psum-dotimes      */
/* (src/utilities.lisp:346-425).  Workers get the contiguous ranges of        */
/* subdivide; partial sums are added in range order (one admissible           */
/* completion order of src/utilities.lisp:422-425).                           */

typedef void (*range_fn)(void *ctx, uint64_t start, uint64_t end, int idx);

typedef struct {
    range_fn f;
    void *ctx;
    uint64_t start, end;
    int idx;
} task_t;

static void *task_main(void *p) {
    task_t *t = (task_t *)p;
    t->f(t->ctx, t->start, t->end, t->idx);
    return NULL;
}

#define MAX_WORKERS 512

/* Run f over [0,count) split by subdivide into `workers()` ranges when
 * count >= 2^threshold (or force), else serially.  Returns #ranges. */
static int par_run(uint64_t count, int force, range_fn f, void *ctx) {
    if (count == 0) return 0;
    int nw = workers();
    if ((!force && count < (UINT64_C(1) << g_par_log2)) || nw <= 1) {
        f(ctx, 0, count, 0);
        return 1;
    }
    if (nw > MAX_WORKERS) nw = MAX_WORKERS;
    uint64_t starts[MAX_WORKERS], ends[MAX_WORKERS];
    int nr = orc_subdivide(count, nw, starts, ends);
    pthread_t th[MAX_WORKERS];
    task_t tasks[MAX_WORKERS];
    for (int i = 0; i < nr; ++i) {
        tasks[i] = (task_t){f, ctx, starts[i], ends[i], i};
        if (i > 0) pthread_create(&th[i], NULL, task_main, &tasks[i]);
    }
    task_main(&tasks[0]);
    for (int i = 1; i < nr; ++i) pthread_join(th[i], NULL);
    return nr;
}

/* ------------------------------------------------------------------------ */
/* Complex helpers: SBCL (complex double-float) arithmetic, no contraction.  */

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline double probability(cplx a) {
    /* src/wavefunction.lisp:44-50 */
    return a.re * a.re + a.im * a.im;
}

/* ------------------------------------------------------------------------ */
/* State initialisation.                                                    */

/* bring-to-zero-state (src/wavefunction.lisp:79-91). */
EXPORT void orc_bring_to_zero_state(double *psi, uint64_t len) {
    memset(psi, 0, len * 2 * sizeof(double));
    psi[0] = 1.0;
}

/* ------------------------------------------------------------------------ */
/* Gate application.                                                        */

/* Address of the amplitude whose gate-local index is `combo` on top of the
 * complement address `base`: bit j of combo sets qubits[j]
 * (set-qubit-components-of-amplitude-address, src/wavefunction.lisp:112-122). */
static inline uint64_t combo_address(uint64_t base, uint32_t combo, uint32_t k,
                                     const uint32_t *qubits) {
    uint64_t a = base;
    for (uint32_t j = 0; j < k; ++j)
        if ((combo >> j) & 1u) a |= UINT64_C(1) << qubits[j];
    return a;
}

/* Complement address of complement index i: inject zeros at the sorted
 * target positions, ascending (map-complement, src/wavefunction.lisp:155-175;
 * generate-complement-iteration, src/compile-gate.lisp:23-49). */
static inline uint64_t complement_address(uint64_t i, uint32_t k, const uint32_t *sorted) {
    uint64_t a = i;
    for (uint32_t j = 0; j < k; ++j) a = orc_inject_bit(a, sorted[j]);
    return a;
}

static void sort_qubits(const uint32_t *q, uint32_t k, uint32_t *out) {
    for (uint32_t i = 0; i < k; ++i) out[i] = q[i];
    for (uint32_t i = 1; i < k; ++i)
        for (uint32_t j = i; j > 0 && out[j - 1] > out[j]; --j) {
            uint32_t t = out[j];
            out[j] = out[j - 1];
            out[j - 1] = t;
        }
}

typedef struct {
    cplx *psi;
    const cplx *U;
    uint32_t k, dim;
    uint32_t qubits[16];
    uint32_t sorted[16];
    int mode; /* 0 serial/generic, 1 compiled (skip 0, elide x1) */
} gate_ctx;

/* One complement group, matrix-vector product with left-to-right row sums
 * (define-serial-kernel src/serial-kernels.lisp:54-67; matrix-multiply
 * src/linear-algebra.lisp:53-131; inlined compiled matvec that skips zero
 * entries and elides multiplication by one, src/compile-gate.lisp:113-154). */
static void gate_range(void *vctx, uint64_t start, uint64_t end, int idx) {
    (void)idx;
    gate_ctx *c = (gate_ctx *)vctx;
    cplx col[64], res[64];
    uint64_t addr[64];
    for (uint64_t i = start; i < end; ++i) {
        uint64_t base = complement_address(i, c->k, c->sorted);
        for (uint32_t j = 0; j < c->dim; ++j) {
            addr[j] = combo_address(base, j, c->k, c->qubits);
            col[j] = c->psi[addr[j]];
        }
        for (uint32_t r = 0; r < c->dim; ++r) {
            const cplx *row = c->U + (size_t)r * c->dim;
            if (c->mode == 1) {
                int have = 0;
                cplx acc = {0, 0};
                for (uint32_t j = 0; j < c->dim; ++j) {
                    cplx m = row[j];
                    if (m.re == 0.0 && m.im == 0.0) continue;
                    cplx term = (m.re == 1.0 && m.im == 0.0) ? col[j] : cmul(col[j], m);
                    acc = have ? cadd(acc, term) : term;
                    have = 1;
                }
                res[r] = have ? acc : (cplx){0, 0};
            } else {
                cplx acc = cmul(col[0], row[0]);
                for (uint32_t j = 1; j < c->dim; ++j) acc = cadd(acc, cmul(col[j], row[j]));
                res[r] = acc;
            }
        }
        for (uint32_t j = 0; j < c->dim; ++j) c->psi[addr[j]] = res[j];
    }
}

static int gate_setup(gate_ctx *c, double *psi, uint32_t nq, const double *U, uint32_t k,
                      const uint32_t *qubits, int mode) {
    if (k < 1 || k > 6 || k > nq) return -1;
    for (uint32_t j = 0; j < k; ++j) {
        if (qubits[j] >= nq) return -1;
        for (uint32_t i = 0; i < j; ++i)
            if (qubits[i] == qubits[j]) return -1;
    }
    c->psi = (cplx *)psi;
    c->U = (const cplx *)U;
    c->k = k;
    c->dim = 1u << k;
    for (uint32_t j = 0; j < k; ++j) c->qubits[j] = qubits[j];
    sort_qubits(qubits, k, c->sorted);
    c->mode = mode;
    return 0;
}

/* Serial kernels apply-1q/2q/3q-operator (src/serial-kernels.lisp:19-168,
 * 271) and the generic apply-operator (src/wavefunction.lisp:271-330): the
 * same per-group arithmetic; serial below 2^19 complement groups, threaded
 * above (map-complement-in-parallel). */
EXPORT int orc_apply_matrix(double *psi, uint32_t nq, const double *U, uint32_t k,
                            const uint32_t *qubits) {
    gate_ctx c;
    if (gate_setup(&c, psi, nq, U, k, qubits, 0)) return -1;
    par_run(UINT64_C(1) << (nq - k), 0, gate_range, &c);
    return 0;
}

/* Compiled mode: generate-gate-application-code-with-matrix with the matrix
 * inlined (src/compile-gate.lisp:23-209), iterated by the QVM's pdotimes
 * (parallel above 2^19 iterations, src/utilities.lisp:374-383). */
EXPORT int orc_apply_compiled(double *psi, uint32_t nq, const double *U, uint32_t k,
                              const uint32_t *qubits) {
    gate_ctx c;
    if (gate_setup(&c, psi, nq, U, k, qubits, 1)) return -1;
    uint64_t n = UINT64_C(1) << (nq - k);
    int force = n > (UINT64_C(1) << g_par_log2);
    if (force)
        par_run(n, 1, gate_range, &c);
    else
        gate_range(&c, 0, n, 0);
    return 0;
}

/* AVX2 kernels (src/serial-kernels.lisp:176-267): k=1 matmul2-simd and
 * k=2 matmul4-simd-half, FMA rounding in the order of
 * src/impl/sbcl-avx-vops.lisp:210-245 and :294-355, split by
 * with-parallel-subdivisions. */
typedef struct {
    cplx *psi;
    const cplx *U;
    uint32_t q0, q1, p0, p1;
} avx_ctx;

/* row r of U applied to (a0..a{m-1}) with the VOP's accumulation:
 * acc = sum_j Ui*swz(a_j) (first a multiply, then FMAs), then
 * fmaddsub with Ur*a0, then FMAs with Ur*a_j. */
static inline cplx avx_row(const cplx *row, const cplx *a, int m) {
    /* imaginary-part products on swizzled (a.im, a.re) */
    double t_re = row[0].im * a[0].im; /* goes to re with a minus */
    double t_im = row[0].im * a[0].re;
    for (int j = 1; j < m; ++j) {
        t_re = fma(row[j].im, a[j].im, t_re);
        t_im = fma(row[j].im, a[j].re, t_im);
    }
    cplx r;
    r.re = fma(row[0].re, a[0].re, -t_re);
    r.im = fma(row[0].re, a[0].im, t_im);
    for (int j = 1; j < m; ++j) {
        r.re = fma(row[j].re, a[j].re, r.re);
        r.im = fma(row[j].re, a[j].im, r.im);
    }
    return r;
}

static void avx1_range(void *vctx, uint64_t start, uint64_t end, int idx) {
    (void)idx;
    avx_ctx *c = (avx_ctx *)vctx;
    uint64_t bit = UINT64_C(1) << c->q0;
    for (uint64_t i = start; i < end; ++i) {
        uint64_t ai = orc_inject_bit(i, c->q0), bi = ai | bit;
        cplx a[2] = {c->psi[ai], c->psi[bi]};
        cplx p = avx_row(c->U, a, 2), q = avx_row(c->U + 2, a, 2);
        c->psi[ai] = p;
        c->psi[bi] = q;
    }
}

static void avx2_range(void *vctx, uint64_t start, uint64_t end, int idx) {
    (void)idx;
    avx_ctx *c = (avx_ctx *)vctx;
    uint64_t m0 = UINT64_C(1) << c->q0, m1 = UINT64_C(1) << c->q1;
    for (uint64_t i = start; i < end; ++i) {
        uint64_t ai = orc_inject_bit(orc_inject_bit(i, c->p0), c->p1);
        uint64_t bi = ai | m0, ci = ai | m1, di = bi | ci;
        cplx a[4] = {c->psi[ai], c->psi[bi], c->psi[ci], c->psi[di]};
        cplx r0 = avx_row(c->U, a, 4), r1 = avx_row(c->U + 4, a, 4);
        cplx r2 = avx_row(c->U + 8, a, 4), r3 = avx_row(c->U + 12, a, 4);
        c->psi[ai] = r0;
        c->psi[bi] = r1;
        c->psi[ci] = r2;
        c->psi[di] = r3;
    }
}

EXPORT int orc_apply_avx(double *psi, uint32_t nq, const double *U, uint32_t k,
                         const uint32_t *qubits) {
    avx_ctx c = {(cplx *)psi, (const cplx *)U, 0, 0, 0, 0};
    if (k == 1) {
        if (qubits[0] >= nq) return -1;
        c.q0 = qubits[0];
        par_run(UINT64_C(1) << (nq - 1), 0, avx1_range, &c);
        return 0;
    }
    if (k == 2) {
        if (qubits[0] >= nq || qubits[1] >= nq || qubits[0] == qubits[1] || nq < 2) return -1;
        c.q0 = qubits[0];
        c.q1 = qubits[1];
        c.p0 = c.q0 < c.q1 ? c.q0 : c.q1;
        c.p1 = c.q0 < c.q1 ? c.q1 : c.q0;
        par_run(UINT64_C(1) << (nq - 2), 0, avx2_range, &c);
        return 0;
    }
    return -1;
}

/* Permutation gates: transpositions of permutation-to-transpositions
 * (src/compile-gate.lisp:259-269) applied as rotatef swaps per complement
 * group (src/compile-gate.lisp:271-309).  Net effect psi'[j] = psi[perm[j]]
 * on gate-local indices (tests/utilities-tests.lisp:41-57). */
EXPORT int orc_permutation_to_transpositions(const uint32_t *perm, uint32_t len,
                                             uint32_t *left, uint32_t *right) {
    int n = 0;
    for (uint32_t dest = 0; dest < len; ++dest) {
        uint32_t src = perm[dest];
        while (src < dest) src = perm[src];
        if (src != dest) {
            left[n] = src;
            right[n] = dest;
            ++n;
        }
    }
    return n;
}

typedef struct {
    cplx *psi;
    uint32_t k, dim, nswaps;
    uint32_t qubits[16], sorted[16];
    uint32_t left[64], right[64];
} perm_ctx;

static void perm_range(void *vctx, uint64_t start, uint64_t end, int idx) {
    (void)idx;
    perm_ctx *c = (perm_ctx *)vctx;
    for (uint64_t i = start; i < end; ++i) {
        uint64_t base = complement_address(i, c->k, c->sorted);
        for (uint32_t s = 0; s < c->nswaps; ++s) {
            uint64_t a = combo_address(base, c->left[s], c->k, c->qubits);
            uint64_t b = combo_address(base, c->right[s], c->k, c->qubits);
            cplx t = c->psi[a];
            c->psi[a] = c->psi[b];
            c->psi[b] = t;
        }
    }
}

EXPORT int orc_apply_permutation(double *psi, uint32_t nq, const uint32_t *perm, uint32_t k,
                                 const uint32_t *qubits) {
    if (k < 1 || k > 6 || k > nq) return -1;
    perm_ctx c;
    c.psi = (cplx *)psi;
    c.k = k;
    c.dim = 1u << k;
    for (uint32_t j = 0; j < k; ++j) {
        if (qubits[j] >= nq) return -1;
        c.qubits[j] = qubits[j];
    }
    sort_qubits(qubits, k, c.sorted);
    c.nswaps = (uint32_t)orc_permutation_to_transpositions(perm, c.dim, c.left, c.right);
    uint64_t n = UINT64_C(1) << (nq - k);
    if (n > (UINT64_C(1) << g_par_log2))
        par_run(n, 1, perm_range, &c);
    else
        perm_range(&c, 0, n, 0);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Reductions (psum-dotimes, src/utilities.lisp:395-425).                   */

typedef struct {
    const cplx *psi;
    uint64_t offset;
    uint32_t qubit;
    int kind; /* 0 plain range, 1 excited half, 2 ground half */
    double partial[MAX_WORKERS];
} sum_ctx;

static void sum_range(void *vctx, uint64_t start, uint64_t end, int idx) {
    sum_ctx *c = (sum_ctx *)vctx;
    double s = 0.0;
    for (uint64_t i = start; i < end; ++i) {
        uint64_t a;
        if (c->kind == 0)
            a = c->offset + i;
        else
            a = orc_index_to_address(i, c->qubit, c->kind == 1 ? 1 : 0);
        s += probability(c->psi[a]);
    }
    c->partial[idx] = s;
}

static double psum_dotimes(sum_ctx *c, uint64_t range) {
    if (range < (UINT64_C(1) << g_par_log2)) {
        sum_range(c, 0, range, 0);
        return c->partial[0];
    }
    int nr = par_run(range, 1, sum_range, c);
    double s = 0.0;
    for (int i = 0; i < nr; ++i) s += c->partial[i];
    return s;
}

/* wavefunction-excited-state-probability (src/wavefunction.lisp:62-70). */
EXPORT double orc_excited_probability(const double *psi, uint32_t nq, uint32_t q) {
    sum_ctx c = {(const cplx *)psi, 0, q, 1, {0}};
    return psum_dotimes(&c, (UINT64_C(1) << nq) / 2);
}

/* wavefunction-ground-state-probability (src/wavefunction.lisp:52-60). */
EXPORT double orc_ground_probability(const double *psi, uint32_t nq, uint32_t q) {
    sum_ctx c = {(const cplx *)psi, 0, q, 2, {0}};
    return psum_dotimes(&c, (UINT64_C(1) << nq) / 2);
}

/* norm: %serial-norm below 19 qubits, psum above
 * (src/wavefunction.lisp:332-346). */
EXPORT double orc_norm(const double *psi, uint64_t len) {
    sum_ctx c = {(const cplx *)psi, 0, 0, 0, {0}};
    int nq = 0;
    while ((UINT64_C(1) << (nq + 1)) <= len) ++nq;
    if (nq < 1) nq = 1; /* wavefunction-qubits clamps at 1 */
    if (nq < g_par_log2) {
        sum_range(&c, 0, len, 0);
        return sqrt(c.partial[0]);
    }
    return sqrt(psum_dotimes(&c, len));
}

/* normalize-wavefunction (src/wavefunction.lisp:349-364). length<=0 means
 * compute the norm. */
EXPORT void orc_normalize(double *psi, uint64_t len, double length) {
    double inv = 1.0 / (length > 0 ? length : orc_norm(psi, len));
    for (uint64_t i = 0; i < 2 * len; ++i) psi[i] *= inv;
}

/* ------------------------------------------------------------------------ */
/* Measurement.                                                             */

/* force-measurement, pure state (src/measurement.lisp:10-41). */
EXPORT void orc_force_measurement(double *psi, uint32_t nq, uint32_t q, int measured,
                                  double excited_probability) {
    int annihilated = 1 - measured;
    double inv_norm = annihilated == 0 ? 1.0 / sqrt(excited_probability)
                                       : 1.0 / sqrt(1.0 - excited_probability);
    uint64_t len = UINT64_C(1) << nq;
    cplx *w = (cplx *)psi;
    for (uint64_t i = 0; i < len; ++i) {
        if ((int)((i >> q) & 1u) == annihilated) {
            w[i].re = 0.0;
            w[i].im = 0.0;
        } else {
            w[i].re = inv_norm * w[i].re;
            w[i].im = inv_norm * w[i].im;
        }
    }
}

/* measure, interpreted (src/measurement.lisp:87-105): p1 == 0 gives bit 0
 * without consuming a draw; otherwise bit = (u <= p1).
 * Compiled projector (src/compile-gate.lisp:221-254): always consumes a draw,
 * keep-zero = (u < p0); every amplitude is multiplied by
 * inv_norm * [bit matches].  Returns the bit; *consumed = draws used. */
EXPORT int orc_measure(double *psi, uint32_t nq, uint32_t q, double u, int compiled,
                       int *consumed) {
    if (!compiled) {
        double p1 = orc_excited_probability(psi, nq, q);
        int bit;
        if (p1 == 0.0) {
            bit = 0;
            *consumed = 0;
        } else {
            *consumed = 1;
            bit = (u <= p1) ? 1 : 0;
        }
        orc_force_measurement(psi, nq, q, bit, p1);
        return bit;
    }
    double p0 = orc_ground_probability(psi, nq, q);
    *consumed = 1;
    int keep_zero = u < p0;
    int bit = keep_zero ? 0 : 1;
    double inv_norm = keep_zero ? 1.0 / sqrt(p0) : 1.0 / sqrt(1.0 - p0);
    uint64_t len = UINT64_C(1) << nq;
    cplx *w = (cplx *)psi;
    for (uint64_t i = 0; i < len; ++i) {
        int abit = (int)((i >> q) & 1u);
        double f = inv_norm * (double)(abit == bit ? 1 : 0);
        w[i].re = f * w[i].re;
        w[i].im = f * w[i].im;
    }
    return bit;
}

/* sample-wavefunction-as-distribution-in-parallel-truly
 * (src/measurement.lisp:170-225): bisection; the lower half [min,mid) is
 * summed with psum-dotimes and added to sum<; the result is
 * min{b : C(b) > p}, or len-1 when no such b exists. */
EXPORT uint64_t orc_sample_bisect(const double *psi, uint64_t len, double p) {
    uint64_t lo = 0, hi = len;
    double sum_lt = 0.0;
    while (lo != hi - 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        sum_ctx c = {(const cplx *)psi, lo, 0, 0, {0}};
        double sum = sum_lt + psum_dotimes(&c, mid - lo);
        if (sum <= p) {
            lo = mid;
            sum_lt = sum;
        } else {
            hi = mid;
        }
    }
    return lo;
}

/* measure-all-state, pure (src/measurement.lisp:128-143): sample, reset to
 * |0...0>, rotatef psi[0] <-> psi[b] (psi[b] = 1 exactly). */
EXPORT uint64_t orc_measure_all(double *psi, uint32_t nq, double p) {
    uint64_t len = UINT64_C(1) << nq;
    uint64_t b = orc_sample_bisect(psi, len, p);
    orc_bring_to_zero_state(psi, len);
    cplx *w = (cplx *)psi;
    cplx t = w[0];
    w[0] = w[b];
    w[b] = t;
    return b;
}

/* cumulative-distribution-function: sequential FP64 running sum
 * (src/wavefunction.lisp:380-395). */
EXPORT void orc_cdf(const double *psi, uint64_t len, double *cdf) {
    const cplx *w = (const cplx *)psi;
    double s = 0.0;
    for (uint64_t i = 0; i < len; ++i) {
        s += probability(w[i]);
        cdf[i] = s;
    }
}

/* The per-draw binary search of sample-wavefunction-multiple-times
 * (src/measurement.lisp:246-288): min{i : cdf[i] >= p} clamped to len-1. */
EXPORT uint64_t orc_cdf_search(const double *cdf, uint64_t len, double p) {
    uint64_t lo = 0, hi = len - 1;
    for (;;) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (lo == hi) return lo;
        if (hi - lo == 1) return cdf[lo] < p ? hi : lo;
        if (cdf[mid] < p)
            lo = mid;
        else
            hi = mid;
    }
}

EXPORT void orc_sample_multiple(const double *psi, uint64_t len, const double *u,
                                uint64_t nshots, uint64_t *out, double *cdf_scratch) {
    orc_cdf(psi, len, cdf_scratch);
    for (uint64_t s = 0; s < nshots; ++s) out[s] = orc_cdf_search(cdf_scratch, len, u[s]);
}

/* sample-wavefunction-as-distribution (src/measurement.lisp:292-311): one
 * running sum; each p (sorted ascending) takes the first idx with
 * p < cumsum.  Entries never reached keep 0. */
EXPORT void orc_sample_as_distribution(const double *psi, uint64_t len, const double *ps,
                                       uint64_t n, uint64_t *out) {
    const cplx *w = (const cplx *)psi;
    double cumsum = 0.0;
    for (uint64_t i = 0; i < n; ++i) out[i] = 0;
    for (uint64_t idx = 0; idx < len; ++idx) {
        cumsum += probability(w[idx]);
        uint64_t i = 0;
        while (i < n && ps[i] < cumsum) {
            out[i] = idx;
            ++i;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Copy (copy-wavefunction, src/wavefunction.lisp:93-109).                  */

EXPORT void orc_copy(const double *src, double *dst, uint64_t len) {
    memcpy(dst, src, len * 2 * sizeof(double));
}
