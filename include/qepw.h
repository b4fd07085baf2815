/*
 * qepw.h -- C ABI of the gfx950 plane-wave h_psi / Davidson kernels.
 *
 * This library replaces, function for function, the native kernels that the
 * reference (sjhong6230/QE_python, qepy-pw 1.0.2) binds from its Cython
 * extension ``qepy_pw/_native_fft.pyx`` on the h_psi / Davidson path.
 * Each entry point below cites the reference interface it replaces.
 *
 * Conventions
 *   - Every pointer named a "device" buffer is HIP device memory; host
 *     buffers are named explicitly.  No entry point allocates device memory
 *     except qepw_plan_create (its index tables).  Scratch is passed in as
 *     ``workspace``; the matching *_workspace() call gives its size.
 *   - Matrices are column-major (Fortran order) with a leading dimension, as
 *     the reference's Davidson storage is.  Complex numbers are interleaved
 *     (re, im) doubles (QEPW_C128) or floats (QEPW_C64).
 *   - ``stream`` is a hipStream_t passed as void* (NULL = default stream).
 *     Every call is asynchronous on that stream.
 *   - Rows of all vectors are in the plan's DEVICE ORDER (stick-major, see
 *     qepw_plan_permutation), except where a function says otherwise.
 *   - Return value: QEPW_OK (0) or a negative error code; the message of the
 *     last failure on the calling thread is given by qepw_last_error().  No
 *     C++ exception crosses this boundary.
 */
#ifndef QEPW_H
#define QEPW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QEPW_C128 0 /* complex128 storage and FP64 arithmetic */
#define QEPW_C64 1  /* complex64 storage and FP32 arithmetic  */

#define QEPW_OP_N 0  /* op(A) = A, op(B) = B        */
#define QEPW_OP_C 1  /* op(A) = A^H                 */
#define QEPW_OP_BC 2 /* op(B) = B^H (or-ed with op) */

#define QEPW_OK 0
#define QEPW_ERR_ARG (-1)         /* invalid argument (ValueError)            */
#define QEPW_ERR_HIP (-2)         /* HIP runtime failure (RuntimeError)       */
#define QEPW_ERR_UNSUPPORTED (-3) /* shape beyond the kernels' limits         */
#define QEPW_ERR_WORKSPACE (-4)   /* workspace smaller than required          */

typedef struct qepw_plan qepw_plan;

/* Message of the last failed call on this thread ("" if none). */
const char* qepw_last_error(void);
/* ABI version, currently 1. */
int qepw_abi_version(void);

/*
 * Plan of the local-potential FFT for one k-point basis.
 * Replaces: LocalPotentialWorkspace.__init__, serial branch
 *   (qepy_pw/basis.py:460-555: linear_slots, z_major_slots,
 *   serial_stick_positions) together with NativeFFTWPlan
 *   (qepy_pw/_native_fft.pyx:1973-2086).
 * miller: host int32 (npw, 3) C-order Miller indices in reference row order;
 * slot of row g = (m0 mod n1, m1 mod n2, m2 mod n3).  ``device`` is the HIP
 * device that holds the plan's tables.
 */
int qepw_plan_create(const int32_t* miller, int64_t npw, int n1, int n2, int n3,
                     int device, qepw_plan** out);
int qepw_plan_destroy(qepw_plan* plan);
/* info[0..7] = npw, n1, n2, n3, active sticks, active x columns,
 *              sticks per z-stage workgroup, full-grid sticks. */
int qepw_plan_query(const qepw_plan* plan, int64_t* info);
/* perm[d] (host int32, npw) = reference row of device row d. */
int qepw_plan_permutation(const qepw_plan* plan, int32_t* perm);
/* QEPW_OK when qepw_apply_local can run this plan in ``precision``, else
 * QEPW_ERR_UNSUPPORTED with a message naming the grid (the FFT kernels hold
 * an xy plane's active columns, or a whole plane for non-square planes, and
 * a group of z sticks in one workgroup's LDS).  The reference's FFTW plans
 * (NativeFFTWPlan, qepy_pw/_native_fft.pyx:1973-2086) take any grid; callers
 * check here once per plan instead of at the first h_psi. */
int qepw_plan_supported(const qepw_plan* plan, int precision);
/* Device copy of the same permutation (int32, npw). */
const int32_t* qepw_plan_device_permutation(const qepw_plan* plan);

/*
 * result[:, b] = FFT^-1-sandwich(V(r) . FFT(vectors[:, b])) / N
 *                + kinetic . vectors[:, b]          (b < nb)
 * Replaces: apply_serial (qepy_pw/_native_fft.pyx:2338-2455) and
 *   apply_serial_spatial (:2458-2534, C :58-135), called from
 *   LocalPotentialWorkspace.apply (qepy_pw/basis.py:1354-1455).
 * potential: device real (n3, n2, n1) grid (plane z, row y, x fastest),
 *   double for QEPW_C128, float for QEPW_C64.
 * kinetic: device real (npw) in device order, same type, or NULL.
 * vectors / result: device complex, column-major, leading dims ldv / ldr.
 * workspace: at least qepw_apply_local_workspace(plan, precision, 1) bytes;
 *   larger workspaces process more bands per launch.
 */
size_t qepw_apply_local_workspace(const qepw_plan* plan, int precision, int nb);
int qepw_apply_local(const qepw_plan* plan, int precision, const void* vectors,
                     int64_t ldv, int nb, const void* potential, const void* kinetic,
                     void* result, int64_t ldr, void* workspace, size_t ws_bytes,
                     void* stream);
/* qepw_apply_local_timed: qepw_apply_local with the HIP events start_event /
 * stop_event (hipEvent_t) recorded on ``stream`` immediately before its first
 * and after its last kernel, so that their elapsed time spans the FFT phase's
 * kernels and not the caller's launch latency (the measurement of the phase's
 * HBM roofline; no counterpart in the reference). */
int qepw_apply_local_timed(const qepw_plan* plan, int precision, const void* vectors,
                           int64_t ldv, int nb, const void* potential, const void* kinetic,
                           void* result, int64_t ldr, void* workspace, size_t ws_bytes,
                           void* start_event, void* stop_event, void* stream);

/*
 * qepw_apply_local for ``count`` independent problems at once (k-points or
 * cells with their own plans, bands, potentials, kinetics and workspaces; one
 * precision): arrays of ``count`` entries, the arguments of qepw_apply_local
 * per problem.  Problems whose grid lengths have a grouped kernel (square
 * planes of side and z length 36, 48, 54, 60, 64, 72, 80 or 96, the team plane
 * stage, all bands in one workspace) share three launches -- z-stage inverse,
 * planes, z-stage forward -- for up to 24 problems; the others run one by
 * one.  Results equal the one-problem calls bit for bit.  The batched form of
 * the reference's per-k-point loop (qepy_pw/pw/scf.py:3340-3537) applying
 * H on each k-point's bands (LocalPotentialWorkspace.apply,
 * qepy_pw/basis.py:1354-1455).
 */
int qepw_apply_local_group(int count, const qepw_plan* const* plans, int precision,
                           const void* const* vectors, const int64_t* ldv, const int* nb,
                           const void* const* potentials, const void* const* kinetics,
                           void* const* results, const int64_t* ldr, void* const* workspaces,
                           const size_t* ws_bytes, void* stream);

/*
 * out[z, x, y] = Re sum_G V(G) exp(+i G.r) (unnormalised backward FFT) of a
 * C-order (n1, n2, n3) complex128 grid, written in (n3, n2, n1) layout.
 * Replaces: inverse_real (qepy_pw/_native_fft.pyx:2188-2202) as used by
 *   LocalPotentialWorkspace.prepare_potential (qepy_pw/basis.py:777-792).
 */
size_t qepw_inverse_real_workspace(const qepw_plan* plan);
int qepw_inverse_real(const qepw_plan* plan, const void* grid_g, double* out,
                      void* workspace, size_t ws_bytes, void* stream);

/*
 * C (m x n) = alpha op(A) op(B) + beta C, complex, column-major, on MFMA
 * (v_mfma_f64_16x16x4f64 for QEPW_C128, v_mfma_f32_16x16x4f32 for QEPW_C64).
 * op_a = QEPW_OP_N: A is m x k;  QEPW_OP_C: A is k x m and op(A) = A^H;
 * or-ed with QEPW_OP_BC: B is n x k and op(B) = B^H (else B is k x n).
 * alpha, beta: host (re, im) pairs.  A^H B with large k is split over k with
 * a fixed-order (deterministic) reduction through ``workspace``.
 * Replaces: the zgemm bridge qepy_zgemm64 (qepy_pw/_native_fft.pyx:595-634)
 *   and its users projector_overlaps (:694-731, becp = beta^H psi) and
 *   add_projector_product (:782-836, Hpsi += beta (D becp)), plus the
 *   NumPy products of davidson's update_projected / ritz_pairs
 *   (qepy_pw/diagonalization.py:658-726) and _matmul_f (:74-80).
 */
size_t qepw_gemm_workspace(int precision, int op_a, int64_t m, int64_t n, int64_t k);
int qepw_gemm(int precision, int op_a, int64_t m, int64_t n, int64_t k,
              const double* alpha, const void* a, int64_t lda, const void* b,
              int64_t ldb, const double* beta, void* c, int64_t ldc, void* workspace,
              size_t ws_bytes, void* stream);

/*
 * Tile choice of qepw_gemm (the autotuner's interface; no counterpart in the
 * reference, whose BLAS picks its own blocking).  The kernel has
 * qepw_gemm_variant_count() tile variants (TM x TN output tiles per
 * workgroup); qepw_gemm_plan writes {variant, K splits, TM, TN} of a shape
 * to out[0..3].  qepw_gemm_set_rule adds a rule -- ranges[0..5] = m_lo,
 * m_hi, n_lo, n_hi, k_lo, k_hi (inclusive) for one precision and op_a --
 * that overrides the built-in choice (first matching rule wins; at most 256
 * rules); qepw_gemm_clear_rules drops them.  Rules apply to qepw_gemm,
 * qepw_gemm_workspace and qepw_gemm_static alike: set them before products
 * run, not while another thread launches GEMMs.
 */
int qepw_gemm_variant_count(void);
int qepw_gemm_plan(int precision, int op_a, int64_t m, int64_t n, int64_t k, int64_t* out);
int qepw_gemm_set_rule(int precision, int op_a, const int64_t* ranges, int variant,
                       int64_t splits);
int qepw_gemm_clear_rules(void);

/*
 * qepw_gemm without a workspace argument and without any allocation: the
 * split-K partial tiles go to a static device array of
 * QEPW_GEMM_STATIC_SCRATCH bytes (32 MiB) that each GPU the library runs on
 * holds until the process ends (qepw_gemm_static_scratch_bytes() returns the
 * size).  Same arguments, kernels and results as qepw_gemm (bitwise: the
 * same split plan and fixed-order reduction).  A product whose split-K
 * partials would exceed the array returns QEPW_ERR_WORKSPACE with a message
 * naming the size.  Calls that overlap on one GPU share its array and race;
 * the array needs no clearing between calls (every partial tile is written
 * before it is summed).  Call with the stream's device current.
 * No counterpart in the reference: this entry point is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).
 */
#define QEPW_GEMM_STATIC_SCRATCH ((size_t)32 << 20)
size_t qepw_gemm_static_scratch_bytes(void);
int qepw_gemm_static(int precision, int op_a, int64_t m, int64_t n, int64_t k,
                     const double* alpha, const void* a, int64_t lda, const void* b,
                     int64_t ldb, const double* beta, void* c, int64_t ldc, void* stream);

/*
 * out[b] = sum_g |v[g, b]|^2 (device double, FP64 accumulation for both
 * precisions, fixed reduction order).
 * Replaces: column_squared_norms (qepy_pw/_native_fft.pyx:1153-1191).
 */
int qepw_column_squared_norms(int precision, const void* v, int64_t ldv, int64_t nrows,
                              int nb, double* out, void* stream);

/*
 * out[b] = sum_g conj(l[g, b]) r[g, b]  (device, nb (re, im) double pairs);
 * ldr = 0 pairs every column of l with the single column r (l^H r).
 * Replaces: column_inner_products (qepy_pw/_native_fft.pyx:1241-1291).
 */
int qepw_column_inner_products(int precision, const void* l, int64_t ldl, const void* r,
                               int64_t ldr, int64_t nrows, int nb, double* out,
                               void* stream);

/*
 * r[:, b] -= e[b] x[:, b]  (e: device double).
 * Replaces: subtract_band_energies (qepy_pw/_native_fft.pyx:1194-1238).
 */
int qepw_subtract_band_energies(int precision, void* r, int64_t ldr, const void* x,
                                int64_t ldx, const double* e, int64_t nrows, int nb,
                                void* stream);

/*
 * QE g_psi: for j < nb, with c = cols[j] (or j when cols is NULL):
 *   x = 2 (diag[g] - e[j]);  d = (1 + x + sqrt(1 + (x-1)^2)) / 2;
 *   out[g, j] = 2 r[g, c] / d.
 * e, diag: device double; cols: device int32 or NULL.
 * Replaces: qe_precondition (qepy_pw/_native_fft.pyx:852-899) as called by
 *   _qe_precondition (qepy_pw/diagonalization.py:475-497, 771-780).
 */
int qepw_qe_precondition(int precision, const void* r, int64_t ldr, const int32_t* cols,
                         const double* e, const double* diag, void* out, int64_t ldo,
                         int64_t nrows, int nb, void* stream);

/*
 * dst[:, j] = src[:, sel[j]] / sqrt(sq[sel[j]])  (sq: device double,
 * sel: device int32).
 * Replaces: normalize_selected_columns (qepy_pw/_native_fft.pyx:1495-1530).
 */
int qepw_normalize_selected_columns(int precision, const void* src, int64_t lds,
                                    const double* sq, const int32_t* sel, void* dst,
                                    int64_t ldd, int64_t nrows, int nb, void* stream);

/*
 * hv[:, b] = x[:, b] e[b] + r[:, b]: H psi of the Ritz vectors rebuilt from
 * the residual at a Davidson restart, without an extra H application.
 * Replaces: the restart block of davidson (qepy_pw/diagonalization.py:843-868).
 */
int qepw_restart_applied(int precision, const void* x, int64_t ldx, const void* r,
                         int64_t ldr, const double* e, void* hv, int64_t ldh,
                         int64_t nrows, int nb, void* stream);

/*
 * Row permutation with optional precision change between the reference row
 * order and the device order.  direction 0 (gather): dst[d] = src[perm[d]];
 * direction 1 (scatter): dst[perm[d]] = src[d].  perm: device int32.
 * No reference counterpart: it is the device layout change at the boundary
 * (the reference indexes rows through linear_slots, basis.py:515-523).
 */
int qepw_permute_rows(int src_precision, int dst_precision, int direction,
                      const int32_t* perm, const void* src, int64_t lds, void* dst,
                      int64_t ldd, int64_t nrows, int nb, void* stream);

/*
 * dst[:, dst_cols[j]] = scale[j] * src[:, src_cols[j]] for j < ncols, nrows
 * rows each (src_cols, dst_cols: device int32, NULL for j; scale: device
 * double, NULL for 1), widening or narrowing between the precisions.  The
 * column selections, merges and conversions of davidson's bookkeeping: the
 * Ritz coefficients of the unconverged roots and C diag(e)
 * (qepy_pw/diagonalization.py:751-757, ritz_pairs 699-726), the restart and
 * locking blocks (843-868).
 */
int qepw_copy_columns(int src_precision, int dst_precision, const void* src, int64_t lds,
                      const int32_t* src_cols, void* dst, int64_t ldd, const int32_t* dst_cols,
                      const double* scale, int64_t nrows, int ncols, void* stream);
/*
 * dst[i, j] = (i == j) ? (d ? d[i] : diag) : 0 over an nrows x ncols block
 * (zero, identity or a real diagonal; d: device double or NULL): the restart
 * block's H = diag(e), S = I and identity coefficients (diagonalization.py:
 * 843-868), and zeroed storage.
 */
int qepw_set_block(int precision, void* dst, int64_t ldd, int64_t nrows, int ncols, double diag,
                   const double* d, void* stream);

/*
 * diagonal[g] += multiplicity * Re sum_i coupled[g, i] conj(beta[g, i])
 * (complex128 inputs, device double output).
 * Replaces: accumulate_projector_diagonal (qepy_pw/_native_fft.pyx:734-779).
 */
int qepw_accumulate_projector_diagonal(const void* coupled, int64_t ldc, const void* beta,
                                       int64_t ldb, int64_t nrows, int nch,
                                       double multiplicity, double* diagonal,
                                       void* stream);

/*
 * out = scale * sum_b weights[b] |psi_b(r)|^2 with psi_b(r) the unnormalised
 * backward FFT of vectors[:, b]; out: device double grid in (n3, n2, n1)
 * layout; weights: device double (nb).  scale = 1/N gives dE/dV(r) of
 * E = sum_b w_b <psi_b|H|psi_b>.  Deterministic (fixed-order band slices).
 * Replaces: accumulate_density_serial / accumulate_density_serial_spatial
 *   (qepy_pw/_native_fft.pyx:2537-2648) as called by
 *   LocalPotentialWorkspace.accumulate_density (qepy_pw/basis.py:1210-1282).
 */
size_t qepw_accumulate_density_workspace(const qepw_plan* plan, int precision, int nb);
int qepw_accumulate_density(const qepw_plan* plan, int precision, const void* vectors,
                            int64_t ldv, int nb, const double* weights, double scale,
                            double* out, void* workspace, size_t ws_bytes, void* stream);

/*
 * dst[g, b] = s[g * sstride] cfac[b] src[g, b]  (s, cfac: device double,
 * either may be NULL for 1).  Forms the (k+G)_c psi blocks of the projector
 * derivative of the nonlocal forces (qepy_pw/pw/scf.py:1580-1590, the
 * ``1j * gk * modulated`` factor) and the occupation-weighted H psi of the
 * band-energy gradient.
 */
int qepw_scale_rows(int precision, const void* src, int64_t lds, const double* s,
                    int64_t sstride, const double* cfac, void* dst, int64_t ldd,
                    int64_t nrows, int nb, void* stream);

/*
 * y[:, j] = a[j] x[:, j] + b[j] y[:, j] for j < ncols (a, b: device double,
 * NULL for 1; y is not read where b[j] == 0).
 * Replaces: the per-column NumPy updates of the iterative solvers
 *   (qepy_pw/diagonalization.py:1105-1160 CG, 1225-1310 bpcg,
 *   1721-1729 RMM-DIIS).
 */
int qepw_column_axpby(int precision, const void* x, int64_t ldx, const double* a, void* y,
                      int64_t ldy, const double* b, int64_t nrows, int ncols, void* stream);

/*
 * out[b] = sum_g d[g] |v[g, b]|^2  (d, out: device double; FP64 accumulation,
 * fixed reduction order).
 * Replaces: column_diagonal_expectations (qepy_pw/_native_fft.pyx:1294-1343).
 */
int qepw_column_diagonal_expectations(int precision, const void* v, int64_t ldv,
                                      const double* d, int64_t nrows, int nb, double* out,
                                      void* stream);

/*
 * out[:, j] = K_j(g) r[:, c] with c = cols[j] (or j when cols is NULL),
 * x = kinetic[g] / (1.5 e[j]), p = 27 + 18x + 12x^2 + 8x^3,
 * K = -4/(3 e[j]) p / (p + 16 x^4)  (kinetic in Ry, e = <T>_j: device double).
 * Replaces: rmm_kinetic_precondition (qepy_pw/_native_fft.pyx:1346-1408).
 */
int qepw_rmm_precondition(int precision, const void* r, int64_t ldr, const int32_t* cols,
                          const double* kinetic, const double* expectations, void* out,
                          int64_t ldo, int64_t nrows, int ncols, void* stream);

/*
 * out[:, j] = sum_{i < count} coef[j, i] hist_i[:, c] with hist_i = hist + i *
 * hstride elements, c = cols[j] (or j), coef: device (ncols, count) complex
 * double pairs, row-major.
 * Replaces: the per-band DIIS combinations phi @ c and H phi @ c of rmm_diis
 *   (qepy_pw/diagonalization.py:1655-1662).
 */
int qepw_history_combine(int precision, const void* hist, int64_t ldh, int64_t hstride,
                         int count, const int32_t* cols, const double* coef, void* out,
                         int64_t ldo, int64_t nrows, int ncols, void* stream);

/*
 * out[d, a*nch + i] = beta[r, i] phases[r, a] with r = perm[d] (perm: device
 * int32, NULL for the identity): the atom blocks of a factorized projector
 * term packed on the device, optionally gathered into a plan's device order.
 * Replaces: the per-atom phase products of FactorizedProjectorTerm applied in
 *   PlaneWaveHamiltonian._apply (qepy_pw/diagonalization.py:21-28, 293-309).
 */
int qepw_phase_blocks(int precision, const void* beta, int64_t ldb, int nch,
                      const void* phases, int64_t ldp, int natom, const int32_t* perm, void* out,
                      int64_t ldo, int64_t nrows, void* stream);

/*
 * out[g, a*nch + i] = beta[g, i] exp(-i q_g . tau_a): the atom-packed
 * projector matrix recomputed from positions (q: device double (npw, 3)
 * Cartesian k+G, tau: device double (natom, 3) bohr).
 * Replaces: phase_matrix (qepy_pw/_native_fft.pyx:1664-1699) and the packed
 *   assembly of _nonlocal_projector_terms (qepy_pw/pw/scf.py:657-686).
 */
int qepw_build_projectors(int precision, const void* beta, int64_t ldb, int nch,
                          const double* q, const double* tau, int natom, void* out,
                          int64_t ldo, int64_t nrows, void* stream);

/*
 * beta[g, j] = table[i, p_j] interpolated at |q_g| times Y_{l_j m_j}(q_g)
 * (-i)^{l_j}: the per-k assembly of one species' projector channels.
 * q: device double (npw, 3) Cartesian k+G (any row order; beta follows it);
 * table: device double, column-major (nq x nradial, leading dimension ldt),
 * radial transforms on the uniform grid i * dq, interpolated with QE's
 * four-point cubic formula (samples i0 .. i0+3, i0 = floor(|q| / dq)): the
 * table must extend to floor(max |q| / dq) + 3; chan_radial, chan_l, chan_m:
 * device int32 (nch), radial index, l <= 3 and QE ylmr2 index m (0, then
 * cos and sin of 1..l); beta: device complex128 (npw x nch, column-major).
 * Replaces: the per-k beta assembly of UPFPseudo.projector_basis
 *   (qepy_pw/upf.py:483-541) with _qe_cubic_interpolate (upf.py:880-912)
 *   and _qe_real_spherical_harmonics (upf.py:722-772) -- QE init_us_2; the
 *   per-atom phases are qepw_build_projectors.
 */
int qepw_projector_basis(const double* q, int64_t npw, const double* table, int64_t ldt,
                         int nq, double dq, const int32_t* chan_radial, const int32_t* chan_l,
                         const int32_t* chan_m, int nch, void* beta, int64_t ldb, void* stream);

/*
 * out[a, c] = 2 sum_n f_n sum_{i in atom a} Re(conj(dO_c[i, n]) DO[i, n]):
 * dE/dtau (the negative force) of the nonlocal energy, with dO_c = dO/dtau_c
 * (three (P, nb) blocks, block stride dstride), DO = D O, atom a owning
 * channels [off[a], off[a+1]) (device int32), f: device double (nb).
 * Replaces: the force branch of _nonlocal_derivatives
 *   (qepy_pw/pw/scf.py:1591-1599).
 */
int qepw_projector_forces(int precision, const void* dO, int64_t lddo, int64_t dstride,
                          const void* DO, int64_t lddO, const double* f, int nb,
                          const int32_t* off, int natom, double* out, void* stream);

/*
 * Pieces of the device Rayleigh-Ritz that replaces the host zhegvd call of
 * davidson (generalized_eigh, qepy_pw/_native_fft.pyx:667-691, via
 * _lowest_generalized_eigh, qepy_pw/diagonalization.py:83-110).  All operate
 * on column-major complex128 matrices in device memory.
 *
 * qepw_hermitian_complete: rows [first, active) of A were just written for
 *   columns [0, active); mirror them into the upper block and symmetrise the
 *   new diagonal block (update_projected, diagonalization.py:658-697).
 * qepw_shift_scale: bound[0] = max_i sum_j |A_ij - sigma delta_ij| (device
 *   double, an upper bound of ||A - sigma I||_2) and X = (A - sigma I)/bound;
 *   ``bound`` holds n + 1 doubles (bound[1 + j] = column sums).
 * qepw_hermitize: X = (X + X^H)/2 in place.
 * qepw_sign_check: out[0] = ||X2 - I||_F^2, out[1] = Re trace(X) (device);
 *   ``out`` holds n + 2 doubles (out[2 + j] = per-column partial sums).
 * qepw_set_diagonal: A[:n, :n] = diag(d), identity when d is NULL.
 * qepw_sign_iteration: the scaled Newton-Schulz steps
 *   X <- (a/2) X (3 I - a^2 X^2), one per entry of alphas[0:nalpha], on the
 *   n x n complex128 X held in x, which must be exactly Hermitian (as
 *   qepw_shift_scale leaves it); x2 and y: scratch of the same shape and
 *   leading dimension.  Every iterate is exactly Hermitian; the result is
 *   left in x2 when *in_x2 is set, else in x, and y is left holding its
 *   square (the input of qepw_sign_check).  Workspace:
 *   qepw_sign_iteration_workspace(n) bytes (none at present).
 */
int qepw_hermitian_complete(void* a, int64_t lda, int first, int active, void* stream);
int qepw_shift_scale(const void* a, int64_t lda, int n, double sigma, void* x, int64_t ldx,
                     double* bound, void* stream);
int qepw_hermitize(void* x, int64_t ldx, int n, void* stream);
int qepw_sign_check(const void* x, int64_t ldx, const void* x2, int64_t ld2, int n,
                    double* out, void* stream);
int qepw_set_diagonal(void* a, int64_t lda, int n, const double* d, void* stream);
size_t qepw_sign_iteration_workspace(int n);
int qepw_sign_iteration(void* x, void* x2, void* y, int64_t ld, int n, const double* alphas,
                        int nalpha, int* in_x2, void* workspace, size_t workspace_bytes,
                        void* stream);

/* ---- native Rayleigh-Ritz steps ------------------------------------------
 * The reduced eigensolve of one Davidson iteration (replacing, with the device
 * spectral division above, the reference's zhegvd of the copied pencil:
 * generalized_eigh, qepy_pw/_native_fft.pyx:667-691, called from
 * qepy_pw/diagonalization.py:699-726) as two host calls that launch their
 * kernels on ``stream``, read back the few small matrices they need (pinned
 * memory, waiting for the stream), factorise them with the host LAPACK and
 * upload the factors.  All matrices are column-major complex128 with leading
 * dimension ``maxsub``.
 *
 * qepw_set_host_lapack: the LAPACK routines used (Fortran calling convention:
 *   zpotrf, ztrtri, zheevd, cheevd), e.g. those of the caller's NumPy/SciPy;
 *   required before the calls below (QEPW_ERR_UNSUPPORTED otherwise).
 * qepw_rr_workspace: device workspace bytes of both calls.
 * qepw_rr_standardize: with S = L L^H and W = L^-1 known for rows and columns
 *   [0, first), extend W to [0, active) and write rows [first, active) of
 *   A = W H W^H (A Hermitian-completed).  *not_pd = 1 (and nothing written)
 *   when the new Schur complement of S is not positive definite.
 * qepw_rr_spectral: the lowest k eigenpairs of the n x n A from the shift
 *   ``sigma`` with ``gap`` the spacing around it: the scaled Newton-Schulz
 *   iteration to sign(A - sigma I) (scratch X, X2, Y), a range sample of the
 *   projector below the shift from the previous eigenvectors ``u_prev``
 *   (k columns, when have_prev) and Gaussian columns ``gauss`` (maxsub
 *   columns), a Rayleigh-Ritz on it.  ``rb``: a device buffer of
 *   2 maxsub + 4 + 2 maxsub^2 doubles (sign check, bound, Gram matrix).  On
 *   status 0: values[0:m_below] (host) the eigenvalues below the shift,
 *   ``coef`` (n x k, leading dimension n) = W^H U for the eigenvectors U of
 *   A, and U stored in u_prev.  status 1: the sign iteration did not
 *   converge; 2: fewer than k eigenvalues below the shift; 3: the sample's
 *   Gram matrix is not positive definite; 4: the host eigensolve failed.
 *   ``bound_prev`` > 0 estimates ||A - sigma I||_1 as twice that value (else
 *   it is read back); ``bound`` returns the exact value of this call.
 */
typedef struct qepw_rr_spectral_args {
  int n;
  int k;
  int maxsub;
  int count_hint;  /* range sample columns to try first */
  int host_single; /* the host eigensolve in complex64 */
  int have_prev;
  double sigma;
  double gap;
  double bound_prev;
  void* A;
  void* W;
  void* X;
  void* X2;
  void* Y;
  void* u_prev;
  void* gauss;
  void* coef;
  double* rb;
  void* work;
  size_t work_bytes;
  void* stream;
} qepw_rr_spectral_args;

typedef struct qepw_rr_spectral_result {
  int status;
  int m_below;
  int ns_steps;
  int slow_shift;
  double bound;
} qepw_rr_spectral_result;

int qepw_set_host_lapack(void* zpotrf, void* ztrtri, void* zheevd, void* cheevd);
size_t qepw_rr_workspace(int maxsub);
int qepw_rr_standardize(void* h, void* s, void* w, void* a, int maxsub, int first, int active,
                        int* not_pd, void* workspace, size_t workspace_bytes, void* stream);
int qepw_rr_spectral(const qepw_rr_spectral_args* args, double* values,
                     qepw_rr_spectral_result* result);
/* qepw_cholesky_qr: orthonormalise the nb columns of an npw x nb block
 * (complex128 or complex64, leading dimension ldb) by Cholesky QR -- Gram
 * matrix on the device, its Cholesky factor in FP64 on the host, out =
 * block L^-H -- for the corrections of the mixed form's complex64 phase
 * (the orthonormalisation step of the reference's Davidson,
 * qepy_pw/diagonalization.py:450-472, in Cholesky form).  *ok = 0 and
 * ``out`` untouched when the Gram matrix is not positive definite or
 * (max/min diag L)^2 exceeds ``condition`` (the caller then falls back to
 * the canonical orthonormalisation).  Workspace: qepw_rr_workspace(nb). */
int qepw_cholesky_qr(int precision, const void* block, int64_t ldb, int64_t npw, int nb,
                     double condition, void* out, int64_t ldo, int* ok, void* workspace,
                     size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEPW_H */
