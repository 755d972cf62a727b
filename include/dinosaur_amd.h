/*
 * dinosaur_amd C ABI: MI355X (gfx950) kernels for the spectral dynamical core
 * of neuralgcm/dinosaur.  Plain pointers and sizes only; every device buffer
 * is allocated by the caller; every compute call is asynchronous on the
 * caller's HIP stream (passed as void*, 0 = default stream) and does no host
 * synchronisation and no device allocation.
 *
 * Layouts (row-major, last index contiguous), matching the reference:
 *   modal  (batch, 2M-1, L)   Real layout, m rows ordered [0, +1, -1, +2, ...]
 *                              (dinosaur/spherical_harmonic.py:192-200)
 *   nodal  (batch, I, J)      longitude x sin(latitude), latitude contiguous
 *                              (dinosaur/spherical_harmonic.py:186-188)
 *   fourier (batch, 2M-1, J)  longitude-transformed intermediate
 *
 * Return codes: DINO_OK, or an error whose message dino_last_error() returns
 * (thread-local).  The Python layer maps DINO_EINVAL -> ValueError,
 * DINO_ESHAPE -> StateShapeError, others -> RuntimeError.
 */
#ifndef DINOSAUR_AMD_H_
#define DINOSAUR_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DINO_OK = 0,
  DINO_EINVAL = 1,
  DINO_ESHAPE = 2,
  DINO_EHIP = 3,
  DINO_ENOMEM = 4,
};

/* element type of caller buffers */
enum { DINO_F32 = 0, DINO_F64 = 1 };

/* arithmetic of the Legendre contraction (FP32 data only for the split forms) */
enum {
  DINO_PREC_NATIVE = 0, /* FP32 FMA for FP32 data, FP64 FMA for FP64 data */
  /* 1 is unused (a split-FP16 form loses small-magnitude fields) */
  DINO_PREC_BF16X3 = 2, /* split-BF16 x3 on MFMA, FP32 accumulation */
};

/* Flag or-ed into `precision` of dino_sht_forward, dino_sht_inverse_adjoint
 * and dino_legendre_forward: each modal row of zonal wavenumber m is written
 * only from degree s(m) = max(0, L - 32 * ceil((L - m) / 32)) on (the first
 * degree of the 32-degree tiles that cover m .. L-1, aligned to end at L),
 * and each entry below (all zero, since P_l^m vanishes for l < m) either
 * keeps what the buffer held or is set to zero: the rows are stored in
 * whole 128-byte lines, so the line that holds degree s(m) and the line a
 * row shares with the start of the next row of the field may zero a few of
 * them (at most 31 per row side).  For callers whose consumers never read
 * them -- the inverse Legendre kernels start at s(m), and the fused
 * explicit-terms kernels treat
 * l < m as zero -- it saves writing about 45% of a TL255 modal array.  No
 * reference counterpart (XLA writes the dense einsum result). */
enum { DINO_SHT_TRIANGLE_ONLY = 0x100 };

typedef struct dino_plan dino_plan;
typedef struct dino_implicit dino_implicit;

/* Last error message of the calling thread ("" if none). */
const char* dino_last_error(void);

/* Library version string. */
const char* dino_version(void);

/* FP64 Gauss-Legendre nodes (ascending) and weights on the host, by the same
 * Newton iteration the plan runs on the GPU.
 * Replaces associated_legendre.gauss_legendre_nodes
 * (dinosaur/associated_legendre.py:104-114). */
int dino_gauss_nodes_host(int n, double* nodes, double* weights);

/* Builds a transform plan: FP64 Gauss nodes/weights (or the given nodes and
 * weights) and the FP64 normalised P_l^m table generated on the GPU, folded
 * onto the northern hemisphere, plus FP32 / split-BF16 copies and the
 * longitude twiddle table.  nodes/weights: host arrays of length J,
 * ascending and symmetric about 0, or NULL for Gauss nodes.
 * Replaces RealSphericalHarmonics.basis
 * (dinosaur/spherical_harmonic.py:219-260) together with
 * associated_legendre.evaluate (:70-101) and fourier.real_basis
 * (dinosaur/fourier.py:24-60). */
int dino_plan_create(int M, int L, int I, int J, const double* nodes,
                     const double* weights, dino_plan** out);
int dino_plan_destroy(dino_plan* plan);

/* Copies the plan's FP64 nodes/weights (length J) to host arrays. */
int dino_plan_nodes(const dino_plan* plan, double* nodes, double* weights);

/* Copies the folded FP64 P table to a host array: for each m, rows l = m..L-1,
 * each of Jh = ceil(J/2) values at the nodes j = J-Jh .. J-1 (test hook). */
int64_t dino_plan_table_size(const dino_plan* plan);
int dino_plan_table(const dino_plan* plan, double* out);

/* Launch schedules of the four transform stages.  A schedule is a kernel
 * family in the low byte plus flags; every family computes the same
 * contraction (FP32 data; FP64 data always runs the FP64 kernels), so a
 * schedule changes speed and FP32 rounding order, never the result beyond
 * rounding.  The project's autotuner (dinosaur_amd/autotune.py) times the
 * candidates for a batch size and sets the fastest.  Set schedules before
 * launching work that uses the plan: the plan is read at launch time, so a
 * change races with transforms running concurrently on other threads.  The
 * reference has no counterpart (XLA picks its einsum kernels internally). */
enum {
  DINO_STAGE_LEGENDRE_INVERSE = 0,
  DINO_STAGE_LEGENDRE_FORWARD = 1,
  DINO_STAGE_DFT_INVERSE = 2,
  DINO_STAGE_DFT_FORWARD = 3,
};
enum {
  DINO_SCHED_AUTO = 0,         /* built-in choice (row-resident / persistent where possible) */
  DINO_SCHED_ROWS = 1,         /* Legendre: row-resident kernel, all degree tiles of one m per workgroup */
  DINO_SCHED_TILED = 2,        /* Legendre: (rows x nodes/degrees) tiled kernel, wide tiles */
  DINO_SCHED_TILED_NARROW = 3, /* Legendre: tiled kernel, narrow tiles (more workgroups) */
  DINO_SCHED_ROWS_WIDE = 4,    /* forward Legendre: row-resident kernel with 128-row
                                  workgroups (half the P-table traffic per row) */
  DINO_SCHED_PERSISTENT = 1,   /* DFT: persistent FP32 FFT workgroups (one wave per CU) */
  DINO_SCHED_PER_TILE = 2,     /* DFT: one workgroup per (field, latitude tile) */
  DINO_SCHED_MPAIRED = 0x100,  /* Legendre flag: m order 0, M-1, 1, M-2, ... (long and
                                  short triangular slabs alternate) instead of 0, 1, 2, ... */
};
int dino_plan_set_schedule(dino_plan* plan, int stage, int schedule);
/* Current schedule of a stage (>= 0), or -1 for an invalid plan or stage. */
int dino_plan_get_schedule(const dino_plan* plan, int stage);

/* Bytes of workspace the transforms below need for `batch` fields. */
size_t dino_sht_workspace_bytes(const dino_plan* plan, int64_t batch,
                                int dtype);

/* modal -> nodal.  Replaces RealSphericalHarmonics.inverse_transform
 * (dinosaur/spherical_harmonic.py:262-269) mapped over a batch by
 * Grid.to_nodal (:1017-1023). */
int dino_sht_inverse(const dino_plan* plan, const void* modal, void* nodal,
                     void* work, int64_t batch, int dtype, int precision,
                     void* stream);

/* dino_sht_inverse without the workspace argument: the (batch, 2M-1, J)
 * Fourier rows live in a static device array of DINO_SHT_SCRATCH_BYTES
 * (64 MiB) that the library holds on every GPU it runs on until the process
 * ends; the call allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Returns
 * DINO_EINVAL when dino_sht_workspace_bytes(plan, batch, dtype) exceeds the
 * array.  Every call on one GPU shares that GPU's array: calls that overlap
 * in time on one GPU (other streams or threads) race. */
#define DINO_SHT_SCRATCH_BYTES ((size_t)64 << 20)
int dino_sht_inverse_scratch(const dino_plan* plan, const void* modal,
                             void* nodal, int64_t batch, int dtype,
                             int precision, void* stream);

/* nodal -> modal with quadrature weights.  Replaces
 * RealSphericalHarmonics.transform (:271-280) via Grid.to_modal (:1025-1031). */
int dino_sht_forward(const dino_plan* plan, const void* nodal, void* modal,
                     void* work, int64_t batch, int dtype, int precision,
                     void* stream);

/* Adjoints (vector-Jacobian products): inverse_adjoint maps a nodal
 * cotangent to a modal one (P^T F^T y), forward_adjoint maps a modal
 * cotangent to a nodal one (w * F P x).  The reference obtains these by XLA
 * transposition of the einsums at :265-279. */
int dino_sht_inverse_adjoint(const dino_plan* plan, const void* nodal_cot,
                             void* modal_cot, void* work, int64_t batch,
                             int dtype, int precision, void* stream);
int dino_sht_forward_adjoint(const dino_plan* plan, const void* modal_cot,
                             void* nodal_cot, void* work, int64_t batch,
                             int dtype, int precision, void* stream);

/* Individual stages (for tests and roofline measurement).
 * dft_forward: nodal (B,I,J) -> fourier (B,2M-1,J); weighted != 0 multiplies
 *   by the quadrature weight (2 pi / I) * w_j (the reference's `w * x`, :275).
 * dft_inverse: fourier -> nodal; weighted != 0 multiplies the output by the
 *   same weight (used by the forward adjoint).
 * legendre_inverse: modal -> fourier; legendre_forward: fourier -> modal. */
int dino_dft_forward(const dino_plan* plan, const void* nodal, void* fourier,
                     int64_t batch, int dtype, int weighted, void* stream);
int dino_dft_inverse(const dino_plan* plan, const void* fourier, void* nodal,
                     int64_t batch, int dtype, int weighted, void* stream);
int dino_legendre_inverse(const dino_plan* plan, const void* modal,
                          void* fourier, int64_t batch, int dtype,
                          int precision, void* stream);
int dino_legendre_forward(const dino_plan* plan, const void* fourier,
                          void* modal, int64_t batch, int dtype,
                          int precision, void* stream);

/* Semi-implicit solve (1 - eta * implicit_terms)^{-1} for the divergence,
 * temperature-variation and log-surface-pressure fields, factorised once per
 * (grid, levels, eta) in FP64 through the vertical modes of the l-independent
 * matrix A = G H + R T_ref dsigma^T:
 *   (I - eta^2 lambda_l A)^{-1} = V diag(1 / (1 - eta^2 lambda_l d)) V^{-1}.
 * All matrices are K x K row-major FP64 host arrays: G (geopotential
 * weights), H (temperature implicit weights, as returned by
 * get_temperature_implicit_weights_sigma), V and Vinv (eigenvectors of A and
 * their inverse), d (eigenvalues of A, length K); tref and dsigma length K;
 * lam length L (laplacian eigenvalues).
 * Replaces PrimitiveEquationsSigma.implicit_inverse with the host
 * np.linalg.inv of _get_implicit_term_matrix_sigma
 * (dinosaur/primitive_equations.py:1251-1458, :544-598). */
int dino_implicit_create(int K, int L, int R, const double* G, const double* H,
                         const double* tref, const double* dsigma,
                         const double* lam, const double* V,
                         const double* Vinv, const double* d, double r_gas,
                         double eta, dino_implicit** out);
int dino_implicit_destroy(dino_implicit* f);

/* Reads div/temp (members, K, R, L) and lnps (members, 1, R, L) and writes
 * the three outputs.  Outputs either do not overlap the inputs or equal them
 * (in place: div_out == div, temp_out == temp, lnps_out == lnps); the same
 * holds for dino_implicit_terms and both adjoints. */
int dino_implicit_solve(const dino_implicit* f, const void* div,
                        const void* temp, const void* lnps, void* div_out,
                        void* temp_out, void* lnps_out, int64_t members,
                        int dtype, void* stream);

/* Implicit tendencies (primitive_equations.py:1196-1249): writes
 *   div_t  = -lambda_l (G T' + R T_ref lnps)
 *   temp_t = -H delta
 *   lnps_t = -dsigma . delta
 * using the same factor object's G, H, tref, dsigma and lam (eta unused). */
int dino_implicit_terms(const dino_implicit* f, const void* div,
                        const void* temp, const void* lnps, void* div_out,
                        void* temp_out, void* lnps_out, int64_t members,
                        int dtype, void* stream);

/* Vector-Jacobian products of dino_implicit_solve and dino_implicit_terms:
 * given the cotangents of (div_out, temp_out, lnps_out), write the
 * cotangents of (div, temp, lnps).  They stand in for what jax.vjp derives
 * from PrimitiveEquationsSigma.implicit_inverse and implicit_terms
 * (dinosaur/primitive_equations.py:1196-1458) when the reference is
 * differentiated; same layouts and no-alias rule as the forward calls. */
int dino_implicit_solve_adjoint(const dino_implicit* f, const void* div_cot,
                                const void* temp_cot, const void* lnps_cot,
                                void* div_out, void* temp_out, void* lnps_out,
                                int64_t members, int dtype, void* stream);
int dino_implicit_terms_adjoint(const dino_implicit* f, const void* div_cot,
                                const void* temp_cot, const void* lnps_cot,
                                void* div_out, void* temp_out, void* lnps_out,
                                int64_t members, int dtype, void* stream);
/* As dino_implicit_terms_adjoint, but adds the cotangents to what the output
 * buffers hold (out += G^T cot): the accumulation jax.vjp performs when a
 * state feeds both explicit_terms and implicit_terms
 * (dinosaur/time_integration.py:361-364), fused into the kernel. */
int dino_implicit_terms_adjoint_accumulate(const dino_implicit* f,
                                           const void* div_cot,
                                           const void* temp_cot,
                                           const void* lnps_cot, void* div_out,
                                           void* temp_out, void* lnps_out,
                                           int64_t members, int dtype,
                                           void* stream);

/* IMEX Runge-Kutta stage combination on flat state buffers of n elements:
 *   out = y0 + dt * (a[0] f[0] + ... + a[n_ex-1] f[n_ex-1])
 *            + dt * (b[0] g[0] + ... + b[n_im-1] g[n_im-1])
 * with the roundings of the reference's elementwise passes, in order
 * (each product, partial sum and scaling rounded to dtype; coefficients
 * rounded to dtype first).  Replaces the tree_map arithmetic of
 * imex_runge_kutta (dinosaur/time_integration.py:351-370).  At most 4
 * terms per group; a group may be empty.  out may alias y0. */
int dino_imex_stage(const void* y0, int n_ex, const void* const* f,
                    const double* a, int n_im, const void* const* g,
                    const double* b, double dt, void* out, int64_t n,
                    int dtype, void* stream);

/* Cotangents of the f and g operands of dino_imex_stage given the output
 * cotangent: f_cot[j] = a[j] * (dt * out_cot), g_cot[k] = b[k] * (dt *
 * out_cot) (the y0 cotangent is out_cot itself).  What jax.vjp derives for
 * the same arithmetic in the reference. */
int dino_imex_stage_adjoint(const void* out_cot, int n_ex, const double* a,
                            int n_im, const double* b, double dt,
                            void* const* f_cot, void* const* g_cot, int64_t n,
                            int dtype, void* stream);

/* dino_imex_stage with implicit terms stored only where they can be nonzero:
 * element i of every g[k] is g[k][i - g_origin] for im_begin <= i < im_end,
 * and the implicit group adds nothing outside that range (the reference
 * adds dt * b * 0 there: the same value).  For the primitive equations the
 * range is the divergence, temperature and ln ps part of the flat state:
 * vorticity and tracers have no implicit tendency
 * (dinosaur/primitive_equations.py:1230-1249).  out may alias y0.
 * degrees > 0 declares the buffers whole Real-layout modal fields of
 * field_size = (2M - 1) * degrees elements: entries with l < m (the part of
 * the dense array outside the triangular truncation, zero in every operand
 * of a masked state) are written as zero, whatever the operands hold there
 * -- a 16-byte vector of them that lies wholly below the triangle without
 * reading any operand.  degrees = 0: plain elementwise.
 * DINO_STAGE_TRIANGLE_ONLY OR-ed into dtype (with degrees > 0): the vectors
 * wholly below the triangle are not written either -- out keeps what it
 * held there -- for a stage value that only kernels reading the triangle
 * consume (the implicit solve and terms are column-local, the explicit
 * terms read l >= m); also accepted by dino_imex_stage_ranged_scaled. */
enum { DINO_STAGE_TRIANGLE_ONLY = 0x100 };
int dino_imex_stage_ranged(const void* y0, int n_ex, const void* const* f,
                           const double* a, int n_im, const void* const* g,
                           const double* b, double dt, void* out, int64_t n,
                           int64_t im_begin, int64_t im_end, int64_t g_origin,
                           int64_t field_size, int degrees, int dtype,
                           void* stream);

/* dino_imex_stage_ranged followed by a per-degree scaling of the result,
 * out[i] = y[i] * degree_scale[i % scale_len] rounded once (scale_len
 * elements of the data type, dividing n): an IMEX step's output combination
 * with the exponential or diffusion step filter (filtering.py:38-100 via
 * time_integration.py:427-486, one elementwise product after the step in
 * the reference) in the same pass. */
int dino_imex_stage_ranged_scaled(const void* y0, int n_ex, const void* const* f,
                                  const double* a, int n_im, const void* const* g,
                                  const double* b, double dt, void* out, int64_t n,
                                  int64_t im_begin, int64_t im_end,
                                  int64_t g_origin, int64_t field_size,
                                  int degrees, const void* degree_scale,
                                  int64_t scale_len, int dtype, void* stream);

/* Cotangent targets of one stage combination, written in one pass over the
 * output cotangent gbar: for each target t,
 *   v = gbar                       (kind has DINO_TARGET_IDENTITY: y0),
 *   v = coef[t] * (dt * gbar)      (otherwise: an f or g term),
 *   out[t] = out[t] + v            (kind has DINO_TARGET_ACCUMULATE),
 *   out[t] = v                     (otherwise);
 * DINO_TARGET_RANGED targets are ranged implicit terms (indexing as in
 * dino_imex_stage_ranged; nothing is written outside the range).  This is
 * the reverse sweep of imex_runge_kutta (time_integration.py:351-370) with
 * the accumulation of each term's cotangent over the stages that use it
 * fused into the pass.  At most 9 targets; out buffers must not overlap
 * gbar or each other.  field_size / degrees as for dino_imex_stage_ranged:
 * where the kernel skips the triangle, gbar is not read, written targets
 * get zero and accumulating ones are left as they are. */
enum {
  DINO_TARGET_ACCUMULATE = 1,
  DINO_TARGET_IDENTITY = 2,
  DINO_TARGET_RANGED = 4,
};
int dino_imex_stage_adjoint_targets(const void* gbar, int n_targets,
                                    void* const* out, const double* coef,
                                    const int* kind, double dt, int64_t n,
                                    int64_t im_begin, int64_t im_end,
                                    int64_t g_origin, int64_t field_size,
                                    int degrees, int dtype, void* stream);

/* One cotangent of a reverse IMEX sweep gathered from all its sources in one
 * pass: out = v_0 + v_1 + ... + v_{n_src-1} in source order, with
 *   x_k = src[k] * degree_scale[i % scale_len]  (flags[k] has
 *         DINO_GATHER_SCALED: the cotangent of a per-degree filter applied
 *         to the step output), else x_k = src[k];
 *   v_k = x_k                      (flags[k] has DINO_GATHER_IDENTITY: a y0
 *         contribution),
 *   v_k = coef[k] * (dt * x_k)     (otherwise: a stage coefficient);
 * the same roundings and order as one dino_imex_stage_adjoint_targets pass
 * writing v_0 followed by passes accumulating v_1, v_2, ... -- so a sweep
 * may wait until a term's cotangent is needed and form it once instead of
 * reading and writing it at every stage (time_integration.py:351-370 under
 * reverse-mode differentiation).  Sources are full-size (n elements);
 * ranged != 0 writes a ranged implicit cotangent (indexing as in
 * dino_imex_stage_ranged, nothing written outside [im_begin, im_end)).
 * field_size / degrees as for dino_imex_stage_ranged: zero where the kernel
 * skips the triangle.  A source flagged DINO_GATHER_IMPLICIT_RANGE holds
 * values only on [im_begin, im_end) (the cotangent of a stage value that
 * only the implicit terms read): it counts as zero elsewhere and is not
 * read there.  1 to 8 sources; out must not overlap them; degree_scale
 * (scale_len elements, scale_len dividing n) may be NULL when no source is
 * scaled.  DINO_STAGE_TRIANGLE_ONLY OR-ed into dtype: the entries that the
 * kernel would write as zero below the triangle are not written (a
 * cotangent that only kernels reading l >= m consume). */
enum { DINO_GATHER_IDENTITY = 1, DINO_GATHER_SCALED = 2, DINO_GATHER_IMPLICIT_RANGE = 4 };
int dino_imex_adjoint_gather(const void* const* src, const double* coef,
                             const int* flags, int n_src, double dt,
                             void* out, int64_t n, int ranged,
                             int64_t im_begin, int64_t im_end,
                             int64_t g_origin, int64_t field_size,
                             int degrees, const void* degree_scale,
                             int64_t scale_len, int dtype, void* stream);

/* Fused nodal nonlinear terms of PrimitiveEquationsSigma.explicit_terms.
 * Replaces compute_diagnostic_state_sigma and the nodal parts of
 * curl_and_div_tendencies, kinetic_energy_tendency,
 * horizontal_scalar_advection, nodal_temperature_vertical_tendency,
 * nodal_temperature_adiabatic_tendency, nodal_log_pressure_tendency and the
 * humidity terms (dinosaur/primitive_equations.py:168-225, 778-805,
 * 861-1124) with sigma_coordinates.centered_vertical_advection (:344-405).
 *
 * Buffers are (fields, I, J).  Input field groups, in order, each
 * (members, levels): Z D T TRACER[0..n) U V GX GY [LAP_LNPS QX QY]
 * (levels K except GX, GY, LAP_LNPS: 1; the bracketed groups only when
 * humidity_index >= 0).  Output groups: CU CV KE TEMP UT VT LSP
 * (TRN TU TV)[0..n) [CURL DC]; with humidity KE holds the kinetic energy
 * plus the virtual-temperature geopotential, which the divergence tendency
 * takes through the same -laplacian (the reference transforms the two
 * separately).  Per-level constants are host FP64
 * arrays of length K (c2c_inv and sum_sigma use the first K-1 entries);
 * sec2 and coriolis have length J; geo_off[k] = G[0][k] (k >= 1, geo_off[0]
 * unused) and geo_diag[k] = G[k][k] of the geopotential matrix (may be NULL
 * for dry equations). */
typedef struct dino_nodal dino_nodal;
int dino_nodal_create(int K, int I, int J, int n_tracers, int humidity_index,
                      const int* cloud_indices, int n_clouds,
                      int include_vertical_advection, const double* sec2,
                      const double* coriolis, const double* dsigma,
                      const double* c2c_inv, const double* sum_sigma,
                      const double* alpha, const double* tref,
                      const double* geo_off, const double* geo_diag,
                      double r_gas, double kappa, double r_vapor, double cp,
                      double cp_vapor, dino_nodal** out);
int dino_nodal_destroy(dino_nodal* n);
/* Number of input (output = 0) or output (output = 1) fields for `members`. */
int64_t dino_nodal_fields(const dino_nodal* n, int64_t members, int output);
int dino_nodal_explicit(const dino_nodal* n, const void* nodal_in,
                        void* nodal_out, int64_t members, int dtype,
                        void* stream);
/* Vector-Jacobian product: input cotangents from output cotangents (the
 * reference obtains this by XLA autodiff of the same elementwise graph). */
int dino_nodal_explicit_adjoint(const dino_nodal* n, const void* nodal_in,
                                const void* out_cot, void* in_cot,
                                int64_t members, int dtype, void* stream);

/* Fused modal-space stages of PrimitiveEquationsSigma.explicit_terms.
 * State buffers are (fields, 2M-1, L) with groups vorticity (K),
 * divergence (K), temperature_variation (K), log_surface_pressure (1),
 * tracers (K each), each (members, levels).  dino_modal_prep writes the
 * inverse-transform input groups of dino_nodal_explicit; dino_modal_epilogue
 * reads the forward-transform output groups of dino_nodal_explicit and
 * writes the state-shaped tendency (orography: (2M-1, L) modal orography or
 * NULL).  The adjoints map cotangents back.  Replaces get_cos_lat_vector,
 * cos_lat_grad (spherical_harmonic.py:1104-1178) and the modal combination
 * of explicit_terms (primitive_equations.py:1001-1059, 778-805, 861-963,
 * 1126-1194).  Coefficients: inv_lap, lap, clip_mask of length L; up, dn
 * ((l+1) a_lm, -l b_lm) and up2, dn2 ((l-1) a_lm, -(l+2) b_lm) of shape
 * (2M-1, L), from Grid._derivative_recurrence_weights (:1063-1070). */
typedef struct dino_modal dino_modal;
int dino_modal_create(int R, int L, int K, int n_tracers, int humidity_index,
                      double radius, double gravity, const double* inv_lap,
                      const double* lap, const double* clip_mask,
                      const double* up, const double* dn, const double* up2,
                      const double* dn2, dino_modal** out);
int dino_modal_destroy(dino_modal* m);
/* Fields of the state (which = 0), prep output (1) or epilogue input (2). */
int64_t dino_modal_fields(const dino_modal* m, int64_t members, int which);
int dino_modal_prep(const dino_modal* m, const void* state, const void* unused,
                    void* out, int64_t members, int dtype, void* stream);
int dino_modal_prep_adjoint(const dino_modal* m, const void* out_cot,
                            const void* unused, void* state_cot,
                            int64_t members, int dtype, void* stream);
int dino_modal_epilogue(const dino_modal* m, const void* in,
                        const void* orography, void* tendency, int64_t members,
                        int dtype, void* stream);
int dino_modal_epilogue_adjoint(const dino_modal* m, const void* tendency_cot,
                                const void* unused, void* in_cot,
                                int64_t members, int dtype, void* stream);

/* dino_modal_create with flags.  DINO_MODAL_HYBRID: the state is on hybrid
 * sigma-pressure levels (PrimitiveEquationsHybrid,
 * dinosaur/primitive_equations.py:1960-2362).  dino_modal_prep then writes
 * the input groups of dino_nodal_hybrid_explicit (the dry sigma groups plus
 * ln ps itself, whatever humidity_index is: the hybrid humidity terms enter
 * only through the virtual temperature, :2491-2518), and dino_modal_epilogue
 * reads its output groups, which are those of the dry sigma kernel. */
enum { DINO_MODAL_HYBRID = 1 };
int dino_modal_create_ex(int R, int L, int K, int n_tracers, int humidity_index,
                         double radius, double gravity, const double* inv_lap,
                         const double* lap, const double* clip_mask,
                         const double* up, const double* dn, const double* up2,
                         const double* dn2, int flags, dino_modal** out);

/* Fused nodal nonlinear terms of PrimitiveEquationsHybrid.explicit_terms.
 * Replaces compute_diagnostic_state_hybrid, the Simmons-Burridge
 * coefficients, get_geopotential_diff_hybrid, _get_pgf_lps_coefficient,
 * hybrid_vertical_advection and the nodal parts of the hybrid
 * curl_and_div_tendencies, nodal_temperature_vertical_tendency,
 * nodal_temperature_adiabatic_tendency and nodal_log_pressure_tendency
 * (dinosaur/primitive_equations.py:1483-1674, 1826-1957, 2032-2362).
 *
 * Input groups, each (members, levels): Z D T TRACER[0..n) U V GX GY LNPS
 * (levels K except GX, GY and the nodal ln ps: 1).  Output groups: CU CV KE
 * TEMP UT VT LSP (TRN TU TV)[0..n), as dino_nodal_explicit for a dry state;
 * KE holds the kinetic energy plus the geopotential of the full virtual
 * temperature at the column's surface pressure minus its linearisation
 * about the reference state (both take the same -laplacian).
 * a_boundaries, b_boundaries: K+1 interface coefficients, top first, with a
 * in the units of exp(ln ps) (nondimensional pressure); ps_ref: reference
 * surface pressure in the same units; tref: K reference temperatures. */
typedef struct dino_nodal_hybrid dino_nodal_hybrid;
int dino_nodal_hybrid_create(int K, int I, int J, int n_tracers,
                             int humidity_index, const int* cloud_indices,
                             int n_clouds, int include_vertical_advection,
                             const double* sec2, const double* coriolis,
                             const double* a_boundaries,
                             const double* b_boundaries, const double* tref,
                             double ps_ref, double r_gas, double kappa,
                             double r_vapor, double cp, double cp_vapor,
                             dino_nodal_hybrid** out);
int dino_nodal_hybrid_destroy(dino_nodal_hybrid* n);
int64_t dino_nodal_hybrid_fields(const dino_nodal_hybrid* n, int64_t members,
                                 int output);
int dino_nodal_hybrid_explicit(const dino_nodal_hybrid* n,
                               const void* nodal_in, void* nodal_out,
                               int64_t members, int dtype, void* stream);

/* Vertical regridding of nodal columns from hybrid levels (source: K_s
 * layers with interface coefficients a, b, p = a + b ps) to target levels.
 * Modes:
 *   DINO_VREGRID_CONSERVATIVE_SIGMA: conservative (overlap-weighted means)
 *     in sigma = a / ps + b; target_a = 0, target_b = sigma boundaries.
 *     Replaces vertical_interpolation.regrid_hybrid_to_sigma
 *     (dinosaur/vertical_interpolation.py:366-401).
 *   DINO_VREGRID_CONSERVATIVE_PRESSURE: conservative in pressure a + b ps
 *     (target: hybrid coefficients).  Replaces regrid_hybrid_to_hybrid
 *     (:404-440).
 *   DINO_VREGRID_LINEAR_SIGMA: linear interpolation between layer centres in
 *     sigma, one linearly extrapolated cell at each end, NaN beyond;
 *     target_a[0 .. K_t - 1] = the target sigma centres (as the sigma
 *     coordinates hold them: the reference's are float32 midpoints),
 *     target_b unused.  Replaces interp_hybrid_to_sigma (:303-326, :96-102).
 * Target cells with no overlap are zero.  a and ps in the same units.
 * Source boundaries must be non-decreasing down each column.
 * dino_vregrid_apply: fields (n_fields, K_s, columns) -> out (n_fields, K_t,
 * columns); surface_pressure (n_surface, columns), field f using slice
 * f / (n_fields / n_surface). */
enum {
  DINO_VREGRID_CONSERVATIVE_SIGMA = 0,
  DINO_VREGRID_CONSERVATIVE_PRESSURE = 1,
  DINO_VREGRID_LINEAR_SIGMA = 2,
};
typedef struct dino_vregrid dino_vregrid;
int dino_vregrid_create(int mode, int source_levels, const double* source_a,
                        const double* source_b, int target_levels,
                        const double* target_a, const double* target_b,
                        dino_vregrid** out);
int dino_vregrid_destroy(dino_vregrid* r);
int dino_vregrid_apply(const dino_vregrid* r, const void* fields,
                       const void* surface_pressure, void* out,
                       int64_t n_fields, int64_t n_surface, int64_t columns,
                       int dtype, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DINOSAUR_AMD_H_ */
