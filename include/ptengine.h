/*
 * ptengine.h -- C ABI of the MI355X process-tensor engine (libptengine.so).
 *
 * The engine replaces the tensor-network backends of OQuPy's process-tensor
 * path. Every entry point below names the reference interface it replaces
 * (paths relative to the OQuPy 0.5.0 source tree). Conventions:
 *
 *  - complex numbers are interleaved (re, im) doubles, as numpy complex128;
 *  - tensors are row-major, with their shape passed explicitly;
 *  - the caller owns every host buffer; the engine copies inputs in and writes
 *    outputs into caller buffers, and owns its device memory until *_destroy;
 *  - every function returns 0 on success and a negative PTE_ERR_* code on
 *    failure; pte_last_error() returns the message of the last failure on the
 *    calling thread;
 *  - calls on one handle must be serialised by the caller; the engine makes no
 *    callbacks.
 */
#ifndef PTENGINE_H
#define PTENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

#define PTE_OK 0
#define PTE_ERR_ARG (-1)       /* invalid argument (shape, index, value)     */
#define PTE_ERR_STATE (-2)     /* call out of order (e.g. step before init)  */
#define PTE_ERR_DEVICE (-3)    /* HIP runtime error                          */
#define PTE_ERR_CAPACITY (-4)  /* a bond would exceed the configured capacity */
#define PTE_ERR_NUMERIC (-5)   /* non-finite input or SVD not converged      */

#define PTE_PRECISION_C128 0  /* complex128 everywhere                          */
#define PTE_PRECISION_C64 1   /* network contractions in complex64 on the matrix
                                 cores with FP32 accumulation; QR, Jacobi SVD,
                                 singular values and the epsrel cut in FP64   */

typedef struct pte_pt_engine pte_pt_engine;
typedef struct pte_pt_batch pte_pt_batch;
typedef struct pte_tempo_engine pte_tempo_engine;

/* Configuration of a PT-TEMPO computation.
 * Replaces PtTempoBackend.__init__ (oqupy/backends/pt_tempo_backend.py:56-93). */
typedef struct {
  int dimension;    /* Hilbert space dimension d                              */
  int num_steps;    /* N, number of time steps of the process tensor (>= 2)   */
  int dkmax;        /* memory cut K (the backend's dkmax, never None)         */
  double epsrel;    /* relative SVD truncation error                          */
  int precision;    /* PTE_PRECISION_C128 or PTE_PRECISION_C64                */
  int chi_cap;      /* limit of every bond (INT_MAX: none, the default of the
                       Python side); a truncation above it is an error      */
  int n_north;      /* length of sum_north (d^2, or the degeneracy classes)   */
} pte_pt_config;

/* Create an engine; *out receives the handle. */
int pte_pt_create(const pte_pt_config* cfg, pte_pt_engine** out);

/* Influence tensor of distance dk (0 <= dk < min(N, K+1)), as returned by
 * influence_matrix (oqupy/tempo.py:969-1020): rows x cols complex values;
 * cols == 0 marks a vector of length rows (dk = 0 with degeneracy maps).
 * Replaces the influence callable given to PtTempoBackend. */
int pte_pt_set_influence(pte_pt_engine* e, int dk, const double* data, int rows,
                         int cols);

/* Degeneracy maps of Bath (oqupy/bath.py:87-89), each of length d^2; only for
 * unique=True (pt_tempo_backend.py:114-140). */
int pte_pt_set_degeneracy_maps(pte_pt_engine* e, const int* north_map,
                               const int* west_map, int length);

/* Build the influence MPO and the initial MPS and run the two truncating
 * sweeps. Replaces PtTempoBackend.initialize (pt_tempo_backend.py:105-190). */
int pte_pt_initialize(pte_pt_engine* e);

/* One PT-TEMPO step. `last_influence` is influence(-step) (rows x cols) when a
 * correlation time is added, else NULL. *more is set to step < num_steps.
 * Replaces PtTempoBackend.compute_step (pt_tempo_backend.py:192-282). */
int pte_pt_step(pte_pt_engine* e, const double* last_influence, int rows,
                int cols, int* more);

/* Current step (0 before initialisation). */
int pte_pt_get_step(pte_pt_engine* e, int* step);

/* Number of sites of the MPS and its bond dimensions (out has n_sites-1
 * entries). */
int pte_pt_num_sites(pte_pt_engine* e, int* n_sites);
int pte_pt_bond_dims(pte_pt_engine* e, int* out, int n);

/* PT-MPO tensor of time step `step` in the layout [bond_past, bond_future,
 * d^2], multiplied by d, into out (capacity in complex values `cap`); shape3
 * receives its shape. Replaces PtTempoBackend.get_mpo_tensor
 * (pt_tempo_backend.py:284-307). */
int pte_pt_get_mpo_tensor(pte_pt_engine* e, int step, double* out, long cap,
                          int* shape3);

/* MPS site i of the current step as stored, [left][physical][right] (no
 * scaling), at any step after initialisation; shape3 receives its shape.
 * Exposes the NodeArray node the reference keeps for that site
 * (oqupy/backends/node_array.py:45, pt_tempo_backend.py:166-169): it hands a
 * computation's state to a checker mid-run. */
int pte_pt_get_site(pte_pt_engine* e, int i, double* out, long cap, int* shape3);

/* Caps of the process tensor, computed on the device: caps[k] has length
 * bond_past of step k (caps[N] = [1]). Replaces SimpleProcessTensor.compute_caps
 * (oqupy/process_tensor.py:380-406). pte_pt_get_cap copies caps[k]. */
int pte_pt_compute_caps(pte_pt_engine* e);
int pte_pt_get_cap(pte_pt_engine* e, int k, double* out, int cap, int* length);

/* Statistics of the last step: number of SVDs, total Jacobi sweeps, largest
 * bond, accumulated device time of the step in milliseconds. */
int pte_pt_last_step_stats(pte_pt_engine* e, int* n_svd, int* sweeps,
                           int* max_bond, double* ms);

/* Accumulated statistics of the Jacobi SVD kernel since creation: launches,
 * device time (HIP events around every launch, milliseconds), algorithmic
 * flops (4 (6 M N^2 + 20 N^3) per M x N SVD, N = min) and algorithmic HBM
 * bytes (16 (2 M N + N^2)). No counterpart in the reference: it feeds the
 * roofline report of the benchmark. */
int pte_pt_svd_kernel_stats(pte_pt_engine* e, long* count, double* ms,
                            double* flops, double* bytes);

/* Log of the truncated SVDs that follow (no counterpart in the reference: it
 * exposes what split_node_full_svd keeps at oqupy/backends/node_array.py:262,
 * 285, 541): per SVD the shape of the matrix the reference factors (rows,
 * columns), the kept rank and the kept singular values, descending.
 * Recording stops once max_values singular values are held; max_values = 0
 * turns the log off. Each call clears the log. */
int pte_pt_svd_log(pte_pt_engine* e, long max_values);

/* Copy the log: shapes receives (m, n, k) per SVD for up to max_svds SVDs,
 * values the kept singular values of those SVDs concatenated, up to
 * max_values; *n_svds and *n_values receive the totals held (NULL buffers:
 * counts only). */
int pte_pt_get_svd_log(pte_pt_engine* e, int* shapes, int max_svds, double* values,
                       long max_values, int* n_svds, long* n_values);

void pte_pt_destroy(pte_pt_engine* e);

/* ---- batched PT-TEMPO (no counterpart in the reference) --------------------
 * Many independent PT-TEMPO engines (each created, given its influences and
 * initialised as above) stepped in lockstep: every truncated SVD of every
 * engine of one step is served by batched cooperative launches, so that
 * hundreds of instances with different baths, memory cuts and bond
 * dimensions share the GPU. Equivalent to calling pte_pt_step(e, NULL, 0, 0,
 * ...) on each engine (PtTempo.compute, oqupy/pt_tempo.py:258-278, for
 * baths without a per-step replacement influence); the batch does not own
 * the engines. */
int pte_pt_batch_create(pte_pt_engine* const* engines, int count, pte_pt_batch** out);

/* Up to max_steps steps of every engine that has steps left, each engine
 * advancing on its own (no lockstep at step boundaries: whatever SVDs the
 * engines have reached are served together); more[i] (count entries)
 * receives 1 while engine i has steps left. */
int pte_pt_batch_run(pte_pt_batch* b, int max_steps, int* more);

/* Served batches and requests (truncated SVDs), cooperative launches, host
 * milliseconds spent serving, and device milliseconds of the batched QR and
 * Jacobi launches since the batch was created. */
int pte_pt_batch_stats(pte_pt_batch* b, long* batches, long* requests, long* launches,
                       double* serve_ms, double* qr_ms, double* jacobi_ms);

void pte_pt_batch_destroy(pte_pt_batch* b);

/* ---- TEMPO (oqupy/backends/tempo_backend.py:314-626) ---------------------- */

typedef struct {
  int dimension;    /* Hilbert space dimension d                              */
  int dkmax;        /* memory cut K, or -1 for no cut (dkmax None)            */
  double epsrel;
  int precision;
  int chi_cap;
  int n_north;      /* length of sum_north                                    */
  int n_west;       /* length of sum_west                                     */
} pte_tempo_config;

/* Replaces BaseTempoBackend.__init__ (tempo_backend.py:343-372). */
int pte_tempo_create(const pte_tempo_config* cfg, pte_tempo_engine** out);

/* Influence of distance dk >= 0 (see pte_pt_set_influence). For dkmax = -1
 * the caller sets dk = len(MPO) before each step, as the reference grows the
 * MPO (tempo_backend.py:487-497). */
int pte_tempo_set_influence(pte_tempo_engine* e, int dk, const double* data,
                            int rows, int cols);

/* Degeneracy maps for unique=True (tempo_backend.py:400-417). */
int pte_tempo_set_degeneracy_maps(pte_tempo_engine* e, const int* north_map,
                                  const int* west_map, int length);

/* Initial state vector (d^2) and the unitary of the coupling's eigenbasis
 * (d x d). Replaces BaseTempoBackend.initialize_mps_mpo
 * (tempo_backend.py:379-437). */
int pte_tempo_initialize(pte_tempo_engine* e, const double* initial_state,
                         const double* unitary);

/* One TEMPO step with the half-step propagators prop_1, prop_2 (d^2 x d^2) and,
 * after the memory is full, influence(dkmax - step) (rows x cols) or NULL.
 * Writes the new state vector (d^2) into state_out. Replaces
 * BaseTempoBackend.compute_system_step (tempo_backend.py:439-575). */
int pte_tempo_step(pte_tempo_engine* e, const double* prop_1,
                   const double* prop_2, const double* extra_influence,
                   int rows, int cols, double* state_out);

int pte_tempo_get_step(pte_tempo_engine* e, int* step);
int pte_tempo_bond_dims(pte_tempo_engine* e, int* out, int n, int* n_sites);
void pte_tempo_destroy(pte_tempo_engine* e);

/* ---- compute_dynamics with one process tensor ------------------------------
 * Replaces the step loop of compute_dynamics (oqupy/system_dynamics.py:41-182,
 * _apply_system_superoperator 631-640, _apply_caps 643-651, _apply_pt_mpos
 * 654-700) for a single environment without controls.
 *   d2          Liouville dimension
 *   num_steps   N
 *   bonds       N+1 bond dimensions (bonds[0] = bonds[N] = 1)
 *   tensors     concatenated rank-3 PT tensors [bonds[k], bonds[k+1], d2]
 *   caps        concatenated caps (lengths bonds[k], k = 0..N)
 *   props       N pairs (P1, P2) of d2 x d2 propagators
 *   transform_in/out  d2 x d2 basis transforms or NULL (process_tensor.py:349-355)
 *   rho0        initial state vector (d2)
 *   states_out  (N+1) x d2 state vectors
 */
int pte_dynamics_run(int d2, int num_steps, const int* bonds,
                     const double* tensors, const double* caps,
                     const double* props, const double* transform_in,
                     const double* transform_out, const double* rho0,
                     double* states_out);

/* Same, reading the PT-MPO tensors and caps held by a finished PT-TEMPO
 * engine (after pte_pt_compute_caps), so the process tensor never leaves the
 * device. */
int pte_dynamics_run_engine(pte_pt_engine* e, const double* props,
                            const double* transform_in,
                            const double* transform_out, const double* rho0,
                            double* states_out);

/* compute_dynamics of nsys systems through the same device-resident process
 * tensor (the same state vector sizes and transforms as
 * pte_dynamics_run_engine): props holds nsys x num_steps x 2 x d^2 x d^2,
 * rho0 nsys x d^2, states_out receives nsys x (num_steps + 1) x d^2. The
 * process tensor is read by all systems in the same kernels. No counterpart
 * in the reference; each system's result equals compute_dynamics
 * (oqupy/system_dynamics.py:41-182) on its own. */
int pte_dynamics_run_engine_batch(pte_pt_engine* e, int nsys, const double* props,
                                  const double* transform_in,
                                  const double* transform_out, const double* rho0,
                                  double* states_out);

/* compute_dynamics of one system coupled to `count` (1..8) environments,
 * each given by a completed, device-resident process tensor of the same
 * length and dimension; transforms_in / transforms_out hold one pointer per
 * environment (NULL entries for none). Replaces the loop of compute_dynamics
 * over a list of process tensors (oqupy/system_dynamics.py:119-171, 654-700). */
int pte_dynamics_run_engines(pte_pt_engine* const* engines, int count, const double* props,
                             const double* const* transforms_in,
                             const double* const* transforms_out, const double* rho0,
                             double* states_out);

/* The same for process tensors held on the host (read from files or built
 * outside the engine): bonds is count x (num_steps + 1); tensors and caps are
 * the environments' [a][b][i] sites and caps concatenated in order, with the
 * layout of pte_dynamics_run (oqupy/system_dynamics.py:119-171). */
int pte_dynamics_run_multi(int d2, int num_steps, int count, const int* bonds,
                           const double* tensors, const double* caps, const double* props,
                           const double* const* transforms_in,
                           const double* const* transforms_out, const double* rho0,
                           double* states_out);

/* Truncated SVD of one matrix by the engine's kernels: M (m x n, row-major)
 * = U diag(S) Vh with the kept rank k = #{j : sqrt(sum_{i>=j} s_i^2) >
 * epsrel * s_0}; U is m x min(m,n) capacity (row-major m x k on return),
 * S has min(m,n) capacity, SVh receives diag(S) Vh (row-major k x n).
 * Replaces tensornetwork's split_node_full_svd(max_truncation_err=epsrel,
 * relative=True) as called at oqupy/backends/node_array.py:262, 285, 541. */
int pte_split_svd(int m, int n, const double* M, double epsrel, int chi_cap,
                  double* U, double* S, double* SVh, int* k_out,
                  int* sweeps_out);

/* The truncated SVD of pte_split_svd without allocating: no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, rule and results as pte_split_svd, for
 * matrices of at most 256 x 256 (larger ones: PTE_ERR_CAPACITY). Its
 * workspace is a 12 MiB file-scope device array (one per GPU, held until the
 * process ends) that every call on a GPU reuses: calls are serialised by a
 * lock inside the library, and a call that overlapped another on the same
 * GPU's array would race. It runs on the null stream and calls no hipMalloc,
 * hipFree or other allocating routine. */
int pte_split_svd_noalloc(int m, int n, const double* M, double epsrel, int chi_cap,
                          double* U, double* S, double* SVh, int* k_out,
                          int* sweeps_out);

/* One network contraction C = A B (A: m x k, B: k x n, C: m x n, row-major
 * complex128) by the kernels of the given precision (PTE_PRECISION_C64
 * rounds the operands to complex64 and accumulates in FP32 on the matrix
 * cores). Replaces one pairwise contraction of tensornetwork's greedy
 * contractor as called at oqupy/backends/node_array.py:519 (zip-up) and the
 * `@` absorption at node_array.py:272, 295. */
int pte_contract(int precision, int m, int n, int k, const double* A, const double* B,
                 double* C);

/* ---- learned truncation-rank form (no counterpart in the reference: the
 * reference always runs the full SVD of split_node_full_svd,
 * oqupy/backends/node_array.py:262, 285, 541) ----
 * The pivoted QR that starts every truncated SVD of the zip-up and the sweep
 * is capped. mode 0: no cap (exact; the default). mode 1: the bond's
 * previous step count + oversample. mode 2: the step count the MLP `params`
 * (pte_rank_param_count() floats) predicts, times (1 + margin), +
 * oversample. mode 3: stop at the acceptance threshold itself. A capped
 * factorisation is accepted when its residual is at most
 * theta * epsrel * |beta_0| (the largest column norm); the residual then
 * counts as discarded weight in the cut. Otherwise the pivoted QR resumes and
 * runs to the exact stop. alpha and temperature are features of the bath. */
int pte_pt_set_rank_mode(pte_pt_engine* e, int mode, const float* params, double alpha,
                         double temperature, int oversample, double margin, double theta);
/* Record up to `capacity` training records (features, target log2(need) / 8,
 * need, kept rank, min(m, n): PTE_RANK_RECORD floats each) of the truncations
 * that follow; need = the pivoted-QR steps that reached the threshold. */
int pte_pt_rank_trace(pte_pt_engine* e, int capacity);
/* Copy up to max_records records into out (NULL: count only). */
int pte_pt_get_rank_trace(pte_pt_engine* e, float* out, int max_records, int* count);
/* Truncations seen, capped, fallen back to the exact SVD; pivoted-QR steps. */
int pte_pt_rank_stats(pte_pt_engine* e, long* svds, long* capped, long* fallbacks,
                      long* qr_steps);

#define PTE_RANK_RECORD 28
/* The rank predictor's network (24 features -> 3 x 128 tanh -> 1, FP32). */
int pte_rank_param_count(void);
int pte_rank_feature_count(void);
/* yhat[i] = network(X[i]) for n feature rows. */
int pte_rank_mlp_predict(const float* params, const float* X, int n, float* yhat);
/* loss = mean (network(X[i]) - y[i])^2 and its gradient w.r.t. params. */
int pte_rank_mlp_grad(const float* params, const float* X, const float* y, int n, float* grad,
                      float* loss);
/* Adam on mini-batches of a seeded shuffle; params updated in place,
 * epoch_loss (optional) receives the mean loss of every epoch. */
int pte_rank_mlp_train(float* params, const float* X, const float* y, int n, int epochs,
                       int batch, float lr, unsigned seed, float* epoch_loss);

/* eta(tau) of power-law baths for many times in one launch: bath b has
 * bath_params[4b..4b+3] = alpha, zeta (>= 1), w_c, T and cutoff_types[b]
 * (0 exponential, 1 Gaussian, 2 hard); time i belongs to bath tau_bath[i];
 * eta_out receives n_tau complex values. gl_nodes / gl_weights: the 16-point
 * Gauss-Legendre rule on [-1, 1]. Replaces the QUADPACK integrals of
 * CustomSD.eta_function (oqupy/bath_correlations.py:518-591) that every
 * influence tensor (oqupy/tempo.py:969-1020) is built from. */
int pte_eta_batch(int n_baths, const double* bath_params, const int* cutoff_types,
                  const double* gl_nodes, const double* gl_weights, int n_tau,
                  const int* tau_bath, const double* taus, double* eta_out);

const char* pte_last_error(void);

/* Library version string. */
const char* pte_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PTENGINE_H */
