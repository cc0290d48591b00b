/*
 * fluidity_amd.h -- C ABI of the MI355X (gfx950) finite-volume Euler step.
 *
 * Drop-in boundary for fluidity's split-solver update call.  Each entry point
 * names the reference interface it replaces (paths relative to the reference's
 * include/fluidity/).  All calls:
 *   - return an fl_status (0 = FL_OK); fl_last_error() describes a failure;
 *   - take device pointers owned by the caller and a hipStream_t passed as
 *     void* (NULL = the null stream) and are asynchronous on that stream
 *     unless stated otherwise;
 *   - allocate nothing except the context's own small control block.
 * A context belongs to one device; calls on one context from several host
 * threads must be serialised by the caller.
 */
#ifndef FLUIDITY_AMD_H
#define FLUIDITY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fl_status {
  FL_OK = 0,
  FL_ERR_ARGUMENT = 1,     /* null pointer, bad size, bad enum, dtdh <= 0 ...  */
  FL_ERR_UNSUPPORTED = 2,  /* a valid combination this library does not run    */
  FL_ERR_HIP = 3,          /* the HIP runtime reported an error                 */
  FL_ERR_NUMERIC = 4       /* NaN or non-positive maximum wavespeed (vacuum)    */
} fl_status;

typedef enum fl_dtype { FL_F64 = 0, FL_F32 = 1 } fl_dtype;

/* AoS is the reference's State<T,conservative,D> array (state/state.hpp:78-98,
 * x fastest: dimension/dimension_info.hpp:72-81); SoA keeps one plane per
 * component, the native layout of the gfx950 kernels. */
typedef enum fl_layout { FL_LAYOUT_AOS = 0, FL_LAYOUT_SOA = 1 } fl_layout;

typedef enum fl_limiter { FL_LIMITER_VAN_LEER = 0, FL_LIMITER_SUPERBEE = 1 } fl_limiter;
typedef enum fl_limit_form { FL_FORM_CONSERVATIVE = 0, FL_FORM_PRIMITIVE = 1 } fl_limit_form;
typedef enum fl_flux {
  FL_FLUX_HLLC = 0,            /* flux_method/flux_hllc.hpp */
  FL_FLUX_FORCE = 1,           /* flux_method/flux_force.hpp */
  FL_FLUX_LAX_FRIEDRICHS = 2,  /* flux_method/flux_lax_friedrichs.hpp (single grids) */
  FL_FLUX_RICHTMYER = 3        /* flux_method/flux_richtmyer.hpp (single grids) */
} fl_flux;
/* FL_EOS_STIFF is the reference's StiffFluid as it computes: its energy
 * e = (p + gamma p_inf)/((gamma - 1) rho) and sound speed
 * sqrt(gamma (p + p_inf)/rho) (material/stiff_fluid.hpp:77-98) with the
 * ideal-gas pressure (gamma - 1)(E - rho |v|^2/2), which ignores p_inf
 * (state/state_impl.hpp:117-128): the bit-parity default.
 * FL_EOS_STIFF_CORRECTED takes the pressure consistent with that energy,
 * p = (gamma - 1)(E - rho |v|^2/2) - gamma p_inf, and HLLC wave-speed ratios
 * (p* + p_inf)/(p_K + p_inf): the stiffened gas of physical runs (an
 * extension). */
typedef enum fl_eos { FL_EOS_IDEAL = 0, FL_EOS_STIFF = 1, FL_EOS_STIFF_CORRECTED = 2 } fl_eos;
/* Arithmetic of the FP64 sweeps.  EXACT keeps the reference's operation
 * order with correctly rounded division and square root: bit-identical to
 * the reference's results.  FAST forms each state's 1/rho once, takes
 * quotients as products with refined reciprocals, uses the closed form of
 * the van Leer slope (minmod) and contracts multiply-adds; it is held to
 * 1e-12 relative L-inf of the reference over a 100-step run.  FP32 sweeps
 * use the fast forms whatever this says.  The single-grid entry points
 * (fl_split_sweep/solve/step, fl_simulate, fl_ctx_autotune) honour it, and so
 * do batches (every FP64 problem of a batch must ask for the same numerics,
 * else FL_ERR_UNSUPPORTED) and the two-material loop (every material the
 * same, else FL_ERR_ARGUMENT; FAST there also takes the level sets' WENO
 * right-hand side and, with two materials, the wavespeed maximum that sets
 * dt in their fast forms); the adjoint entry points run FP64 exact
 * and reject FAST with FL_ERR_UNSUPPORTED.  An extension: the reference has
 * one arithmetic. */
typedef enum fl_numerics { FL_NUMERICS_EXACT = 0, FL_NUMERICS_FAST = 1 } fl_numerics;
typedef enum fl_boundary {
  FL_BC_TRANSMISSIVE = 0,  /* solver/boundary_loader.hpp:29-32 */
  FL_BC_REFLECTIVE = 1,
  FL_BC_PERIODIC = 2       /* extension: not in the reference */
} fl_boundary;

/* A grid of conservative states [rho, E, m_x, (m_y)] holding interior cells
 * only (simulator/simulator_impl.hpp:272); ghost cells never live in memory. */
typedef struct fl_field {
  void*   data;      /* device pointer to component 0 of cell 0 */
  int32_t dtype;     /* fl_dtype  */
  int32_t layout;    /* fl_layout */
  int32_t ndim;      /* 1 or 2    */
  int32_t reserved;
  int64_t n[3];      /* interior cells per dimension, x first; unused = 1 */
} fl_field;

/* The scheme a reference SplitSolver<FaceFlux<MHReconstructor<Limiter<Form>>,
 * Flux, Eos>, BoundaryLoader<4>, Num<D>> instantiation fixes at compile time
 * (solver/split_solver.hpp:37-126, solver/flux_solver.hpp:27-41). */
typedef struct fl_scheme {
  int32_t limiter;   /* fl_limiter    */
  int32_t form;      /* fl_limit_form */
  int32_t flux;      /* fl_flux       */
  int32_t eos;       /* fl_eos        */
  double  gamma;     /* adiabatic index (material/ideal_gas.hpp:118) */
  double  p_inf;     /* stiffened-gas constant (material/stiff_fluid.hpp:132) */
  int32_t bc[3];     /* fl_boundary per dimension, both ends (BoundarySetter) */
  int32_t numerics;  /* fl_numerics (0 = exact) */
} fl_scheme;

typedef struct fl_sim_params {
  double  cfl;        /* Parameters::cfl (simulator/parameters.hpp) */
  double  dh;         /* Domain::resolution (simulator/domain.hpp)   */
  double  sim_time;   /* Parameters::sim_time                         */
  int64_t max_iters;  /* Parameters::max_iters                        */
} fl_sim_params;

typedef struct fl_sim_result {
  double  run_time;   /* Parameters::run_time after the loop */
  int64_t iters;      /* Parameters::iters after the loop    */
  double  last_dt;    /* dt of the last step                 */
  double  max_wavespeed; /* max wavespeed of the final state */
} fl_sim_result;

/* Launch geometry; zeros select the built-in defaults. */
typedef struct fl_tuning {
  int32_t x_block;    /* lanes per x-sweep workgroup (64..256, multiple of 64) */
  int32_t x_rows;     /* rows marched per x-sweep workgroup                   */
  int32_t y_block;    /* lanes per y-sweep workgroup (64..256, multiple of 64) */
  int32_t y_rows;     /* rows marched per y-sweep workgroup                   */
} fl_tuning;

typedef struct fl_context fl_context;

const char* fl_status_string(int status);
/* Message of the last failure on the calling thread ("" if none). */
const char* fl_last_error(void);
/* Version string of this library. */
const char* fl_version(void);

/* Creates a context on HIP device `device`; replaces the state the reference
 * keeps in SimulatorImpl (simulator/simulator_impl.hpp:43-211). */
int fl_ctx_create(int device, fl_context** ctx);
int fl_ctx_destroy(fl_context* ctx);
int fl_ctx_set_tuning(fl_context* ctx, const fl_tuning* tuning);
int fl_ctx_get_tuning(const fl_context* ctx, fl_tuning* tuning);
/* Launch-shape autotuner: times every candidate (block width 64/128/256 x
 * rows per workgroup: 8..128 and the heights that make a launch fill whole
 * rounds of the workgroups the device holds at once) of each sweep on this
 * grid (in -> work, a fixed small dt/dh; `in` is not modified), keeps the
 * fastest in the context and returns it.  Synchronises.  Replaces the fixed
 * thread_sizes()/block_sizes() (solver/split_solver.hpp:117-125). */
int fl_ctx_autotune(fl_context* ctx, const fl_field* in, const fl_field* work,
                    const fl_scheme* scheme, int reps, fl_tuning* chosen, void* stream);

/* One sweep in dimension `dim`: out = U + dtdh (F_{i-1/2} - F_{i+1/2}).
 * Replaces solve_dim<S> -> SplitSolver::invoke
 * (solver/cuda/split_solver.cuh:39-44, solver/split_solver.hpp:138-162). */
int fl_split_sweep(fl_context* ctx, const fl_field* in, const fl_field* out, int dim,
                   const fl_scheme* scheme, double dtdh, void* stream);

/* All sweeps of one step, x then y.  Mirrors SplitSolver::solve
 * (solver/split_solver.hpp:105-115, solver/cuda/split_solver.cuh:57-80),
 * including its swap of the caller's in/out between sweeps: on return
 * out->data holds the result (for 2D, in->data and out->data are exchanged). */
int fl_split_solve(fl_context* ctx, fl_field* in, fl_field* out, const fl_scheme* scheme,
                   double dtdh, void* stream);

/* fl_split_solve with the CFL reduction fused into the last sweep: the
 * proposed fused split step of the update call (solver/split_solver.hpp:
 * 105-115 followed by set_wavespeeds + max_element on its result,
 * simulator/simulator_impl.hpp:241-246).  max_wavespeed (host, nullable)
 * receives the FP64 maximum wavespeed of the new states, which synchronises
 * `stream`; NULL skips the reduction and the synchronisation. */
int fl_split_step(fl_context* ctx, fl_field* in, fl_field* out, const fl_scheme* scheme,
                  double dtdh, double* max_wavespeed, void* stream);

/* max over cells of max_d |u_d| + a, computed in FP64 and returned on the
 * host (synchronises `stream`).  Replaces set_wavespeeds + max_element
 * (simulator/cuda/wavespeed_initialization.cuh:41-97,
 * algorithm/max_element.hpp:28-38).  A NaN result is returned as NaN. */
int fl_max_wavespeed(fl_context* ctx, const fl_field* in, const fl_scheme* scheme,
                     double* max_wavespeed, void* stream);

/* Per-cell wavespeeds max_d |u_d| + a of `in`, in FP64, into `wavespeeds`
 * (device, one double per cell, x fastest).  Replaces set_wavespeeds
 * (simulator/wavespeed_initialization.hpp:34-53,
 * simulator/cuda/wavespeed_initialization.cuh:41-97) as SimulatorImpl::simulate
 * calls it (simulator/simulator_impl.hpp:241). */
int fl_set_wavespeeds(fl_context* ctx, const fl_field* in, const fl_scheme* scheme,
                      double* wavespeeds, void* stream);

/* The maximum of `count` doubles at `first` (device), returned on the host
 * (synchronises `stream`).  A NaN among them gives NaN; count must be > 0.
 * Replaces max_element over a device tensor's range (algorithm/max_element.hpp:
 * 28-38, algorithm/cuda/reduce.cuh) as SimulatorImpl::simulate calls it
 * (simulator/simulator_impl.hpp:242-243). */
int fl_max_element(fl_context* ctx, const double* first, int64_t count, double* result,
                   void* stream);

/* The simulator loop of SimulatorImpl::simulate
 * (simulator/simulator_impl.hpp:222-258): while run_time < sim_time and
 * iters < max_iters, dt = cfl*dh/max_ws, all sweeps, run_time += dt.  dt is
 * computed on the device from the max wavespeed fused into the last sweep of
 * the previous step; the loop runs max_iters steps' worth of launches and
 * steps past the stop condition do nothing.  `state` holds the initial data
 * and receives the result; `work` is a same-shaped buffer.  Synchronises
 * `stream` once at the end.  result may be NULL. */
int fl_simulate(fl_context* ctx, fl_field* state, fl_field* work, const fl_scheme* scheme,
                const fl_sim_params* params, fl_sim_result* result, void* stream);

/* fl_simulate without a context and without allocating: for callers that
 * cannot allocate device or pinned memory (serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The loop, its
 * arguments and results are fl_simulate's; tuning (nullable: the built-in
 * launch shapes) gives the launch shapes, and steps are launched one by one
 * (no graphs).  The step-control block the kernels share lives in a
 * file-scope device array of 256 bytes, one per GPU, held until the process
 * ends; the loop needs one such block whatever the grid size, so no grid is
 * too large for it.  The host reads the block into a stack copy once per 256
 * steps.  Every call on one GPU uses that GPU's block: calls that overlap on
 * one GPU (from several host threads or on several streams) race and give
 * undefined results.  Synchronises `stream`. */
int fl_simulate_noalloc(const fl_tuning* tuning, fl_field* state, fl_field* work,
                        const fl_scheme* scheme, const fl_sim_params* params,
                        fl_sim_result* result, void* stream);

/* fl_simulate captures a pair of steps (control kernel + sweeps, both
 * parities) in a hipGraph on first use for a given grid, scheme, parameters
 * and tuning, and replays it (default); a 1D line of at most 2^16 cells
 * (exact FP64 or FP32) instead runs the whole loop in one cooperative launch
 * with a grid barrier per step.  enable = 0 launches every kernel. */
int fl_ctx_set_graphs(fl_context* ctx, int enable);

/* FP64 sums of each conserved component over the grid (synchronises
 * `stream`); sums has 2+ndim entries.  Conservation diagnostics beside the
 * FP32 bulk path. */
int fl_conservation_sums(fl_context* ctx, const fl_field* in, double* sums, void* stream);

/* ---- initial conditions and result output -----------------------------------
 * The device half of SimulatorImpl::fill_data / MultimaterialSimulator's
 * set_state_data, and the reference's result formats. */

/* Primitive planes of a grid, for output: `out` (device, FP64) receives 3+ndim
 * planes of cell_count doubles, x fastest: rho, p, v_x, (v_y), and the
 * internal energy e = eos(primitive state).  Replaces the per-cell
 * state.primitive(material) and material.eos(state) of stream_output_ascii
 * (simulator/simulator_impl.hpp:478-490, state/state_impl.hpp:339-357,
 * material/ideal_gas.hpp:60-63, material/stiff_fluid.hpp:77-80). */
int fl_primitive_planes(fl_context* ctx, const fl_field* state, const fl_scheme* scheme,
                        double* out, void* stream);

/* Conservative grid from primitive planes (device, FP64: rho, p, v_x, (v_y),
 * x fastest): fill_state.conservative(material) of SimulatorImpl::fill_data
 * (simulator/simulator_impl.hpp:329-331, state/state_impl.hpp:135-142,
 * 388-406).  The host evaluates the fillers. */
int fl_from_primitive(fl_context* ctx, const double* prim, const fl_scheme* scheme,
                      fl_field* state, void* stream);

/* Cells inside the level set `phi` (device, one double per cell, phi <= 0)
 * take the conservative form of the primitive state `prim_state` (host,
 * 2+ndim doubles: rho, p, v...); others are untouched.  Replaces set_states
 * (state/cuda/state_initialization.cuh:72-118) as called by
 * MultimaterialSimulationData::set_state_data
 * (simulator/multimaterial_simulation_data.hpp:166-189). */
int fl_set_states(fl_context* ctx, const double* phi, const double* prim_state,
                  const fl_scheme* scheme, fl_field* state, void* stream);

/* The ASCII result file of SimulatorImpl::write_results / print_results
 * (simulator/simulator_impl.hpp:343-357, 435-508): a "# t = <run_time>"
 * header, column names, then one line per cell (x fastest) of coordinates
 * ((i + 0.5) * resolution), rho, p, v..., e, each as "%-12.8f ".  path NULL
 * writes to stdout.  With `phi` (device, nullable) the file is the
 * two-material one of MultimaterialSimulator::stream_output_ascii
 * (simulator/multimaterial_simulator.hpp:882-971): one level-set column per
 * dimension after the coordinates.  Synchronises `stream`. */
int fl_write_results(fl_context* ctx, const fl_field* state, const fl_scheme* scheme,
                     const double* phi, double run_time, double resolution, const char* path,
                     void* stream);

/* SimulatorImpl::write_results_separate_raw (simulator/simulator_impl.hpp:
 * 360-420): one file per primitive component, named
 * <prefix>_<rho|p|v_x|v_y>_<sim_time as %f>.dat, holding the values as "%g "
 * one grid row per line (1D: one more newline).  Synchronises `stream`. */
int fl_write_results_separate_raw(fl_context* ctx, const fl_field* state,
                                  const fl_scheme* scheme, double sim_time, const char* prefix,
                                  void* stream);

/* ---- batched problems -------------------------------------------------------
 * Many independent grids of uneven size (1D and 2D, each with its own
 * material, CFL dt and stop condition) advanced together: one step is a
 * control kernel plus one persistent launch per sweep phase whose workgroups
 * take tiles from a shared counter; steps are captured once in a hipGraph and
 * replayed.  Each problem's result is bit-identical to fl_simulate on that
 * problem alone.  The batch as a whole replaces a host loop of
 * SimulatorImpl::simulate calls (simulator/simulator_impl.hpp:222-258), one
 * per problem, with Parameters::update_time_delta / dt_dh per problem
 * (simulator/parameters.hpp:60-99). */
typedef struct fl_batch fl_batch;

typedef struct fl_batch_problem {
  fl_field      state;   /* initial data; the result after fl_batch_finish     */
  fl_field      work;    /* same-shaped scratch buffer                         */
  fl_scheme     scheme;  /* material, limiter, BCs; one flux for the batch     */
  fl_sim_params params;  /* cfl, dh, sim_time, max_iters of this problem       */
} fl_batch_problem;

/* Builds the device tables of a batch (all problems share one dtype; the
 * batch owns its tables, allocated here) and computes each problem's initial
 * wavespeed on `stream`.  The problems' buffers stay owned by the caller. */
int fl_batch_create(fl_context* ctx, const fl_batch_problem* problems, int nprob,
                    fl_batch** batch, void* stream);
int fl_batch_destroy(fl_batch* batch);
/* Tile shapes of the two sweep phases (x_block/x_rows: x tiles of 2D problems
 * and 1D tiles; y_block/y_rows: y tiles).  Zeros keep the current values. */
int fl_batch_set_tuning(fl_batch* batch, const fl_tuning* tuning);
int fl_batch_get_tuning(const fl_batch* batch, fl_tuning* tuning);
/* Times every candidate tile shape of each phase on this batch (`reps` timed
 * launches each, states untouched) and keeps the fastest.  Synchronises. */
int fl_batch_autotune(fl_batch* batch, int reps, fl_tuning* chosen, void* stream);
/* Restarts every problem's clock and iteration count from its current state. */
int fl_batch_reset(fl_batch* batch, void* stream);
/* Enqueues `steps` steps (problems past their stop condition do nothing);
 * use_graph replays a captured hipGraph of one step.  Asynchronous. */
int fl_batch_step(fl_batch* batch, int64_t steps, int use_graph, void* stream);
/* Puts every result in its state buffer and reports each problem's run time,
 * iterations, last dt and final max wavespeed (results: nprob entries, may be
 * NULL).  Synchronises.  A problem whose maximum wavespeed turns NaN takes
 * that step with a NaN dt and then stops, as the reference loop does, while
 * the others go on; FL_ERR_NUMERIC then names the first such problem (the
 * results are filled all the same). */
int fl_batch_finish(fl_batch* batch, fl_sim_result* results, void* stream);
/* FP64 sums of each conserved component per problem (4 slots per problem,
 * the 4th 0 for 1D) of the state buffers; call after fl_batch_finish. */
int fl_batch_conservation_sums(fl_batch* batch, double* sums, void* stream);

/* ---- adjoint of the split step ------------------------------------------------
 * Hand-written reverse mode of the FP64 split MUSCL-Hancock step with the
 * HLLC or FORCE flux (either limiter, either limiting form, any boundary),
 * including the dependence of dt = cfl dh / max wavespeed and of the
 * flattening factor chi on the state and on gamma.  Branches follow the
 * primal (the derivative of the branch taken).  The reference has no adjoint; these
 * entry points differentiate its update call
 * (solver/split_solver.hpp:105-162 inside simulator/simulator_impl.hpp:222-258).
 * Adjoint fields use the state's layout. */

/* Bytes of the checkpoint buffer of a run of `steps` steps with a checkpoint
 * every `checkpoint_every` steps, and of the scratch of one segment. */
int64_t fl_adj_checkpoint_bytes(const fl_field* state, int64_t steps, int checkpoint_every);
int64_t fl_adj_work_bytes(const fl_field* state, int checkpoint_every);

/* Runs `steps` steps from `state` (the result is left in `state`) and stores
 * the state of every checkpoint_every-th step in `checkpoints`.  Synchronises
 * once per segment; result->run_time is the accumulated simulated time. */
int fl_adj_forward(fl_context* ctx, fl_field* state, const fl_scheme* scheme, double cfl,
                   double dh, int64_t steps, int checkpoint_every, void* checkpoints,
                   int64_t ckpt_bytes, void* work, int64_t work_bytes, fl_sim_result* result,
                   void* stream);

/* Given dLoss/d(final state) in `grad`, replaces it with dLoss/d(initial
 * state) and returns dLoss/dgamma, recomputing each segment's steps from its
 * checkpoint (memory: one segment, not the whole run).  Synchronises. */
int fl_adj_backward(fl_context* ctx, fl_field* grad, const fl_scheme* scheme, double cfl,
                    double dh, int64_t steps, int checkpoint_every, const void* checkpoints,
                    int64_t ckpt_bytes, void* work, int64_t work_bytes, double* dgamma,
                    void* stream);

/* Vector-Jacobian product of one step U_{n+1} = step(U_n):
 * lambda = (dU_{n+1}/dU_n)^T lambda_next, dgamma = lambda_next . dU_{n+1}/dgamma.
 * work: fl_adj_work_bytes(state, 1).  Synchronises. */
int fl_split_step_vjp(fl_context* ctx, const fl_field* state, const fl_field* lambda_next,
                      fl_field* lambda, const fl_scheme* scheme, double cfl, double dh,
                      double* dgamma, void* work, int64_t work_bytes, void* stream);

/* ---- level set and Riemann ghost fluid -------------------------------------
 * Level-set fields phi are FP64 device arrays of n[0] (1D) or n[0]*n[1] (2D)
 * values, x fastest, interior cells only; phi <= 0 inside the material.  State
 * fields given with them are FP64 or FP32 (AoS or SoA) of the same shape; the
 * level-set and ghost-fluid arithmetic runs in FP64 either way (FP32 states
 * are widened on load and rounded on store).  `work`
 * arguments are caller-owned device scratch of the stated size, 256-byte
 * aligned, not aliasing any other argument. */

/* Interface masks, one byte per cell: 1 inside (phi <= 0), 2 outside
 * (phi > 0), 4 on the boundary (phi == 0), 8 inside-interfacial
 * (inside && phi >= -dh), 16 outside-interfacial (outside && phi < |dh|).
 * Replaces the predicates of levelset/levelset.hpp:166-217. */
int fl_ls_masks(const double* phi, int ndim, const int64_t* n, double dh, uint8_t* masks,
                void* stream);

/* Level-set speed V = u . n, u the velocity of the first material whose
 * phi <= 0 (V = 0 where none is), n the normalised central-difference
 * gradient: of that material's phi when target < 0 (the reference's field),
 * of phis[target] otherwise (the field level set `target` is advected with).
 * phis and states are host arrays of nmat (1 or 2) entries.  Replaces
 * set_velocities (levelset/cuda/velocity_initialization.cuh:108-143). */
int fl_ls_velocity(int nmat, const double* const* phis, const fl_field* states, double dh,
                   int target, double* V, void* stream);

/* One SSP-RK3 step of phi_t + V |grad phi| = 0 with WENO-HJ5 one-sided
 * derivatives and Godunov upwinding; ghost cells mirror the interior.
 * work: 2 * cells doubles.  Replaces scheme::evolve with the level-set
 * evolver (scheme/cuda/evolve.cuh:51-157, scheme/updaters/runge_kutta_3.hpp,
 * scheme/stencils/weno_hj_5.hpp, scheme/schemes/godunov_upwind.hpp,
 * scheme/evaluators/levelset_hamiltonian.hpp). */
int fl_ls_advect(const double* phi, const double* V, int ndim, const int64_t* n, double dt,
                 double dh, double* out, double* work, void* stream);

/* Two-material projection: avg = (phi0 + phi1) / 2; phi0 -= avg;
 * phi1 = -phi0.  Replaces levelset::project (levelset/cuda/levelset_projection.cuh:54-66). */
int fl_ls_project(double* phi0, double* phi1, int64_t cells, void* stream);

/* Reinitialises phi to signed distance within `band` cells of the
 * interface: sources beside each sign change keep |phi|, every other cell
 * solves the Godunov eikonal update by Jacobi sweeps to the fixed point;
 * cells not reached get +-(band + 1) dh.  band_mask (may be NULL) marks the
 * cells reached; *iterations (may be NULL) receives the sweep count.
 * work: fl_ls_reinit_work_bytes() bytes.  Synchronises `stream` once per few
 * sweeps.  Replaces solver::eikonal with the fast iterative method
 * (scheme/cuda/fast_iterative_method.cuh:48-461,
 * scheme/schemes/godunov_upwind_scheme.hpp:37-79). */
int64_t fl_ls_reinit_work_bytes(int ndim, const int64_t* n);
int fl_ls_reinit(fl_context* ctx, const double* phi, int ndim, const int64_t* n, double dh,
                 int band, double* out, uint8_t* band_mask, void* work, int64_t* iterations,
                 void* stream);

/* Riemann ghost-fluid star states of material m against material o: at each
 * inside-interfacial cell of m not beside a phi_m == 0 cell whose phi_o is
 * outside-interfacial or zero, the states interpolated at -+1.5 dh along the
 * normal are solved with HLLC in the normal frame; star (cells * (2+ndim)
 * doubles, AoS) receives the left star state with the tangential velocity
 * of m, has_star marks those cells.  eos_m/eos_o give each side's material.
 * Replaces RiemannGFM::invoke (ghost_fluid/riemann_ghost_fluid.hpp:50-195). */
int fl_gfm_star(const double* phi_m, const double* phi_o, const fl_field* state_m,
                const fl_field* state_o, const fl_scheme* eos_m, const fl_scheme* eos_o,
                double dh, double* star, uint8_t* has_star, void* stream);

/* Fills the ghost band 0 < phi_m <= 6 dh / sqrt(2) of state_m by `sweeps`
 * Jacobi sweeps of the normal-weighted upwind average, seeded with the star
 * states (the sweeps walk a list of the band cells).  work:
 * fl_gfm_extrapolate_work_bytes() bytes, 256-byte aligned.  Replaces
 * GhostCellExtrapolator::invoke (ghost_fluid/ghost_fluid_extrapolator.hpp:54-202). */
int64_t fl_gfm_extrapolate_work_bytes(int ndim, const int64_t* n);
int fl_gfm_extrapolate(const double* phi_m, fl_field* state_m, const double* star,
                       const uint8_t* has_star, double dh, int sweeps, void* work,
                       void* stream);

/* out = state0 where phi0 <= 0, else state1.  Replaces combine_into_first
 * (material/cuda/combine_materials.cuh:39-63). */
int fl_mm_combine(const double* phi0, const fl_field* state0, const fl_field* state1,
                  fl_field* out, void* stream);

/* One material of a ghost-fluid run: its states over the whole grid, a
 * same-shaped buffer, its level set and its scheme (eos, gamma, p_inf and
 * the solver options). */
typedef struct fl_material {
  fl_field  state;
  fl_field  work;
  double*   phi;
  fl_scheme scheme;
} fl_material;

typedef struct fl_mm_params {
  int32_t band;           /* eikonal band in cells (default 6)                 */
  int32_t extrap_sweeps;  /* ghost extrapolation sweeps (default 8)            */
  int32_t reinit_every;   /* reinitialise after every k-th step (default 1,    */
                          /* the reference's every step)                       */
  int32_t reserved;
} fl_mm_params;

/* The multimaterial loop of MultimaterialSimulator::simulate_mm
 * (simulator/multimaterial_simulator.hpp:227-303, 367-483) for nmat = 2..8
 * materials.  Per step: star states of every material (material m's from
 * the last other material whose level set is outside-interfacial or on the
 * boundary at the cell, ghost_fluid/riemann_ghost_fluid.hpp:56-190), then
 * extrapolation of each; dt = cfl dh / max wavespeed over every material's
 * grid; each material's split solve; level set m advected with u . n_m;
 * projection (two level sets: levelset_projection.cuh:54-66; more: the
 * average of the two smallest values subtracted from each,
 * levelset_projection.cuh:87-118); reinitialisation of every level set
 * after every reinit_every-th step.  Each pass covers both materials in
 * one launch (the sweeps too, when both schemes use the same flux method).
 * Between reinitialisations the level sets alternate between phi and a work
 * buffer; they are back in phi on FL_OK and on the FL_ERR_NUMERIC of a NaN or
 * vacuum wavespeed (the fields then hold the last completed step, with the
 * ghost cells the failed step had written, as the reference's loop would).  After
 * other errors (a HIP failure, an eikonal solve that does not converge) the
 * states and level sets are unspecified.  Stops when run_time >= sim_time or
 * iters >= max_iters.  dt is computed on the device (in the context's control
 * block, which the sweeps and the level-set advection read); the host reads a
 * copy of that block once per step, after the step's sweeps are queued, to
 * test the loop condition, and waits for the eikonal solve's convergence once
 * per reinitialisation (more often only when a solve needs more sweeps than
 * the previous one).  Both materials' states (and work grids) share one
 * dtype: FP64, or FP32 bulk storage (FP32 sweeps; dt, the wavespeed maximum,
 * the level sets, the ghost-fluid states and the eikonal solve in FP64).
 * work: fl_mm_work_bytes_n(nmat, ...) (fl_mm_work_bytes() for two).  mp may
 * be NULL (defaults). */
int64_t fl_mm_work_bytes(int ndim, const int64_t* n);
/* fl_mm_simulate's workspace for nmat (2..8) materials (fl_mm_work_bytes is
 * the nmat = 2 case); -1 for a bad argument. */
int64_t fl_mm_work_bytes_n(int nmat, int ndim, const int64_t* n);
int fl_mm_simulate(fl_context* ctx, fl_material* mats, int nmat, const fl_sim_params* params,
                   const fl_mm_params* mp, void* work, int64_t work_bytes,
                   fl_sim_result* result, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDITY_AMD_H */
