/*
 * alsvinn_hip.h -- C ABI of the MI355X finite-volume engine (platform "hip").
 *
 * One RK time step of a hyperbolic conservation law on a Cartesian grid:
 * reconstruction (none / weno2 / eno2 / eno3) -> numerical flux (godunov for
 * burgers; hll / hll3 for euler1/2/3) -> net-flux divergence -> RK stage
 * combination (forwardeuler / rungekutta2 / rungekutta3), with the CFL time
 * step from an on-device max-wave-speed reduction and periodic / neumann
 * ghost cells resolved on the device.  All arithmetic is IEEE FP64 in the
 * operation order of the reference CPU path.
 *
 * Plain C types only.  Every call returns an avh_status; avh_last_error()
 * gives a thread-local message for the last failing call on this thread.
 * A solver is driven by one host thread (it owns one HIP stream); distinct
 * solvers may be driven concurrently.
 *
 * Reference interfaces each entry point stands in for (files of
 * alsvinn/alsvinn v0.5.4):
 *   avh_solver_create   -> Simulator::Simulator + NumericalFluxFactory /
 *                          ReconstructionFactory / IntegratorFactory /
 *                          BoundaryFactory dispatch on platform
 *                          (alsfvm/src/simulator/Simulator.cpp:26-64,
 *                           alsfvm/src/numflux/NumericalFluxFactory.cpp:43-200,
 *                           alsfvm/src/reconstruction/ReconstructionFactory.cpp:41-355,
 *                           alsfvm/src/integrator/IntegratorFactory.cpp:31-44,
 *                           alsfvm/src/boundary/BoundaryFactory.cpp:48-75)
 *   avh_set_state       -> Simulator::setSimulationState
 *                          (alsfvm/src/simulator/Simulator.cpp:100-103)
 *   avh_get_state       -> the ghosted conserved Volume the writers receive
 *                          (alsfvm/src/simulator/Simulator.cpp:164-170)
 *   avh_add_fixed_interval_adjuster
 *                       -> FixedIntervalWriter as a TimestepAdjuster
 *                          (alsfvm/src/io/FixedIntervalWriter.cpp:24-53,
 *                           alsfvm/src/config/SimulatorSetup.cpp:222-228)
 *   avh_step            -> Simulator::performStep = incrementSolution +
 *                          checkConstraints + callWriters
 *                          (alsfvm/src/simulator/Simulator.cpp:81-85,180-214)
 *   avh_advance         -> the driver loop `while (!atEnd()) performStep()`
 *                          (alsvinncli/src/main.cpp:210-232)
 *   avh_compute_flux    -> NumericalFlux::computeFlux over the whole domain
 *                          (alsfvm/include/alsfvm/numflux/NumericalFlux.hpp:44-52)
 */
#ifndef ALSVINN_HIP_H
#define ALSVINN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum avh_status {
    AVH_OK = 0,
    AVH_EINVAL = 1,      /* bad argument / unsupported combination        */
    AVH_EHIP = 2,        /* a HIP runtime call failed                     */
    AVH_ECONSTRAINT = 3, /* state violates the equation's constraints     */
    AVH_ENOMEM = 4       /* device allocation failed                      */
} avh_status;

/* equation names as in the reference XML <equation> */
enum { AVH_EQ_BURGERS = 0, AVH_EQ_EULER1 = 1, AVH_EQ_EULER2 = 2, AVH_EQ_EULER3 = 3 };
/* <reconstruction> */
enum { AVH_RECON_NONE = 0, AVH_RECON_WENO2 = 1, AVH_RECON_ENO2 = 2, AVH_RECON_ENO3 = 3 };
/* <flux> */
enum { AVH_FLUX_GODUNOV = 0, AVH_FLUX_HLL = 1, AVH_FLUX_HLL3 = 2 };
/* <integrator> */
enum { AVH_INT_FORWARDEULER = 0, AVH_INT_RUNGEKUTTA2 = 1, AVH_INT_RUNGEKUTTA3 = 2 };
/* <boundary> */
enum { AVH_BC_PERIODIC = 0, AVH_BC_NEUMANN = 1 };

typedef struct avh_config {
    int equation;       /* AVH_EQ_*                                        */
    int reconstruction; /* AVH_RECON_*                                     */
    int flux;           /* AVH_FLUX_*                                      */
    int integrator;     /* AVH_INT_*                                       */
    int boundary;       /* AVH_BC_*                                        */
    int n[3];           /* interior cells per axis; 1 on inactive axes      */
    double lower[3];    /* <lowerCorner>                                   */
    double upper[3];    /* <upperCorner>                                   */
    double gamma;       /* euler gamma (reference default 1.4)             */
    double cfl;         /* CFL number                                      */
} avh_config;

typedef struct avh_ctx avh_ctx;
typedef struct avh_solver avh_solver;

/* Step counters and flags read back from the device. */
typedef struct avh_progress {
    double time;          /* current simulation time                        */
    double last_dt;       /* dt of the last completed step                  */
    int64_t steps;        /* number of completed steps                      */
    int32_t violated;     /* 1 if the last step broke the constraints       */
    int32_t saves_pending;/* number of fixed-interval saves hit, not acked  */
} avh_progress;

const char* avh_last_error(void);

avh_status avh_ctx_create(int device, avh_ctx** ctx);
avh_status avh_ctx_destroy(avh_ctx* ctx);

avh_status avh_solver_create(avh_ctx* ctx, const avh_config* cfg, avh_solver** solver);
avh_status avh_solver_destroy(avh_solver* solver);

/* Number of ghost cells per active axis (2 weno2/eno2, 3 eno3, 1 none). */
avh_status avh_ghost_cells(const avh_solver* solver, int* ghosts);
/* Conserved-variable count (1 burgers, 2+nsd euler). */
avh_status avh_num_vars(const avh_solver* solver, int* nvars);

/* Host SoA state, reference Volume layout: nvars consecutive arrays, each
 * (n[0]+2g)*(n[1]+2g')*(n[2]+2g'') doubles, x fastest, g' = g on active
 * axes and 0 on inactive ones.  Ghost entries are ignored by set and filled
 * with the boundary condition by get. */
avh_status avh_set_state(avh_solver* solver, const double* host_ghosted);
avh_status avh_get_state(avh_solver* solver, double* host_ghosted);
/* Same, but the state is already in device memory (interior-only SoA,
 * nvars * n[0]*n[1]*n[2] doubles).  The copy is device-to-device. */
avh_status avh_set_state_device(avh_solver* solver, const double* dev_interior);
avh_status avh_get_state_device(avh_solver* solver, double* dev_interior);

/* Time bookkeeping: current time and step count (reference
 * TimestepInformation).  Resets the device counters. */
avh_status avh_set_time(avh_solver* solver, double time, int64_t steps);

/* Register a FixedIntervalWriter-equivalent timestep adjuster: next save
 * at numberSaved*interval + start_time, numberSaved starts at 0 when
 * write_initial != 0 and at 1 otherwise.  At most 4 per solver. */
avh_status avh_add_fixed_interval_adjuster(avh_solver* solver, double interval,
                                           double start_time, int write_initial);

/* One full time step (all RK stages, CFL dt, adjusters, constraint check,
 * save counters).  Blocks until done.  Returns AVH_ECONSTRAINT if the new
 * state violates the constraints. */
avh_status avh_step(avh_solver* solver, double* dt_out);

/* Advance until time >= t_end (reference atEnd), at most max_steps steps,
 * enqueueing chunk steps at a time between host synchronisations.  Steps
 * past the end are no-ops on the device.  If pause_on_save != 0 it returns
 * after the step that reached a fixed-interval save time (progress.
 * saves_pending > 0) so the caller can read the state; call
 * avh_ack_saves() before continuing. */
avh_status avh_advance(avh_solver* solver, double t_end, int64_t max_steps,
                       int chunk, int pause_on_save, avh_progress* progress);
avh_status avh_ack_saves(avh_solver* solver);
avh_status avh_get_progress(avh_solver* solver, avh_progress* progress);

/* Set the end time used by the device-side loop test (time >= t_end makes
 * further enqueued steps no-ops).  Synchronises. */
avh_status avh_set_end_time(avh_solver* solver, double t_end);
/* Enqueue `steps` full time steps on the solver's stream without any host
 * synchronisation (for timing and for callers that poll progress). */
avh_status avh_enqueue_steps(avh_solver* solver, int64_t steps);
avh_status avh_synchronize(avh_solver* solver);
/* Steps per captured hipGraph used by avh_enqueue_steps / avh_advance
 * (default 16; 0 or 1: ordinary launches).  A chunk of steps is captured
 * once and replayed; results are identical either way.  The reference
 * launches each kernel separately (alsfvm/src/simulator/Simulator.cpp:191-214). */
avh_status avh_set_graph_steps(avh_solver* solver, int steps);
/* One full time step of cfg's scheme (all RK stages, CFL dt, constraint
 * check) on a device state, allocating nothing: `state` is the interior-only
 * SoA (nvars * n[0]*n[1]*n[2] doubles, x fastest) and is advanced in place;
 * dt = min(cfl / sum_d(ws_d / dx_d), dt_cap) (dt_cap = +inf: no clip);
 * dt_out[0] (device double, may be NULL) receives dt and violated_out[0]
 * (device int, may be NULL) 1 if the new state breaks the constraints.
 * Everything is enqueued on `stream` (NULL: the null stream) with no host
 * synchronisation, no memset and no allocation, so the call can be
 * captured into a hipGraph.  The stage buffers live in a static device
 * array of 2 x 4194304 doubles (64 MiB, held by every GPU the library runs
 * on): states of more than 4194304 values (nvars * cells) are rejected with
 * AVH_EINVAL.  All calls on one GPU share that array: calls that overlap in
 * time on one GPU (other streams, other threads) race.  The tile shape is
 * the autotuner's choice if a solver of the same problem made one, else a
 * fixed default.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops). */
avh_status avh_step_noalloc(avh_ctx* ctx, const avh_config* cfg, double* state, double dt_cap, double* dt_out,
                            int* violated_out, void* stream);
/* Resolve the tile shape (autotuning if none is fixed) and capture and
 * instantiate the step-chunk graph now, without launching it, so that a
 * later avh_enqueue_steps / avh_advance only replays.  No reference
 * counterpart (set-up cost kept out of a timed or latency-bound loop). */
avh_status avh_prepare_graph(avh_solver* solver);
/* Tile shape of the fused stage kernel: -1 (default) lets the autotuner
 * pick on first use (cached per equation, reconstruction, flux and grid for
 * the process), 0..nshapes-1 fixes one (1D: 256 or 64 cells; 2D: 32x16,
 * 16x15, 64x4, 32x15 or 24x20; 3D columns marching over z-planes: 32x8,
 * 16x15 and 32x7 columns over 32 planes with one plane buffer, then 16x15
 * over 32 and over 16 planes and 32x7 over 32 planes with a double-buffered
 * plane and two barriers per plane, then 16x15 over 32 planes, two
 * barriers, fluxes with shared reciprocals).  Results are identical for
 * every shape.
 * No reference counterpart: the reference's CUDA kernels use fixed block
 * sizes (alsfvm/src/numflux/NumericalFluxCUDA.cu:330-414). */
avh_status avh_set_tile_shape(avh_solver* solver, int shape);
/* Time every tile shape on the solver's current state now (a flux-only
 * launch, and for RK3 the second stage on its output, into the solver's
 * scratch buffers; the solution and step state are unchanged) and select
 * the fastest.  ms_per_shape (may be NULL) receives *nshapes <= 8 timings. */
avh_status avh_autotune(avh_solver* solver, int* best_shape, double* ms_per_shape, int* nshapes);

/* Run `steps` full time steps with a HIP event pair around every kernel
 * launch on the solver's stream (the hipEvent analogue of the reference's
 * ALSVINN_TIME_CUDA_BLOCK timers, alsutils/include/alsutils/timer/CudaTimer.hpp:11-48).
 * ms_per_launch[k] receives the average duration in milliseconds of the
 * k-th kernel of a step (k < *kernels_per_step <= 4), in launch order: the
 * stages and the element-wise pass of forward euler / RK2, and forward
 * euler's end-of-step kernel (the last RK2 / RK3 stage ends the step in its
 * last workgroup). */
avh_status avh_time_step_kernels(avh_solver* solver, int64_t steps, double* ms_per_launch,
                                 int* kernels_per_step);

/* HIP stream the solver launches on (a hipStream_t as void*). */
void* avh_stream(avh_solver* solver);

/* NumericalFlux::computeFlux over the whole domain on the current state:
 * writes the unscaled net-flux divergence (interior-only SoA) to
 * host_interior_out and the per-direction max wave speeds to ws[3]. */
avh_status avh_compute_flux(avh_solver* solver, double* host_interior_out, double ws[3]);
/* Unfused reconstruction of the current state along one active axis:
 * left[v][cell] / right[v][cell] = the value at the cell's lower / upper
 * face, interior SoA, ghosts through the boundary conditions
 * (Reconstruction::performReconstruction,
 * alsfvm/include/alsfvm/reconstruction/Reconstruction.hpp:46-58; the stage
 * kernels fuse this step).  Synchronises. */
avh_status avh_reconstruct(avh_solver* solver, int direction, double* host_left, double* host_right);

/* ---------------------------------------------------------------------------
 * The reference's per-call plug-in interfaces on device Volumes.
 *
 * avh_volume describes a reference Volume in device memory: one array per
 * conserved variable (Volume::getScalarMemoryArea(v)->getPointer()), each
 * (n[0]+2*ghosts[0]) * (n[1]+2*ghosts[1]) * (n[2]+2*ghosts[2]) doubles, x
 * fastest, ghosts[d] = 0 on inactive axes (alsfvm/src/volume/Volume.cpp:44-60).
 * Ghost cells are read as the caller filled them.  These are the calls the
 * hip classes behind the reference's factory dispatch make (HipNumericalFlux,
 * HipReconstruction, HipBoundary, HipCellComputer, HipIntegrator).  Each
 * runs on `stream` (a hipStream_t; NULL: the null stream) and returns after
 * the work is done.  Calls that share a context serialise on it.
 * ------------------------------------------------------------------------- */
typedef struct avh_volume {
    double* var[5]; /* device array of each conserved variable (nvars used) */
    int nvars;
    int n[3];       /* interior cells per axis, 1 on inactive axes */
    int ghosts[3];  /* ghost cells per axis, 0 on inactive axes */
} avh_volume;

/* NumericalFlux::computeFlux(in, waveSpeed, computeWaveSpeed, out, start,
 * end) (alsfvm/include/alsfvm/numflux/NumericalFlux.hpp:44-52,
 * alsfvm/src/numflux/NumericalFluxCPU.cpp:170-221): the cells of `out` whose
 * interior coordinates lie in [start[d], n[d] + end[d]) receive the net flux
 * sum over active d of ((-F(i+1/2)) + F(i-1/2)) (not divided by the cell
 * size), and ws[d] the largest wave speed over the interfaces of that box
 * along d (host array).  start >= 0 and end <= 0 on active axes, 0 on
 * inactive ones; the volumes need at least the reconstruction's ghost
 * cells.  cfg supplies equation, reconstruction, flux and gamma. */
avh_status avh_numerical_flux(avh_ctx* ctx, const avh_config* cfg, const avh_volume* in, avh_volume* out,
                              const int start[3], const int end[3], double ws[3], void* stream);

/* Reconstruction::performReconstruction(in, direction, indicatorVariable,
 * left, right, start, end) (alsfvm/include/alsfvm/reconstruction/
 * Reconstruction.hpp:46-58, alsfvm/src/reconstruction/ReconstructionCPU.cpp:
 * 37-97, ENOCPU.cpp:41-187): left/right receive the values at the lower /
 * upper face along `direction` of every cell with interior coordinates in
 * [start[d] - (d == direction), n[d] + end[d] + (d == direction)).  ENO
 * chooses its stencils on `indicator_variable`; WENO2 weights on the
 * equation's weight variable (rho / u), as the reference.  Reconstruction
 * "none" copies every cell of the volume, ghost cells included, whatever
 * start/end say (alsfvm/src/reconstruction/NoReconstruction.cpp:46-60). */
avh_status avh_reconstruction(avh_ctx* ctx, const avh_config* cfg, const avh_volume* in, int direction,
                              int indicator_variable, avh_volume* left, avh_volume* right,
                              const int start[3], const int end[3], void* stream);

/* Boundary::applyBoundaryConditions(volume, grid) (alsfvm/include/alsfvm/
 * boundary/Boundary.hpp:30-31, alsfvm/src/boundary/BoundaryCPU.cpp:30-121):
 * every ghost cell of every variable, corners included, from the interior
 * by the periodic or neumann map (AVH_BC_*), in place. */
avh_status avh_apply_boundary(avh_ctx* ctx, int boundary, avh_volume* volume, void* stream);

/* CellComputer::obeysConstraints(conserved, extra) (alsfvm/include/alsfvm/
 * equation/CellComputer.hpp:61, alsfvm/src/equation/CPUCellComputer.cpp:
 * 79-125): *obeys = 1 if every cell, ghost cells included, has finite
 * conserved and derived variables and (euler) rho > 0 and p > 0. */
avh_status avh_obeys_constraints(avh_ctx* ctx, const avh_config* cfg, const avh_volume* volume, int* obeys,
                                 void* stream);

/* Integrator::performSubstep(inputs, cellLengths, dt, cfl, out, substep,
 * timestepInformation) -> dt (alsfvm/include/alsfvm/integrator/
 * Integrator.hpp:76-80; ForwardEuler.cpp:32-51, RungeKutta2.cpp:37-67,
 * RungeKutta3.cpp:49-86) for cfg->integrator: out = the system applied to
 * inputs[substep] over the whole interior (ConservedSystem::operator(),
 * alsfvm/src/simulator/ConservedSystem.cpp:31-103), then the substep's
 * combination with inputs[0..substep] on every cell of `out`, ghosts
 * included, as the reference's Volume operators do.  At substep 0 the step
 * is dt = cfl / sum_d(ws_d / dx_d) (Integrator::computeTimestep,
 * alsfvm/src/integrator/Integrator.cpp:20-46) with dx from cfg's corners,
 * then min(dt, dt_cap) -- dt_cap carries the caller's timestep adjusters
 * (FixedIntervalWriter: next save time - t; +inf if none); later substeps
 * use dt as given.  *dt_out receives the dt used. */
avh_status avh_substep(avh_ctx* ctx, const avh_config* cfg, const avh_volume* const* inputs, int substep,
                       avh_volume* out, double dt, double dt_cap, double* dt_out, void* stream);

/* Device self-test of the FP64 arithmetic the stage kernels use in place of
 * the compiler's sqrt and divide where the operands admit it: n operands
 * drawn from a counter-based generator (seed) over the admitted range, each
 * result compared bit for bit with sqrt() or '/' on the device.
 * kind 0: sqrt over [2^-767, DBL_MAX]; 1: shared-reciprocal quotient with
 * denominators in [2^-100, 2^100] and numerators 0 or +-[2^-800, 2^600];
 * 2: WENO2 weight a0 / (a0 + a1), a0, a1 in [2^-400, 2^53];
 * 3: Roe quotient, denominators in [2^-50, 2^51]; 4: the binary exponent
 * the range checks read (v_frexp_exp_i32_f64) against the one the bits
 * give, on signed zeros, subnormals and normals.  Bounds are included.
 * mismatches = count of
 * differing results, first_bad[2] (may be NULL) = operands of one of them.
 * No reference counterpart (the reference divides and takes square roots
 * directly; this checks that the shortcuts give the same bits). */
avh_status avh_arith_selftest(avh_ctx* ctx, int kind, uint64_t seed, int64_t n, int64_t* mismatches,
                              double* first_bad);

/* ------------------------------------------------------------------
 * Batched Monte Carlo (alsuq on the hip platform).
 *
 * A batch holds up to 4 resolution groups of samples of one configuration
 * (avh_config.n is ignored; each group has its own grid).  Every launch
 * advances all samples of all groups by one RK stage; each sample keeps its
 * own time, dt, save counters, constraint flag and loop test, as the
 * reference's Runner advances one Simulator per sample
 * (alsuq/src/run/Runner.cpp:39-90).  Statistics (meanvar) are accumulated on
 * the device at the statistics save times of FixedIntervalStatistics
 * (alsuq/src/stats/FixedIntervalStatistics.cpp:33-86) for the conserved
 * variables (MeanVariance, alsuq/src/stats/MeanVariance.cpp:30-67) and for
 * rho, u, p.  States are interior-only SoA, nvars * n[0]*n[1]*n[2] doubles. */
typedef struct avh_batch avh_batch;
typedef struct avh_group {
    int n[3];      /* grid of this resolution group */
    int nsamples;  /* samples in the group */
} avh_group;

/* stats_saves = the statistics' numberOfSaves (0: no statistics); the
 * statistics save interval is t_end / stats_saves (alsuq/src/config/
 * Setup.cpp:270-285).  stats_write_initial records the initial state as
 * save 0. */
avh_status avh_batch_create(avh_ctx* ctx, const avh_config* cfg, int ngroups, const avh_group* groups,
                            double t_end, int stats_saves, int stats_write_initial, avh_batch** batch);
avh_status avh_batch_destroy(avh_batch* batch);
/* A further FixedIntervalWriter adjuster on every sample (the fvm
 * <writer numberOfSaves>, alsfvm/src/config/SimulatorSetup.cpp:222-228). */
avh_status avh_batch_add_fixed_interval_adjuster(avh_batch* batch, double interval, double start_time,
                                                 int write_initial);
avh_status avh_batch_set_state(avh_batch* batch, int group, int sample, const double* host_interior);
avh_status avh_batch_get_state(avh_batch* batch, int group, int sample, double* host_interior);
/* 2D Kelvin-Helmholtz initial data on the device from per-sample band
 * interfaces: host_interfaces[s][0][x] = lower interface y, [s][1][x] =
 * upper interface y, at the x cell midpoints (euler2 only). */
avh_status avh_batch_set_kh2d(avh_batch* batch, int group, const double* host_interfaces);
/* Advance all samples until each reaches t_end (or violates its
 * constraints), at most max_steps launches of full steps, synchronising
 * every chunk steps.  active_left = samples still running. */
avh_status avh_batch_advance(avh_batch* batch, int64_t max_steps, int chunk, int64_t* steps_done,
                             int64_t* active_left);
avh_status avh_batch_get_progress(avh_batch* batch, int group, int sample, avh_progress* progress);
/* Samples of the group whose state failed the constraint check
 * (Simulator::checkConstraints, alsfvm/src/simulator/Simulator.cpp:180-189,
 * throws there; a batched sample stops instead): count and the first such
 * sample (-1 if none). */
avh_status avh_batch_violations(avh_batch* batch, int group, int64_t* count, int64_t* first_sample);
/* Set every sample's end time (Simulator::setEndTime analogue): samples with
 * time < t_end and no constraint failure become active again.  With a
 * fixed-interval adjuster whose save times include t_end, every sample stops
 * exactly at t_end, so observables can be read there. */
avh_status avh_batch_set_end_time(avh_batch* batch, double t_end);
/* out[s] = (sum over cells of 0.5 * rho * u_c^2) * dV for each sample s of the
 * group at its current state, u_c = m_c / rho, c < dimensions (euler only); a
 * functional modelled on alsfvm/src/functional/Moment.cpp:31-45. */
avh_status avh_batch_kinetic_energy(avh_batch* batch, int group, int component, double* out);
/* Structure function of variable `var` per sample at its current state, the
 * structure_cube statistic (alsuq/src/stats/StructureCube.cpp:37-76,
 * alsfvm/include/alsfvm/functional/structure_common.hpp:36-154): out[s][h] =
 * sum over cells and over the cells on the square ring at distance h of
 * |u - u'|^p, divided by the number of cells, h < number_of_h (out[s][0] = 0);
 * ring cells through the grid's boundary condition.  1D and 2D grids. */
avh_status avh_batch_structure_cube(avh_batch* batch, int group, int var, int number_of_h, double p,
                                    double* out);
/* Steps per captured hipGraph in avh_batch_advance (default 8; 0 or 1:
 * ordinary launches). */
avh_status avh_batch_set_graph_steps(avh_batch* batch, int steps);
/* Tile shape and autotuner of the batched stage kernel, as
 * avh_set_tile_shape / avh_autotune. */
avh_status avh_batch_set_tile_shape(avh_batch* batch, int shape);
avh_status avh_batch_autotune(avh_batch* batch, int* best_shape, double* ms_per_shape, int* nshapes);
/* Advance every sample `steps` full steps with a HIP event after every
 * launch; ms_per_launch[k] = average duration of the k-th launch of a step
 * (k < *kernels_per_step <= 6: stages, element-wise pass, end-of-step
 * kernel, statistics kernel), as avh_time_step_kernels. */
avh_status avh_batch_time_step_kernels(avh_batch* batch, int64_t steps, double* ms_per_launch,
                                       int* kernels_per_step);
/* Mean and population variance (M2/M) of statistics save `save`;
 * primitive = 0: conserved variables, 1: rho, u..., p.  count = samples. */
avh_status avh_batch_get_statistics(avh_batch* batch, int group, int save, int primitive, double* mean,
                                    double* variance, int64_t* count);

/* Sample parameters as the reference's STLMersenne + Uniform generate them
 * (alsuq/src/generator/STLMersenne.cpp:25-85, alsuq/src/distribution/
 * Uniform.cpp:27-34, alsuq/src/samples/SampleGenerator.cpp:29-56): one
 * default-seeded std::mt19937_64, uniform_real_distribution<double>.
 * out[(s - sample_start)][...] holds the parameters of sample s in the
 * given order, each dims[p] long, scaled to [lower[p], upper[p]). */
avh_status avh_uq_generate(int nparams, const int* dims, const double* lower, const double* upper,
                           int64_t sample_start, int64_t nsamples, double* out);

/* ---------------------------------------------------------------------------
 * Device-side meanvar statistics for one-solver-per-sample runs: the hip
 * MeanVariance statistics (alsuq/src/stats/MeanVariance.cpp:30-67) a
 * REGISTER_STATISTICS(hip, meanvar, ...) class would wrap.  Each save keeps
 * a sample count and Welford mean / M2 arrays of the conserved variables and
 * of rho, u, p, all interior-only SoA.
 * ------------------------------------------------------------------------- */
typedef struct avh_stats avh_stats;
avh_status avh_stats_create(avh_ctx* ctx, int equation, const int n[3], double gamma, int nsaves,
                            avh_stats** stats);
avh_status avh_stats_destroy(avh_stats* stats);
/* zero every save (count, mean, M2) */
avh_status avh_stats_reset(avh_stats* stats);
/* fold the solver's current state into save `save` */
avh_status avh_stats_accumulate(avh_stats* stats, avh_solver* solver, int save);
/* mean and population variance (M2 / count) of save `save`, conserved
 * (primitive = 0) or rho, u, p (primitive = 1) */
avh_status avh_stats_read(avh_stats* stats, int save, int primitive, double* mean, double* variance,
                          int64_t* count);

/* ---------------------------------------------------------------------------
 * Surrogate network: an ensemble of multilayer perceptrons n_in -> width
 * (GELU) x hidden -> n_out, bf16 weights and activations on the matrix
 * cores with fp32 accumulation, fp32 master weights trained by Adam on the
 * normalised mean-squared error.  No reference counterpart: the reference
 * has no learned model; this maps the sample parameters of its Monte Carlo
 * runs (alsuq/src/run/Runner.cpp:39-90) to their observables.
 * ------------------------------------------------------------------------- */
typedef struct avh_mlp avh_mlp;

typedef struct avh_mlp_config {
    int n_in;      /* inputs (sample parameters) */
    int width;     /* hidden width, multiple of 16 */
    int hidden;    /* hidden layers, 1..7 */
    int n_out;     /* outputs (observables) */
    int members;   /* ensemble members, 1..64 */
    int batch;     /* minibatch rows, multiple of 16 */
    float beta1, beta2, eps;  /* Adam */
    uint64_t seed; /* initial weights and shuffles */
} avh_mlp_config;

avh_status avh_mlp_create(avh_ctx* ctx, const avh_mlp_config* cfg, avh_mlp** out);
avh_status avh_mlp_destroy(avh_mlp* mlp);
/* parameters per member: sum over layers of in * out + out */
avh_status avh_mlp_num_params(avh_mlp* mlp, int64_t* count);
/* inputs and targets are standardised as (x - mean) / std inside */
avh_status avh_mlp_set_normalization(avh_mlp* mlp, const float* in_mean, const float* in_std,
                                     const float* out_mean, const float* out_std);
/* fp32 parameters of one member: per layer W[out][in] row-major, then b[out] */
avh_status avh_mlp_get_params(avh_mlp* mlp, int member, float* params);
avh_status avh_mlp_set_params(avh_mlp* mlp, int member, const float* params);
/* loss and gradients of every member on the same `batch` rows, no update:
 * X[rows][n_in], Y[rows][n_out], grads[member][num_params], loss[member] */
avh_status avh_mlp_gradients(avh_mlp* mlp, const float* X, const float* Y, int64_t rows, float* grads,
                             float* loss);
/* decoupled weight decay (AdamW) of the weights in every later Adam step:
 * p -= lr * weight_decay * p beside the Adam update; biases are not
 * decayed; 0 (the default) is plain Adam */
avh_status avh_mlp_set_weight_decay(avh_mlp* mlp, float weight_decay);
/* one Adam step of every member from the last gradients */
avh_status avh_mlp_adam_step(avh_mlp* mlp, float lr);
/* minibatch Adam over `epochs` passes, cosine learning-rate decay from lr to
 * lr * final_lr_fraction over this call; the shuffles continue the epoch
 * count of earlier calls; epoch_loss[epoch][member] (may be NULL) */
avh_status avh_mlp_train(avh_mlp* mlp, const float* X, const float* Y, int64_t rows, int epochs, float lr,
                         float final_lr_fraction, float* epoch_loss);
/* out[member][row][n_out] in data units */
avh_status avh_mlp_predict(avh_mlp* mlp, const float* X, int64_t rows, float* out);

#ifdef __cplusplus
}
#endif

#endif /* ALSVINN_HIP_H */
