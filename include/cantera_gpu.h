/*
 * cantera_gpu.h -- C ABI of libcantera_amd.so, the MI355X (gfx950) engine for
 * batched ideal-gas kinetics and constant-pressure reactor integration.
 *
 * Conventions (mirroring the reference CLib, interfaces/clib + the generated
 * recipes in interfaces/sourcegen/headers/):
 *   - objects are integer handles (reference: Cabinet<T>, src/clib/Cabinet.h:64-121);
 *   - every call returns >= 0 on success and a negative code on failure; the
 *     message is fetched with cg_last_error (reference: ERR = -999 and the
 *     Application error stack, include/cantera/clib/clib_defs.h:27-33,
 *     src/clib/clib_utils.h:30-52).  No C++ exception crosses this boundary.
 *   - units are SI with kmol (reference ct_defs.h); species order is mechanism
 *     order.
 *   - batched arrays are species-major SoA: Y[k*n + s] is species k of state s,
 *     unless a function says otherwise.
 *   - one reactor batch owns one HIP stream; calls on different handles may run
 *     concurrently, calls on one handle must be serialised by the caller
 *     (reference objects are not thread-safe either, openmp_ignition.cpp:28-31).
 */
#ifndef CANTERA_GPU_H
#define CANTERA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_ERR (-999)

enum cg_precision {
    CG_FP64 = 0,   /* plain: everything in FP64 */
    CG_MIXED = 1,  /* FP32 iteration matrix / LU / solves, FP64 residual, state and step control */
    CG_FP32 = 2    /* rates only: the rate constants' exp/log10/exp10 (Arrhenius, Troe/Tsang/Lindemann
                      falloff) in FP32; PLOG, Chebyshev and SRI constants, thermo, equilibrium
                      constants, mass action and wdot in FP64; FP64 I/O */
};

/* ---- library ------------------------------------------------------------ */
int32_t cg_version(void);
/* Copy the last error message of the calling thread into buf (NUL-terminated).
 * Returns the full message length.  Reference: clib_utils.h:30-52 handleAllExceptions. */
int32_t cg_last_error(char* buf, int32_t len);
int32_t cg_device_count(void);
/* Device memory helpers (so callers need no HIP runtime of their own). */
void* cg_dev_alloc(int64_t bytes);
int32_t cg_dev_free(void* p);
int32_t cg_memcpy_h2d(void* dst, const void* src, int64_t bytes);
int32_t cg_memcpy_d2h(void* dst, const void* src, int64_t bytes);
int32_t cg_memcpy_d2d(void* dst, const void* src, int64_t bytes);
int32_t cg_device_synchronize(void);

/* ---- mechanisms ----------------------------------------------------------- */
/* Load a compiled mechanism (.mech, written by cantera_amd/mechc.py from a
 * Cantera YAML file) and upload its tables.  Replaces newSolution(file, phase)
 * for the ideal-gas/bulk-kinetics path (src/base/Solution.cpp:242-282;
 * KineticsFactory "gas"/"bulk", src/kinetics/KineticsFactory.cpp:24-37).
 * At most 254 species: rates and mole-based reactor batches take all of
 * them; the dense reactor batches, Jacobians, equilibrium and the surrogate
 * take at most 126 and refuse larger mechanisms with an error.
 * Returns a mechanism handle. */
int32_t cg_mech_load(const char* path);
int32_t cg_mech_free(int32_t mech);
/* Kinetics::nTotalSpecies / nReactions / Phase::nElements */
int32_t cg_mech_nspecies(int32_t mech);
int32_t cg_mech_nreactions(int32_t mech);
int32_t cg_mech_nelements(int32_t mech);
/* Phase::speciesIndex (returns -1 if absent) / speciesName */
int32_t cg_mech_species_index(int32_t mech, const char* name);
int32_t cg_mech_species_name(int32_t mech, int32_t k, char* buf, int32_t len);
/* Phase::getMolecularWeights (bufLen/buf like kin_getNetProductionRates) */
int32_t cg_mech_molecular_weights(int32_t mech, int32_t len, double* mw);

/* ---- batched rates ---------------------------------------------------------- */
/* Net production rates wdot [kmol/m^3/s] of n independent states given T [K],
 * P [Pa] and (unnormalised, possibly negative) mass fractions Y.  The batched
 * form of Kinetics::getNetProductionRates (include/cantera/kinetics/Kinetics.h:572,
 * src/kinetics/Kinetics.cpp:447-457; CLib kin_getNetProductionRates,
 * interfaces/sourcegen/headers/ctkin.yaml:46) with the state set as
 * IdealGasConstPressureReactor::updateState does
 * (src/zeroD/IdealGasConstPressureReactor.cpp:34-44).
 * Host pointers; copies in and out.  precision: CG_FP64 or CG_FP32. */
int32_t cg_rates_batch(int32_t mech, int64_t n, const double* T, const double* P,
                       const double* Y, double* wdot, int32_t precision);
/* Same on device pointers, no copies, asynchronous on `stream` (hipStream_t,
 * NULL for the default stream). */
int32_t cg_rates_batch_device(int32_t mech, int64_t n, const double* T, const double* P,
                              const double* Y, double* wdot, int32_t precision, void* stream);

/* ---- batched reactor Jacobians ---------------------------------------------- */
/* J = d(ydot)/dy of an IdealGasConstPressureReactor at each of n states, the
 * matrix Reactor::jacobian() returns for it by finite differences
 * (Reactor::finiteDifferenceJacobian, src/zeroD/Reactor.cpp:233-276; eval
 * src/zeroD/IdealGasConstPressureReactor.cpp:46-110), here analytic in the
 * species columns (the chain rule of BulkKinetics::netProductionRates_ddCi,
 * src/kinetics/BulkKinetics.cpp:761-811) and a one-sided difference in T.
 * T[n], P[n], Y[K*n] species-major as cg_rates_batch; J[n*N*N] with N = K+2,
 * row-major per state, state order [m, T, Y_0..Y_{K-1}].  ydot is independent
 * of the mass, so the m row and column are zero.  precision: CG_FP64 (J holds
 * doubles) or CG_FP32 (J holds floats: the FP64 assembly rounded once at the
 * store, half the bytes written; K + 2 <= 64). */
int32_t cg_jac_batch(int32_t mech, int64_t n, const double* T, const double* P,
                     const double* Y, void* J, int32_t precision);
/* Same on device pointers, asynchronous on `stream`. */
int32_t cg_jac_batch_device(int32_t mech, int64_t n, const double* T, const double* P,
                            const double* Y, void* J, int32_t precision, void* stream);
/* Chemical equilibrium of n states (ThermoPhase::equilibrate(XY),
 * src/thermo/ThermoPhase.cpp; for one ideal-gas phase the reference's default
 * is the element-potential solver ChemEquil, src/equil/ChemEquil.cpp:329-337):
 * XY = "HP" holds each state's specific enthalpy and pressure, "TP" its
 * temperature and pressure.  T[n], P[n], Y[K*n] species-major as
 * cg_rates_batch; T_out[n], Y_out[K*n] the equilibrium states; info[n] (may
 * be NULL): Newton iterations, or -1 (no convergence in 500 iterations) / -3
 * (singular system).  Elements and mass are conserved exactly up to rounding.
 * Returns CG_ERR if any state failed (the reference throws). */
int32_t cg_equilibrate_batch(int32_t mech, int64_t n, const char* XY, const double* T, const double* P,
                             const double* Y, double* T_out, double* Y_out, int32_t* info);
/* Same on device pointers, asynchronous on `stream`, at most max_iter
 * iterations; failures are reported in info only. */
int32_t cg_equilibrate_batch_device(int32_t mech, int64_t n, const char* XY, const double* T, const double* P,
                                    const double* Y, double* T_out, double* Y_out, int32_t* info, int32_t max_iter,
                                    void* stream);
/* Duration [ms] of the last batched-Jacobian kernel on `mech`. */
double cg_jac_last_kernel_ms(int32_t mech);

/* ---- batched reactors ------------------------------------------------------- */
/* Create n IdealGasConstPressureReactor + ReactorNet pairs (one per state) with
 * V = 1 m^3, integrator tolerances rtol/atol for every component
 * (ReactorNet::setTolerances, src/zeroD/ReactorNet.cpp:233-245).  Y0 is AoS
 * n x K (row s = mass fractions of reactor s).  Returns a batch handle.
 * Reference: IdealGasConstPressureReactor(sol) + ReactorNet(reactor),
 * samples/cxx/openmp_ignition/openmp_ignition.cpp:36-40. */
int32_t cg_reactors_create(int32_t mech, int64_t n, const double* T0, const double* P,
                           const double* Y0, double rtol, double atol, int32_t precision);
/* Same for n closed, rigid, adiabatic IdealGasReactor + ReactorNet pairs:
 * V = 1 m^3 and mass rho(T0, P0, Y0) V, state (T, rho) thereafter
 * (IdealGasReactor::updateState / eval, src/zeroD/IdealGasReactor.cpp:38-117:
 * m c_v dT/dt = -sum_k u_k wdot_k V).  The state returned by
 * cg_reactors_get_state is [m, T, Y] as for the constant-pressure reactor (the
 * reference's [m, V, T, Y] without the constant V); the WRMS norm counts V, as
 * CVODES does on the reference's state.  The surrogate path rejects such a
 * batch.  Reference: IdealGasReactor(sol) + ReactorNet(reactor). */
int32_t cg_reactors_create_const_volume(int32_t mech, int64_t n, const double* T0, const double* P,
                                        const double* Y0, double rtol, double atol, int32_t precision);
int32_t cg_reactors_free(int32_t batch);
/* n IdealGasConstPressureMoleReactors (state [T, n_1..n_K], kmol; V = 1 m^3
 * at (T0, P0, Y0)) in ReactorNets with the AdaptivePreconditioner and the
 * GMRES linear solver (IdealGasConstPressureMoleReactor.cpp; ReactorNet::
 * setPreconditioner / preconditionerSetup, src/zeroD/ReactorNet.cpp:330-339,
 * 881-908; CVodesIntegrator.cpp:424-434 with SPGMR, right preconditioning).
 * The preconditioner Jacobian is IdealGasConstPressureMoleReactor::
 * getJacobianElements (:108-174); M = I - gamma J is factored completely on
 * a fixed fill pattern (minimum-degree order) where the reference uses an
 * incomplete LUT.  flags: 1 skip third-body derivatives, 2 skip falloff
 * derivatives (BulkKinetics derivative settings), 4 atol applies to every
 * component (ReactorNet::setTolerances); without 4 the species' atol is atol
 * times the initial total moles (MoleReactor::updateDefaultTolerances,
 * src/zeroD/MoleReactor.cpp:64-76); 8 the constant-volume IdealGasMoleReactor
 * (src/zeroD/IdealGasMoleReactor.cpp, state [T, V, n_1..n_K]).  FP64; at
 * most 254 species (the dense batches take 126); the preconditioner's fill
 * pattern must fit in one CU's LDS (160 KiB with the reactor's scratch; the
 * error names the size, and flags 1 | 2 make it sparser).  Advance, status, stats and
 * cg_reactors_get_state (which returns [T, n_1..n_K], or [T, V, n_1..n_K])
 * work as for the other batches. */
int32_t cg_reactors_create_mole(int32_t mech, int64_t n, const double* T0, const double* P, const double* Y0,
                                double rtol, double atol, int32_t flags);
/* Linear-solver counters per reactor, n x 6 int64: lin_iters, prec_evals,
 * prec_solves, jt_vec_prod_evals, lin_rhs_evals, lin_conv_fails
 * (CVodesIntegrator::solverStats, src/numerics/CVodesIntegrator.cpp:
 * 572-621); zero for batches with a direct linear solver. */
int32_t cg_reactors_linear_stats(int32_t batch, int64_t* stats);
/* Advance every reactor s to t_out[s] (host array of n).  History persists
 * between calls; each reactor steps past t_out[s] and interpolates back, as
 * ReactorNet::advance does (src/zeroD/ReactorNet.cpp:351-357;
 * CVodesIntegrator::integrate, src/numerics/CVodesIntegrator.cpp:479-529).
 * Returns 0, or CG_ERR if any reactor failed (see cg_reactors_status). */
int32_t cg_reactors_advance(int32_t batch, const double* t_out);
/* Same with t_out a device pointer; asynchronous (no error check on reactor
 * status until cg_reactors_status / cg_reactors_get_state). */
int32_t cg_reactors_advance_device(int32_t batch, const double* t_out_dev);
/* n_steps successive advances, to t_out_dev[k*n + r] for k = 0..n_steps-1, as
 * repeated ReactorNet::advance calls (src/zeroD/ReactorNet.cpp:351-357) would
 * do them.  The n_steps integrator launches are captured once as a HIP graph
 * and replayed by later calls with the same n_steps and t_out_dev pointer (the
 * times behind it may change between calls; they are read when the graph
 * runs).  The longest-first order is updated between the captured launches as
 * between separate calls.  Asynchronous on the batch stream like
 * cg_reactors_advance_device. */
int32_t cg_reactors_advance_graph(int32_t batch, int32_t n_steps, const double* t_out_dev);
/* Measured choice of the integrator launch (no counterpart in the reference,
 * whose integrator has no device launch to tune): for 32 < N <= 64 the
 * iteration matrix in LDS (one wavefront per SIMD) or in the reactor's global
 * factor slot (two per SIMD), with the natural or a capped number of
 * wavefronts per CU; otherwise the cap alone.  Each candidate advances the
 * whole batch from its current state to t_probe_dev[n] once; the state is
 * restored after every trial and at the end, so the call leaves the
 * integration where it was and keeps the fastest setting for later advances
 * (the first candidate is the default launch; another must be at least 3%
 * faster to replace it).
 * choice[2] (may be NULL): matrix storage as cg_reactors_set_matrix_storage,
 * waves per CU (0 = as many as fit).  times_ms[max_candidates] (may be NULL):
 * each candidate's time, -1 for a candidate identical to an earlier one.
 * Returns the number of candidates. */
int32_t cg_reactors_autotune(int32_t batch, const double* t_probe_dev, int32_t* choice, double* times_ms,
                             int32_t max_candidates);
/* State y = [m, T, Y_1..Y_K] per reactor, AoS n x (K+2)
 * (ReactorNet::getState, src/zeroD/ReactorNet.cpp:754-759). */
int32_t cg_reactors_get_state(int32_t batch, double* y);
/* The same states written to device memory (n x (K+2) doubles), ordered on
 * the batch's stream after the work queued before it, without host
 * synchronisation (no counterpart in the reference: for callers keeping
 * trajectories on the device). */
int32_t cg_reactors_get_state_device(int32_t batch, double* y_dev);
/* Net production rates [kmol/m^3/s] at every reactor's current state into
 * device memory, species-major wdot[k*n + r] as cg_rates_batch writes them
 * (the Kinetics::getNetProductionRates call the reference's kinetics1 sample
 * makes after each ReactorNet::advance, samples/cxx/kinetics1/kinetics1.cpp:
 * 42-60); stream-ordered on the batch stream. */
int32_t cg_reactors_net_production_rates_device(int32_t batch, double* wdot_dev);
/* Integrator counters per reactor, n x 9 int64: steps, rhs evaluations,
 * Jacobian evaluations, LU factorisations, Newton iterations, nonlinear
 * convergence failures, error-test failures, current order, extra rhs
 * evaluations for Jacobians (CVodesIntegrator::solverStats,
 * src/numerics/CVodesIntegrator.cpp:572-621). */
int32_t cg_reactors_stats(int32_t batch, int64_t* stats);
/* Integrator control block per reactor, n x 64 doubles: [0] internal time
 * t_n, [1] current step h, [2] next step, [28] order q, [32] steps, [37] user
 * time, [38..43] rhs/jacobian/LU/Newton/failure counters, [44] status, [45]
 * the pressure of a constant-pressure reactor, the volume (1 m^3) of a
 * constant-volume one (CVodesIntegrator::getCurrentStepSize / getLastOrder and
 * the solverStats fields, src/numerics/CVodesIntegrator.cpp:572-621). */
int32_t cg_reactors_get_control(int32_t batch, double* ctl);
/* Status per reactor: 0 ok, negative = integration failure code, as the
 * CVODES flags the reference reports (CVodesIntegrator::integrate,
 * src/numerics/CVodesIntegrator.cpp:491-512): -3 repeated Newton
 * convergence failures, -4 repeated error-test failures, -5 a non-finite
 * right-hand side at the first evaluation (FuncEval's fatal error,
 * src/numerics/FuncEval.cpp:13-43), -1 / -2 / -6 no usable initial step,
 * -10 tolerances too small for the state, -20 an output time before the
 * current time, -22 max_steps internal steps taken before reaching the
 * output time.  A reactor with a negative status is left untouched by later
 * advances. */
int32_t cg_reactors_status(int32_t batch, int32_t* status);
/* Forward sensitivity of every reactor's state to the rate multiplier of
 * reaction rxn (mechanism order), one parameter per call for all reactors of
 * the batch (Reactor::addSensitivityReaction, src/zeroD/Reactor.cpp:278-289;
 * parameter value and scale 1).  Allowed before the first advance only
 * (ReactorNet::registerSensitivityParameter, src/zeroD/ReactorNet.cpp:832-
 * 840), FP64 batches only, at most 32 parameters.  The sensitivities are
 * integrated with the state by CVODES' staggered corrector, excluded from the
 * error test (CVodesIntegrator::sensInit, src/numerics/CVodesIntegrator.cpp:
 * 263-279), right-hand sides by centred difference quotients.  Returns the
 * parameter's index. */
int32_t cg_reactors_add_sensitivity_reaction(int32_t batch, int32_t rxn);
/* Sensitivity tolerances (ReactorNet::setSensitivityTolerances,
 * src/zeroD/ReactorNet.cpp:247-256; defaults rtol 1e-4, atol 1e-6). */
int32_t cg_reactors_set_sensitivity_tolerances(int32_t batch, double rtol, double atol);
/* Sensitivities at the last output time, n x (K+2) x n_params doubles in the
 * state order of cg_reactors_get_state: dy_i/dp_j, or with normalized != 0
 * (dy_i/dp_j) / y_i with 1e-300 for a zero y_i, as ReactorNet::sensitivity
 * returns them (src/zeroD/ReactorNet.cpp:676-688); zero before the first
 * step (CVodesIntegrator::sensitivity, src/numerics/CVodesIntegrator.cpp:
 * 623-644). */
int32_t cg_reactors_get_sensitivities(int32_t batch, int32_t normalized, double* S);
/* Options: max internal steps per advance call (default 20000,
 * CVodesIntegrator.h:120-132) and the persistent grid size in wavefronts per
 * CU (0 = from occupancy). */
int32_t cg_reactors_set_options(int32_t batch, int32_t max_steps, int32_t waves_per_cu);
/* Work-queue schedule (no counterpart in the reference, whose reactors run one
 * per thread): 1 (default) = each advance call takes reactors in decreasing
 * order of the internal steps they needed in the previous call; 0 = index
 * order.  Results do not depend on it. */
int32_t cg_reactors_set_schedule(int32_t batch, int32_t longest_first);
/* Where the iteration matrix lives for 32 < N <= 64 (no reference
 * counterpart): 0 default (= 1), 1 LDS, 2 the reactor's slot in global
 * memory; factorisation and solves run in registers either way.  Takes effect
 * at the next advance; results do not depend on it. */
int32_t cg_reactors_set_matrix_storage(int32_t batch, int32_t storage);
/* Duration [ms] of the last integrator kernel of this batch, from HIP events
 * on the batch stream. */
double cg_reactors_last_kernel_ms(int32_t batch);
/* Duration [ms] of the last batched-rates kernel on `mech`. */
double cg_rates_last_kernel_ms(int32_t mech);

/* ---- learned surrogate (no counterpart in the reference) ---------------------
 * An MLP (2+K inputs: T, ln P, ln max(Y_k, 1e-30); `depth` GELU layers of
 * `hidden` units; 1+K outputs: the increments of T and Y_k over one step dt)
 * with BF16 weights/activations and FP32 accumulation on the matrix cores,
 * trained with Adam on state pairs (y0 -> y1 after dt) that the FP64
 * integrator produced.  Rows of y0/y1 are reactor states [m, T, Y_0..Y_{K-1}]
 * (device arrays, n x (K+2)); P per row.  It serves callers that advance many
 * reactors by the same short step and accept a guarded approximation. */
int32_t cg_surrogate_create(int32_t mech, double dt, int32_t hidden, int32_t depth, uint64_t seed);
int32_t cg_surrogate_free(int32_t surrogate);
int64_t cg_surrogate_nparams(int32_t surrogate);
/* Feature/target means and deviations from (up to 2^20 of) the pairs. */
int32_t cg_surrogate_fit_normalization(int32_t surrogate, int64_t n, const double* y0, const double* P,
                                       const double* y1);
/* One epoch of minibatch Adam (learning rate lr) over the pairs; mean loss out. */
int32_t cg_surrogate_train_epoch(int32_t surrogate, int64_t n, const double* y0, const double* P,
                                 const double* y1, int32_t batch, double lr, double* loss);
/* Loss and parameter gradient of the pairs as one batch, no update (grad:
 * host array of cg_surrogate_nparams floats, may be NULL). */
int32_t cg_surrogate_loss_grad(int32_t surrogate, int64_t n, const double* y0, const double* P,
                               const double* y1, double* loss, float* grad);
/* y1 = y0 + net(y0) (device arrays); m carried over. */
int32_t cg_surrogate_predict(int32_t surrogate, int64_t n, const double* y0, const double* P, double* y1);
/* cg_surrogate_predict without any allocation (no counterpart in the
 * reference: for callers that cannot allocate device memory -- code under
 * HIP graph capture, serving loops).  Same results as cg_surrogate_predict;
 * asynchronous on `stream` (NULL: the default stream), no synchronisation.
 * Intermediate activations live in one fixed 16 MiB workspace per GPU, part
 * of the library and held until the process ends; rows pass through it in
 * chunks (11,904 rows for a 4 x 256 gri30 network).  Every call on a GPU
 * shares that GPU's workspace, so calls that overlap on one GPU (different
 * streams or threads) race.  Rejects a network too wide for 128 rows per
 * pass (per row: 2 (K+2 padded to 64) + 4 hidden + 4 (K+1 padded to 64)
 * bytes). */
int32_t cg_surrogate_predict_noalloc(int32_t surrogate, int64_t n, const double* y0, const double* P, double* y1,
                                     void* stream);
/* Parameters (host arrays): FP32 weights and biases layer by layer, weights
 * row-major [out][in] with inputs/outputs padded to multiples of 64;
 * normalisation = feature means, feature deviations, target means, target
 * deviations, target scales (padded lengths; a target is asinh(increment /
 * scale), standardised). */
int32_t cg_surrogate_get_params(int32_t surrogate, float* params);
int32_t cg_surrogate_set_params(int32_t surrogate, const float* params);
int32_t cg_surrogate_get_normalization(int32_t surrogate, double* norm);
int32_t cg_surrogate_set_normalization(int32_t surrogate, const double* norm);
/* cg_reactors_advance with one guarded surrogate step (no counterpart in the
 * reference; its path is always the integrator): every reactor whose output
 * time is exactly one surrogate step dt ahead takes the network's prediction
 * if it passes the FP64 guard -- physical range, element conservation, the
 * regime gate (T > gate_T K and the H + O + OH pool being destroyed) and the
 * step-error estimate from the trapezoidal defect of the reactor ODE
 * (|r_T| <= tol_T K; |r_k| / (1 + dt/2 lambda_k) <= tol_Y for species, with
 * lambda_k the species' loss frequency) -- and then restarts from it, as
 * after CVodeReInit.  Every other reactor (rejected, or a different output
 * span) is advanced by the BDF integrator in the same call.
 * _device: t_out in device memory, stream-ordered, no host synchronisation
 * and no allocation after the first call on a batch.  Counters accumulate on
 * the device: cg_reactors_surrogate_counts returns [eligible reactor-steps,
 * kept, integrated] (reset = 1 zeroes them).  The host form returns the
 * counts of its own call in counts[3] (may be NULL) and an error if any
 * reactor failed. */
int32_t cg_reactors_advance_surrogate_device(int32_t batch, int32_t surrogate, const double* t_out_dev, double tol_T,
                                             double tol_Y, double gate_T);
int32_t cg_reactors_surrogate_counts(int32_t batch, int64_t* counts, int32_t reset);
int32_t cg_reactors_advance_surrogate(int32_t batch, int32_t surrogate, const double* t_out, double tol_T,
                                      double tol_Y, double gate_T, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif
