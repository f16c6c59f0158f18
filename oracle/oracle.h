/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm for the batched
 * gas-kinetics + constant-pressure reactor path.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load this library; the
 * product path (libcantera_amd.so) never links or calls it.
 *
 * Restated reference functions (file:line in the Cantera sources):
 *   thermo      NasaPoly1::updateProperties  include/cantera/thermo/NasaPoly1.h:103-122
 *               NasaPoly2::updateProperties  include/cantera/thermo/NasaPoly2.h:105-111
 *               IdealGasPhase::updateThermo  src/thermo/IdealGasPhase.cpp:289-306
 *               IdealGasPhase::getStandardChemPotentials  src/thermo/IdealGasPhase.cpp:59-66
 *               Phase::setMassFractions_NoNorm / getConcentrations  src/thermo/Phase.cpp:362-372, 501-505
 *   kinetics    BulkKinetics::updateROP  src/kinetics/BulkKinetics.cpp:457-545
 *               Kinetics::getNetProductionRates  src/kinetics/Kinetics.cpp:447-457
 *               Kinetics::getRevReactionDelta  src/kinetics/Kinetics.cpp:408-418
 *               StoichManagerN (C1/C2/C3/C_AnyN)  include/cantera/kinetics/StoichManager.h:131-366, 676-744
 *               ThirdBodyCalc::update/multiply  include/cantera/kinetics/ThirdBodyCalc.h:75-91
 *               ArrheniusRate::evalFromStruct  include/cantera/kinetics/Arrhenius.h:196-198
 *               FalloffRate::evalFromStruct  include/cantera/kinetics/Falloff.h:154-176
 *               TroeRate::updateTemp / F  src/kinetics/Falloff.cpp:283-300
 *               SriRate::updateTemp / F  src/kinetics/Falloff.cpp:389-403
 *               TsangRate::updateTemp / F  src/kinetics/Falloff.cpp:480-494
 *               PlogRate::setRates / updateFromStruct / evalFromStruct
 *                 src/kinetics/PlogRate.cpp:92-123, include/cantera/kinetics/PlogRate.h:110-164
 *               ChebyshevRate::setLimits / updateFromStruct / evalFromStruct
 *                 src/kinetics/ChebyshevRate.cpp:96-112, include/cantera/kinetics/ChebyshevRate.h:133-171
 *   reactor     IdealGasConstPressureReactor::getState/updateState/eval
 *                 src/zeroD/IdealGasConstPressureReactor.cpp:16-110
 *               ReactorNet::eval  src/zeroD/ReactorNet.cpp:617-635
 *               Reactor::finiteDifferenceJacobian (FD Jacobian check)  src/zeroD/Reactor.cpp:233-276
 *   integrator  CVodesIntegrator::integrate/step semantics  src/numerics/CVodesIntegrator.cpp:479-549
 *               The BDF itself lives in SUNDIALS CVODES 7.5 (an un-vendored
 *               submodule, absent here).  It is restated from the published
 *               method: Byrne & Hindmarsh, ACM TOMS 1 (1975); Hindmarsh et al.,
 *               ACM TOMS 31 (2005); the CVODE user guide's "Mathematical
 *               considerations" chapter (fixed-leading-coefficient Nordsieck BDF,
 *               q <= 5, Newton with a dense difference-quotient Jacobian,
 *               local-error test and step/order selection).
 *
 * Pinning (see tests/): NASA7 against the CHEMKIN-II values of
 * test/thermo/nasapoly.cpp:62-78; rates against test/kinetics/rates.cpp:128-150
 * (noxNeg.yaml); Kc against the closed form of rates.cpp:99-115; Troe against
 * its exact special cases; PLOG and Chebyshev against the known answers of
 * test/kinetics/pdep.cpp:52-220 (pdep-test.yaml: pressure limits against the
 * closed-form Arrhenius expressions, CHEMKIN rates of progress, Chebyshev
 * rate constants); SRI and Tsang against their exact cases (SRI exponent 1
 * at Pr = 1 and 1/5 at Pr = 100, Tsang F = Fcent at log10 Pr = -c) and their
 * [M] derivatives against finite differences; invariants (mass and element conservation, detailed
 * balance); the integrator against test_problems/cxx_samples/kin1_blessed.csv
 * and opmenmp_ignition_blessed.txt and against scipy Radau on the same RHS;
 * the analytic Jacobian against a finite-difference Jacobian of the RHS.
 */
#ifndef CG_ORACLE_H
#define CG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mech or_mech;
typedef struct or_reactor or_reactor;

/* mechanism ------------------------------------------------------------ */
or_mech* or_mech_load(const char* path);
void or_mech_free(or_mech* m);
int or_mech_nspecies(const or_mech* m);
int or_mech_nreactions(const or_mech* m);
int or_mech_nelements(const or_mech* m);
void or_mech_mw(const or_mech* m, double* mw);                 /* K */
void or_mech_composition(const or_mech* m, double* comp);      /* K x E, row-major */
int or_mech_species_index(const or_mech* m, const char* name);
const char* or_last_error(void);

/* thermo: cp/R, h/RT, s/R for every species at T (NASA7) */
void or_thermo(const or_mech* m, double T, double* cp_R, double* h_RT, double* s_R);

/* single state, AoS: Y has K entries; outputs as named (any may be NULL) */
int or_rates(const or_mech* m, double T, double P, const double* Y,
             double* wdot, double* ropf, double* ropr, double* kf, double* rkcn);
/* equilibrium constants Kc (all reactions), BulkKinetics::getEquilibriumConstants */
int or_equilibrium_constants(const or_mech* m, double T, double P, const double* Y, double* Kc);

/* batch over n states, species-major SoA: Y[k*n + s], wdot[k*n + s] */
int or_rates_batch(const or_mech* m, int n, const double* T, const double* P,
                   const double* Y, double* wdot, int nthreads);

/* chemical equilibrium at fixed (h, P) (hp = 1) or (T, P) (hp = 0) of the
 * mixture (T_in, P, Y_in): ThermoPhase::equilibrate, element-potential
 * Gibbs minimisation.  Returns 0, -1 (max_iter reached), -3 (singular). */
int or_equilibrate(const or_mech* m, int hp, double T_in, double P, const double* Y_in, double* T_out,
                   double* Y_out, int max_iter, int* iters);

/* analytic reactor Jacobians (m = 1 kg) of n states, inputs as or_rates_batch,
 * J[s*N*N + r*N + c] */
int or_jacobian_batch(const or_mech* m, int n, const double* T, const double* P, const double* Y, double* J,
                      int nthreads);

/* reactor RHS: y = [m, T, Y_1..Y_K] (neq = K+2), P held fixed */
int or_igcpr_rhs(const or_mech* m, double P, const double* y, double* ydot);
/* Jacobian d(ydot)/dy, row-major neq x neq.  mode 0: finite differences
 * (increment sqrt(eps)*max(|y_j|, 1e-... floor)); mode 1: analytic (the same
 * closed form the GPU kernels use). */
/* IdealGasReactor RHS for y = [m, T, Y] at volume V (the constant V omitted). */
int or_igr_rhs(const or_mech* m, double V, const double* y, double* ydot);
int or_igcpr_jacobian(const or_mech* m, double P, const double* y, double* J, int mode);

/* integrator ------------------------------------------------------------ */
/* jac_mode: OR_JAC_DQ or OR_JAC_ANALYTIC, plus OR_CONST_VOLUME for a closed
 * rigid IdealGasReactor (V = 1 m^3, mass and V from (T0, P, Y0);
 * src/zeroD/IdealGasReactor.cpp:16-117) instead of the constant-pressure
 * reactor.  Its WRMS norms count V (CVODES's state holds it). */
enum { OR_JAC_DQ = 0, OR_JAC_ANALYTIC = 1, OR_CONST_VOLUME = 2 };
or_reactor* or_reactor_new(const or_mech* m, double T0, double P, const double* Y0,
                           double rtol, double atol, int jac_mode);
void or_reactor_free(or_reactor* r);
int or_reactor_advance(or_reactor* r, double tout);   /* 0 ok, <0 error */
int or_reactor_step(or_reactor* r);                   /* one internal step */
double or_reactor_time(const or_reactor* r);          /* user time (after advance: tout) */
void or_reactor_state(const or_reactor* r, double* y); /* neq values */
/* stats: [nsteps, nfe, nje, nsetups, nni, ncfn, netf, last_q, nfe_dq] */
void or_reactor_stats(const or_reactor* r, long* stats);
/* forward sensitivities dy/dp to reaction-rate multipliers (at most 32),
 * added before the first step; S is neq x ns, row-major */
int or_reactor_add_sensitivity(or_reactor* r, int rxn);
/* IdealGasConstPressureMoleReactor with the adaptive preconditioner and SPGMR
 * (state [T, n_1..n_K]); flags: 1 skip third-body derivatives, 2 skip falloff
 * derivatives, 4 atol for every component as given, 8 the constant-volume
 * IdealGasMoleReactor (state [T, V, n_1..n_K]) */
or_reactor* or_mole_reactor_new(const or_mech* m, double T0, double P, const double* Y0, double rtol, double atol,
                                int flags);
/* lin_iters, prec_evals, prec_solves, jt_vec_prod_evals, lin_rhs_evals, lin_conv_fails */
void or_reactor_lin_stats(const or_reactor* r, long* stats);
/* mole-reactor right-hand side and preconditioner Jacobian (state order) */
int or_mole_rhs(const or_mech* m, double P, const double* y, double* ydot, int cv);
int or_mole_jacobian(const or_mech* m, double P, const double* y, double* J, int flags);
/* n mole reactors to k dt[i], k = 1..nout (OpenMP); final states n x N */
int or_mole_reactors_run(const or_mech* m, int n, const double* T0, const double* P, const double* Y0,
                         const double* dt, int nout, double rtol, double atol, int flags, double* y_out,
                         int nthreads);
void or_reactor_set_sensitivity_tolerances(or_reactor* r, double rtol, double atol);
void or_reactor_sensitivities(const or_reactor* r, double* S);
/* every reaction's rate multiplier (Kinetics::setMultiplier) for this reactor */
void or_reactor_set_multipliers(or_reactor* r, const double* mult);

/* batch: n reactors, each advanced through nout outputs at t_k = k*dt[i],
 * k = 1..nout, states written AoS y_out[(i*(nout+1) + k)*neq + j] (k = 0 is the
 * initial state).  stats_out: n x 9.  Returns 0 or the first failing index+1
 * negated. */
int or_reactors_run(const or_mech* m, int n, const double* T0, const double* P,
                    const double* Y0 /* n x K AoS */, const double* dt, int nout,
                    double rtol, double atol, int jac_mode, double* y_out,
                    long* stats_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
