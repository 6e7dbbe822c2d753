/*
 * pdgpu — MI355X (gfx950) peridynamic force path behind a plain C ABI.
 *
 * The linear peridynamic solid ("Elastic") material with critical-stretch bond
 * damage, explicit velocity-Verlet integration, the matrix-free tangent of the
 * quasi-static solve and its preconditioners, all computed by hand-written HIP
 * kernels. Every array crossing this boundary is a host pointer in Peridigm's
 * own layout:
 *   - node vectors: FP64, xyz-interleaved, length 3*N, owned points 0..N-1;
 *   - node scalars: FP64, length N;
 *   - the neighbourhood list: int32 [n_0, j.., n_1, j.., ...] of local ids;
 *   - bond fields (bond damage): FP64, length B, in neighbourhood-list order.
 * Device mirrors are owned by the context. Every call returns 0 on success or
 * a negative PDG_E* code; pdg_last_error() gives the message. Calls on one
 * context must be serialised; use one context per host thread.
 *
 * Reference interfaces replaced (paths relative to the Peridigm source root):
 *   Material::initialize / computeForce   src/materials/Peridigm_Material.hpp:120-141
 *   ElasticMaterial::initialize/computeForce src/materials/Peridigm_ElasticMaterial.cpp:124-172
 *   DamageModel::computeDamage            src/damage/Peridigm_DamageModel.hpp:106-111
 *   CriticalStretchDamageModel            src/damage/Peridigm_CriticalStretchDamageModel.cpp:83-184
 *   NeighborhoodData layout               src/core/Peridigm_NeighborhoodData.hpp:56-148
 *   Peridigm::executeExplicit (Verlet)    src/core/Peridigm.cpp:1290-1593
 *   ComputeCriticalTimeStep               src/core/Peridigm_CriticalTimeStep.cpp:54-126
 *   QuickGrid / radius search / filters   src/io/mesh_input/quick_grid/QuickGrid.cxx:510-647,
 *                                         src/io/Peridigm_ZoltanSearchTree.cpp:214-247,
 *                                         src/io/pdneigh/BondFilter.cxx:59-153
 *   QS residual / 3x3 block preconditioner src/core/Peridigm.cpp:4176-4317, :1603-1647
 */
#ifndef PDGPU_H
#define PDGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PDG_OK 0
#define PDG_E_ARG -1       /* invalid argument or call order */
#define PDG_E_HIP -2       /* HIP runtime error */
#define PDG_E_UNSUPPORTED -3 /* input outside the supported model */
#define PDG_E_NONFINITE -4 /* NaN/Inf in a force evaluation */
#define PDG_E_NOCONV -5    /* iterative solve did not converge */

/* material models: "Elastic" (the linear peridynamic solid,
 * src/materials/Peridigm_ElasticMaterial.cpp) and "Elastic Bond Based"
 * (src/materials/Peridigm_ElasticBondBasedMaterial.cpp) */
#define PDG_MATERIAL_ELASTIC 0
#define PDG_MATERIAL_ELASTIC_BOND_BASED 1
#define PDG_MATERIAL_LINEAR_LPS 2

/* influence functions (src/core/Peridigm_InfluenceFunction.hpp:69-88) */
#define PDG_INFLUENCE_ONE 0
#define PDG_INFLUENCE_PARABOLIC_DECAY 1
#define PDG_INFLUENCE_GAUSSIAN 2

typedef struct pdg_ctx pdg_ctx;

/* Context on HIP device `device` (one stream, default priority). */
int pdg_ctx_create(int device, pdg_ctx** out);
int pdg_ctx_destroy(pdg_ctx* ctx);
const char* pdg_last_error(const pdg_ctx* ctx);

/* Lanes per node used by the bond kernels (4, 8, 16, 32 or 64; default 4;
 * the untiled kernels use 16 when 4 or 8 is set; the tiled kernels use at
 * least 512 / tile size, so one workgroup round covers a tile). */
int pdg_set_lanes_per_node(pdg_ctx* ctx, int lanes);

/* Nodes per LDS tile: 32, 64 or 128, or 0 (default) to choose from the mesh
 * at pdg_set_constants: 128, halved down to 32 while the mesh gives fewer
 * than two workgroups per CU. Takes effect at the next pdg_set_constants.
 * No reference counterpart (a tuning control of this implementation). */
int pdg_set_tile_nodes(pdg_ctx* ctx, int nodes);
/* Tile size and lanes per node the bond kernels run with (tile size 0 when
 * the untiled kernels run). No reference counterpart. */
int pdg_tile_config(const pdg_ctx* ctx, int* tileNodes, int* lanes);

/* Bond-kernel variant: 0 (default) LDS-staged Hilbert tiles when every tile
 * halo fits in LDS, else the untiled kernels; 1 forces the untiled kernels
 * (one 64-byte neighbour-record gather per bond); 2 is variant 0 with plain
 * (branching, unpipelined) bond loops, same arithmetic. */
int pdg_set_kernel_variant(pdg_ctx* ctx, int variant);
/* Number of Hilbert tiles (0: tiling unavailable) and total halo entries. */
int pdg_get_tiling(const pdg_ctx* ctx, int* numTiles, long long* haloTotal);
/* LDS bank-conflict figures of the tiled bond loops: the number of (16-lane
 * group, iteration) pairs and the sum over them of the largest number of
 * distinct halo slots sharing one residue modulo 16 (1 per pair when
 * conflict-free), for the rows in neighbour order and in the order chosen at
 * pdg_set_constants. No reference counterpart (a diagnostic). */
int pdg_lds_schedule(const pdg_ctx* ctx, long long* groups, long long* costBefore,
                     long long* costAfter);

/* ---- topology and constants -------------------------------------------
 * pdg_set_topology: the block's NeighborhoodData (numOwnedPoints, ownedIDs,
 * neighborhoodList of listSize ints). All neighbour ids must be owned points
 * (single device, no ghosts) and the list must be symmetric (j in N(i) iff
 * i in N(j)): forces are gathered from both bond ends without atomics.
 * pdg_set_constants: Model_Coordinates x (3N), Volume (N), material and
 * damage constants (s_crit <= 0: no damage model). Builds the Hilbert node
 * order, the CSR bond arrays, the bond-damage bit masks and the device state.
 */
int pdg_set_topology(pdg_ctx* ctx, int numOwnedPoints, const int* ownedIDs,
                     const int* neighborhoodList, long long listSize);
int pdg_set_constants(pdg_ctx* ctx, const double* x, const double* volume,
                      double bulkModulus, double shearModulus, double horizon,
                      int influenceKind, double criticalStretch);
/* Critical stretch per node (N) for multi-specimen runs; the damage model
 * must have been enabled by a positive criticalStretch in pdg_set_constants.
 * A bond i->j uses node i's value. */
int pdg_set_critical_stretch(pdg_ctx* ctx, const double* criticalStretch);
/* Material model for the next pdg_set_constants (default
 * PDG_MATERIAL_ELASTIC). The bond-based model (pair force
 * t = 0.5 (1 - d) c s with c = 18K/(pi delta^4),
 * src/materials/elastic_bond_based.cxx:55-130) takes the bulk modulus only;
 * planar = 1 (plane strain) or 2 (plane stress) uses c = 12K/(pi h delta^3)
 * with thickness `height` (:80-84). Its dilatation output is zero and the
 * quasi-static calls reject it. PDG_MATERIAL_LINEAR_LPS is the linearized
 * LPS of "Linear LPS Partial Volume" without partial volumes
 * (src/materials/Peridigm_LinearLPSPVMaterial.cpp:229-394: theta and force
 * linear in u, the force along the reference bond); pdg_compute_force and
 * the quasi-static calls evaluate it in FP64, the explicit runs reject it. */
int pdg_set_material_model(pdg_ctx* ctx, int model, int planar, double height);

/* Critical stretch used by every damage-enabled node from the next damage
 * evaluation on (> 0), or back to the per-node values (<= 0): the current
 * value of "Time Dependent Critical Stretch"
 * (src/damage/Peridigm_UserDefinedTimeDependentCriticalStretchDamageModel.cpp:93-131),
 * which the caller evaluates at each step's time. */
int pdg_set_critical_stretch_value(pdg_ctx* ctx, double criticalStretch);
/* The same as a function of time for pdg_verlet_run: at each damage
 * evaluation (time t0 + k dt) the critical stretch is the function's value,
 * linear between the n points (times increasing), constant beyond them; a
 * point time gives its value exactly (src/core/Peridigm.cpp:1447-1463
 * evaluates it at every step). n = 0 removes it. */
int pdg_set_critical_stretch_function(pdg_ctx* ctx, int n, const double* times,
                                      const double* values);

/* Density per node (N) for the explicit integrator. */
int pdg_set_density(pdg_ctx* ctx, const double* density);

/* Number of bonds B and nodes N after pdg_set_topology. */
int pdg_get_sizes(const pdg_ctx* ctx, int* numNodes, long long* numBonds);

/* ---- drop-in material / damage calls ----------------------------------- */
/* ElasticMaterial::initialize: Weighted_Volume (N), FP64, reference order of
 * summation (bit-identical to the CPU loop for the "One" influence function). */
int pdg_initialize(pdg_ctx* ctx, double* weightedVolume);

/* CriticalStretchDamageModel::computeDamage: Coordinates NP1 y (3N),
 * Bond_Damage N (B) -> Bond_Damage NP1 (B) and Damage (N). nearCritical
 * (optional) receives the number of intact bonds whose FP32 stretch lies
 * within 1e-6*s_crit of s_crit. */
int pdg_compute_damage(pdg_ctx* ctx, const double* y, const double* bondDamageN,
                       double* bondDamageNP1, double* damage, long long* nearCritical);

/* ElasticMaterial::computeForce: y (3N), Bond_Damage NP1 (B, values 0 or 1,
 * symmetric) -> Dilatation (N) and Force_Density (3N). */
int pdg_compute_force(pdg_ctx* ctx, double dt, const double* y, const double* bondDamageNP1,
                      double* dilatation, double* forceDensity);

/* ---- explicit dynamics ------------------------------------------------- */
/* Prescribed-displacement boundary conditions u_c(t) = c1*t + c0 on a node
 * set, applied to the velocity after the half kick as DirichletIncrementBC
 * does (src/core/Peridigm_BoundaryCondition.cpp:202-241). */
typedef struct {
  const int* nodes; /* node ids */
  int numNodes;
  int coord;        /* 0,1,2 */
  double c0, c1;
} pdg_bc;

/* State upload/download (reference layout). u, v, a: 3N; bondDamage: B. */
int pdg_state_upload(pdg_ctx* ctx, const double* u, const double* v, const double* a,
                     const double* bondDamage);
int pdg_state_download(pdg_ctx* ctx, double* u, double* v, double* a, double* bondDamage,
                       double* damage, double* dilatation, double* forceDensity);
int pdg_set_bcs(pdg_ctx* ctx, const pdg_bc* bcs, int numBcs);

/* Peridigm::executeExplicit: if `initialEval`, evaluates damage and force in
 * the current configuration first (:1388-1424); then runs nsteps Verlet steps
 * from time t0 with step dt on the device-resident state. */
int pdg_verlet_run(pdg_ctx* ctx, int nsteps, double dt, double t0, int initialEval);

/* Same run as pdg_verlet_run in ONE persistent cooperative launch (grid-wide
 * barriers between the two passes of every step), for batched multi-specimen
 * sweeps: the specimens are concatenated into one node/bond space. Needs the
 * tiled kernels and 4, 8 or 16 lanes per node. *gridSize receives the workgroups used. */
int pdg_verlet_run_persistent(pdg_ctx* ctx, int nsteps, double dt, double t0, int initialEval,
                              int* gridSize);

/* ---- short-range contact ------------------------------------------------
 * src/contact/Peridigm_ShortRangeForceContactModel.cpp:98-258 and
 * src/core/Peridigm_ContactManager.cpp:700-750. Every node takes part, with
 * the context's horizon. */
/* ContactManager (Search Radius, Search Frequency) with one
 * ShortRangeForceContactModel (Contact Radius, Spring Constant, Friction
 * Coefficient); contactRadius <= 0 removes the model. Set after the
 * constants; pdg_set_topology and pdg_set_constants remove it. With a model
 * set, pdg_verlet_run searches at its initial evaluation and at every step
 * whose number (counted from the last initial evaluation) is a multiple of
 * searchFrequency, in the previous step's configuration (src/core/
 * Peridigm.cpp:1468-1472), and adds the contact force density to the
 * internal one before the acceleration (:1559-1565); Force_Density then
 * holds the sum, as in the reference. */
int pdg_set_contact(pdg_ctx* ctx, double searchRadius, int searchFrequency,
                    double contactRadius, double springConstant,
                    double frictionCoefficient);
/* ContactManager::contactSearch: for each node of the configuration y (3N),
 * the nodes within the search radius (squared distance <= R^2), itself and
 * its bonded neighbours (broken or not) excluded; flat list
 * [n, j1..jn]..., neighbours ascending. *listSize: capacity in, size out
 * (contactList may be null to query the size). */
int pdg_contact_search(pdg_ctx* ctx, const double* y, int* contactList, long long* listSize);
/* ShortRangeForceContactModel::computeForce: Contact_Force_Density (3N) of
 * configuration y and velocity v (3N) for a flat contact list (any list: each
 * listed pair acts on both of its nodes, as in the reference). */
int pdg_contact_compute_force(pdg_ctx* ctx, double dt, const double* y, const double* v,
                              const int* contactList, long long listSize,
                              double* contactForceDensity);
/* Contact_Force_Density (3N) of the last pdg_verlet_run evaluation. */
int pdg_contact_force_density(pdg_ctx* ctx, double* contactForceDensity);

/* Stable time step (ComputeCriticalTimeStep, constant horizon,
 * src/core/Peridigm_CriticalTimeStep.cpp:54-126) for density rho, computed on
 * the device: min over nodes of sqrt(2 rho / sum_j V_j 18K/(pi delta^4) / |xi|). */
int pdg_critical_time_step(pdg_ctx* ctx, double density, double* dtOut);
/* The same time step without allocating: written to *dtOutDevice (device
 * memory, 1 double) before the call returns, with the per-workgroup partial
 * minima in a fixed device array of the library (128 KiB per device, kept
 * until the process ends). Blocks of at most 4,194,304 nodes
 * (PDG_E_UNSUPPORTED beyond). Calls that overlap on one device race on that
 * array. No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). */
int pdg_critical_time_step_noalloc(pdg_ctx* ctx, double density, double* dtOutDevice);

/* Per-kernel timing of the last pdg_verlet_run (ms, HIP events on the
 * context's stream): total, damage+dilatation pass, force+Verlet pass. */
int pdg_last_timing(const pdg_ctx* ctx, double* totalMs, double* pass1Ms, double* pass2Ms);

/* 1: pdg_verlet_run records per-pass HIP events and synchronises after each
 * step so pdg_last_timing reports per-pass times; 0 (default): no per-step
 * synchronisation, only the total time is recorded. */
int pdg_set_per_pass_timing(pdg_ctx* ctx, int on);

/* Near-critical count of the last pdg_verlet_run / pdg_verlet_run_persistent
 * (or of the last pdg_compute_damage that asked for one): the number of
 * (bond, damage evaluation) pairs in which an intact bond's FP32 stretch lay
 * within 1e-6*s_crit of s_crit, summed over every evaluation of the run. Its
 * break decision may differ from an FP64 evaluation's. No reference
 * counterpart (src/damage/Peridigm_CriticalStretchDamageModel.cpp:147-161
 * makes the FP64 decision these bonds approximate). */
int pdg_near_critical_count(const pdg_ctx* ctx, long long* count);

/* ---- quasi-static (src/core/Peridigm.cpp:2470-3043, :4176-4317) --------
 * FP64 throughout. Vectors are 3N, reference order. Kinematic (prescribed)
 * DOFs are node*3+coord ids; their residual rows are zero and their tangent
 * rows are identity. */
#define PDG_PRECOND_NONE 0
#define PDG_PRECOND_BLOCK_JACOBI 1 /* 3x3 diagonal blocks, :1603-1647 */
#define PDG_PRECOND_LEARNED 2      /* message-passing network, flexible CG */
#define PDG_PRECOND_TWO_LEVEL 3    /* block Jacobi + Galerkin coarse correction on rigid-body
                                      modes of node aggregates (no reference counterpart) */
int pdg_qs_set_fixed(pdg_ctx* ctx, const int* dofs, int numDofs);
/* r = f(u) V with kinematic rows zeroed; *norm = ||r||_2 + 20 ||r||_inf (:4316). */
int pdg_qs_residual(pdg_ctx* ctx, const double* u, double* residual, double* norm);
/* K w = -V (df/du) w on free DOFs (forward-mode directional derivative of
 * the force), w on kinematic DOFs. */
int pdg_qs_tangent_apply(pdg_ctx* ctx, const double* u, const double* w, double* Kw);
/* FP64 internal force density f(u) on every DOF (3N, kinematic ones
 * included) and dilatation (N, optional) at displacement u, from the
 * quasi-static kernels (src/materials/elastic.cxx:55-146,
 * src/materials/material_utilities.cxx:235-288 with y = x + u). */
int pdg_qs_force_dilatation(pdg_ctx* ctx, const double* u, double* force, double* dilatation);
/* Inverses of the 3x3 diagonal blocks of K at u (9N, row-major per node). */
int pdg_qs_block_diagonal(pdg_ctx* ctx, const double* u, double* blocksInv);
/* Solve K(u) x = rhs (x: initial guess in, solution out) to ||r||/||rhs|| <=
 * relTol with CG (block Jacobi) or flexible CG (learned preconditioner). */
int pdg_qs_solve(pdg_ctx* ctx, const double* u, const double* rhs, double* x, int precond,
                 double relTol, int maxIter, int* iters, double* relRes);
/* Newton at fixed prescribed values (u holds them on kinematic DOFs): until
 * ||r||_2 + 20||r||_inf <= tol * ||reactions||_2, the reactions (f V on the
 * kinematic DOFs) taken once at the entry state, as the reference takes them
 * once per load step (:2700-2732); each linear solve to linTol. */
int pdg_qs_newton(pdg_ctx* ctx, double* u, double tol, int maxNewton, int precond,
                  double linTol, int maxLin, int* newtonIters, int* linIters,
                  double* residualNorm);
/* 1: the Newton tolerance is absolute (multiplier 1), the reference's
 * "Absolute Tolerance" / useAbsoluteTolerance (src/core/Peridigm.cpp:2514-2517,
 * :2716-2732); 0 (default): scaled by the reaction norm. */
int pdg_qs_set_absolute_tolerance(pdg_ctx* ctx, int on);
/* The tolerance multiplier (reaction norm, or 1) of the last pdg_qs_newton. */
int pdg_qs_tolerance_multiplier(const pdg_ctx* ctx, double* multiplier);
/* Controls of the quasi-static load-stepping driver, named after the
 * reference's "Solver"/"QuasiStatic" parameters (src/core/Peridigm.cpp:2483-2630);
 * pdg_qs_params_default fills the reference's defaults. */
typedef struct {
  int maxSolverIterations;     /* "Maximum Solver Iterations" (10) */
  double tolerance;            /* "Relative Tolerance" (1e-6), or the absolute one */
  int absoluteTolerance;       /* 1: `tolerance` is "Absolute Tolerance" (no reaction scaling) */
  int disableHeuristics;       /* "Disable Heuristics": no predictor, line search or damping */
  double dampedNewtonDiagonalScaleFactor; /* "Damped Newton Diagonal Scale Factor" (1.0001) */
  double dampedNewtonDiagonalShiftFactor; /* "Damped Newton Diagonal Shift Factor" (1e-5) */
  int adaptiveLoadStepping;    /* the "Adaptive Load Stepping" sublist is present */
  int maxLoadStepReductionsInOneStep; /* "Maximum load step reductions in one step" */
  int maxTotalLoadStepReductions;     /* "Maximum total load step reductions" */
  int reduceAllRemainingLoadSteps;    /* "Reduce all remaining load steps" */
  int precond;                 /* PDG_PRECOND_* of the linear solves (the reference: none) */
  double linearTolerance;      /* "Belos Relative Tolerance" (1e-4) */
  int maxLinearIterations;     /* "Belos Maximum Iterations" (0: 3N) */
} pdg_qs_params;

/* One accepted load step of pdg_qs_run. */
typedef struct {
  double timePrevious, timeCurrent;
  int newtonIterations;        /* solver iterations, the predictor step included */
  int linearIterations;
  double residualNorm;         /* ||r||_2 + 20 ||r||_inf at the end of the step */
  double toleranceMultiplier;  /* reaction norm at the step's entry, or 1 */
  int status;                  /* 0 converged; 1 accepted below 100x the tolerance */
  int dampedNewton;            /* 1 if the step switched to damped Newton */
  double lastAlpha;            /* line-search step of the last iteration */
} pdg_qs_step;

int pdg_qs_params_default(pdg_qs_params* params);

/* Peridigm::executeQuasiStatic (src/core/Peridigm.cpp:2470-3043) from the
 * displacement u (3N, in: initial, out: final): load steps at times[0..n-1];
 * prescribed displacements u_c(t) = c1 t + c0 on the bcs' node sets applied
 * as increments per load step (DirichletIncrementBC, :2674-2686); the
 * reaction-scaled Newton tolerance (:2716-2732); the previous step's velocity
 * as the first iterate (:2764-2769); damped Newton after 50 iterations or on a
 * failed linear solve, retried without preconditioner (:2772-2822); the
 * candidate-alpha line search (:3364-3618); acceptance below 100x the
 * tolerance, and adaptive halving of failed load steps (:2880-2934). The
 * tangent is applied matrix-free at the current state. Returns
 * PDG_E_NOCONV when the run aborts (u then holds the last accepted step). */
int pdg_qs_run(pdg_ctx* ctx, const pdg_bc* bcs, int numBcs, const double* times, int numTimes,
               const pdg_qs_params* params, double* u, pdg_qs_step* steps, int maxSteps,
               int* numSteps);
/* The same run, also writing the displacement of every accepted load step
 * (maxSteps x 3N, reference order) to uSteps, where the reference writes
 * its output frames (:2984-2990). */
int pdg_qs_run_record(pdg_ctx* ctx, const pdg_bc* bcs, int numBcs, const double* times,
                      int numTimes, const pdg_qs_params* params, double* u, pdg_qs_step* steps,
                      int maxSteps, int* numSteps, double* uSteps);

/* ---- thermal strains and partial stress (elastic LPS) --------------------
 * ElasticMaterial with "Thermal Expansion Coefficient"
 * (src/materials/Peridigm_ElasticMaterial.cpp:78-81): every bond's extension
 * loses alpha dT_i zeta, dT_i the Temperature_Change of the bond's owner
 * (src/materials/elastic.cxx:112-113, material_utilities.cxx:281-282).
 * Evaluated in FP64 by pdg_compute_force and the quasi-static calls; the
 * explicit runs reject it. coefficient 0 turns it off. */
int pdg_set_thermal_expansion(pdg_ctx* ctx, double coefficient);
/* Temperature_Change per node as dT(t) = dT1 t + dT0 (N each, reference
 * order; dT1 may be null; both null clears the field). pdg_qs_run evaluates
 * it at every load step's time, as a prescribed temperature's change from
 * its initial value (src/core/Peridigm_BoundaryAndInitialConditionManager.cpp:
 * 148-159); the other calls at the time set by pdg_set_temperature_time or
 * the last load step (0 at first). */
int pdg_set_temperature_change(pdg_ctx* ctx, const double* dT0, const double* dT1);
int pdg_set_temperature_time(pdg_ctx* ctx, double t);
/* Partial_Stress (9N, per node row-major: force component x reference bond
 * component) of ElasticMaterial::computeForce with "Compute Partial Stress"
 * (src/materials/elastic.cxx:132-142): sum over the node's own bonds of
 * f_ij (x_j - x_i)^T V_j at configuration y with bond damage (B, may be
 * null), thermal strains included; FP64. */
int pdg_compute_partial_stress(pdg_ctx* ctx, const double* y, const double* bondDamage,
                               double* partialStress);
/* 1: linear solves replay each CG iteration's two fixed launch sequences as
 * HIP graphs, captured once per context; 0 (default): every kernel is
 * launched directly. Same arithmetic, bitwise. */
int pdg_qs_set_graphs(pdg_ctx* ctx, int enable);
/* Number of learned-preconditioner solves that stalled (residual not halved
 * in 50 iterations) and were finished by block-Jacobi CG, the fallback. */
int pdg_qs_fallback_count(pdg_ctx* ctx, int* count);

/* ---- learned preconditioner (no counterpart in the reference) -----------
 * A 3-layer message-passing network (width 32, BF16 MFMA, FP32 accumulate)
 * over the bond graph; used by PDG_PRECOND_LEARNED inside flexible CG. */
int pdg_gnn_init(pdg_ctx* ctx, unsigned seed);
int pdg_gnn_num_params(void);
int pdg_gnn_get_params(pdg_ctx* ctx, float* params);
int pdg_gnn_set_params(pdg_ctx* ctx, const float* params);
/* Forward on node features (N x 8) -> out (N x 3); with gout (N x 3) also the
 * gradient of sum(gout * out) w.r.t. the parameters. */
int pdg_gnn_forward_backward(pdg_ctx* ctx, const float* features, const float* gout,
                             float* out, float* dparams);
/* Adam training on (r, e = K^-1 r) pairs (nsamples x 3N each) at tangent u. */
int pdg_gnn_train(pdg_ctx* ctx, const double* u, int nsamples, const double* r,
                  const double* e, int epochs, float lr, unsigned seed, float* lastLoss);

/* ---- neighbour search (integer results, bit-exact predicates) ----------- */
/* QuickGrid tensor-product lattice: node coordinates (3N), volumes (N) and
 * the flattened neighbourhood list (ascending global id, self excluded).
 * Call with nlist == NULL to get *listSize, then again with capacity. */
int pdg_quickgrid(int nx, int ny, int nz, const double origin[3], const double length[3],
                  double horizon, int spherical, double* x, double* volume, int* nlist,
                  long long* listSize);
/* Fixed-radius search |x_j - x_i|^2 <= R^2 (rows sorted by id, self excluded),
 * then removal of bonds cut by any of the finite planes (11 doubles each:
 * normal, lower-left corner, bottom unit vector, bottom length, side length). */
int pdg_radius_search(int N, const double* x, double radius, int numPlanes,
                      const double* planes, int* nlist, long long* listSize);

#ifdef __cplusplus
}
#endif
#endif /* PDGPU_H */
