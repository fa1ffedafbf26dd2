/*
 * hoefem: MI355X (gfx950) implementation of the implicit elastodynamics step of the
 * higher-order embedded FEM reference (InteractiveComputerGraphics/higher_order_embedded_fem).
 *
 * C ABI. Setup calls take HOST arrays (mesh, quadrature, boundary conditions); hot-path
 * calls take DEVICE pointers by default, or HOST pointers after
 * hoefem_model_set_pointer_mode(model, HOEFEM_POINTERS_HOST) (every call that takes the model or a
 * handle built on it), and are ordered on the handle's HIP stream (the null stream unless
 * hoefem_*_set_stream is called). Handles are not re-entrant; different handles
 * may be used from different threads. Every function returns a hoefem_status; the
 * message of the last failure on the calling thread is hoefem_last_error().
 *
 * Conventions shared with the reference:
 *  - DOF vectors are node-major xyz (DVector of 3 * num_vertices entries).
 *  - CSR matrices use int32 row offsets / column indices (the reference's MKL_INT boundary,
 *    simulation_toolbox/src/fem/newton_cg.rs:621-658) and the pattern of
 *    CsrParAssembler::assemble_pattern (fenris/src/assembly.rs:177-249).
 *  - 3x3 matrices (F, P) are column-major; dP/dF is the 9x9 column-major matrix in the
 *    vec() ordering of fenris/src/solid/materials.rs:745-761.
 */
#ifndef HOEFEM_H
#define HOEFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hoefem_status {
    HOEFEM_OK = 0,
    HOEFEM_ERROR_INVALID_ARGUMENT = 1,
    HOEFEM_ERROR_DEVICE = 2,
    HOEFEM_ERROR_SINGULAR = 3,                     /* singular F or J (reference panics) */
    HOEFEM_ERROR_PATTERN = 4,                      /* missing CSR entry (reference panics) */
    HOEFEM_ERROR_CG_MAX_ITERATIONS = 5,            /* SolveErrorKind::MaxIterationsReached */
    HOEFEM_ERROR_CG_INDEFINITE_OPERATOR = 6,       /* SolveErrorKind::IndefiniteOperator */
    HOEFEM_ERROR_CG_INDEFINITE_PRECONDITIONER = 7, /* SolveErrorKind::IndefinitePreconditioner */
    HOEFEM_ERROR_NEWTON_MAX_ITERATIONS = 8,        /* NewtonError::MaximumIterationsReached */
    HOEFEM_ERROR_LINE_SEARCH = 9,                  /* NewtonError::LineSearchError */
    HOEFEM_ERROR_SUBSTEPPING = 10,                 /* integrate_model_with_substeps gave up */
    HOEFEM_ERROR_NOT_FINITE = 11,                  /* non-finite state (newton_cg.rs:224-225) */
    HOEFEM_ERROR_JACOBIAN = 12                     /* NewtonError::JacobianError */
} hoefem_status;

typedef enum hoefem_element_type {
    HOEFEM_TET4 = 0,
    HOEFEM_TET10 = 1,
    HOEFEM_TET20 = 2,
    HOEFEM_HEX8 = 3,
    HOEFEM_HEX20 = 4,
    HOEFEM_HEX27 = 5
} hoefem_element_type;

/* EmbeddedQuadrature (fenris/src/embedding/mod.rs:67-99): one rule shared by the interior
 * elements [0, num_interior) and one rule per interface element. A NodalModel
 * (fenris/src/model.rs:32-46) is the case num_interior == num_elements. */
typedef struct hoefem_quadrature {
    int32_t interior_num_points;
    const double* interior_weights;        /* [interior_num_points] */
    const double* interior_points;         /* [interior_num_points * 3], reference coordinates */
    const int64_t* interface_offsets;      /* [num_interface + 1], or NULL if no interface elements */
    const double* interface_weights;
    const double* interface_points;
} hoefem_quadrature;

typedef struct hoefem_model_desc {
    int32_t element_type;                  /* hoefem_element_type */
    int64_t num_vertices;
    const double* vertices;                /* [num_vertices * 3] */
    int64_t num_elements;
    const int32_t* connectivity;           /* [num_elements * nodes_per_element], reference node order */
    int64_t num_interior;                  /* interior elements first (separately colored) */
    hoefem_quadrature mass_quadrature;
    hoefem_quadrature stiffness_quadrature;
    hoefem_quadrature elliptic_quadrature;
    double mass_regularization_factor;     /* EmbeddedModel::mass_regularization_factor, 0 = off */
} hoefem_model_desc;

/* Where the vectors and CSR values passed to a model's hot-path calls live (the pointer flag of the
 * reference-side binding: a Rust caller hands over host DVector / CsrMatrix storage). */
typedef enum hoefem_pointer_mode {
    HOEFEM_POINTERS_DEVICE = 0,            /* device pointers, stream-ordered (the default) */
    HOEFEM_POINTERS_HOST = 1               /* host pointers: staged through device buffers; calls return
                                              with their outputs written */
} hoefem_pointer_mode;

/* Value type of a CSR matrix written by hoefem_model_stiffness_add_prec. */
typedef enum hoefem_precision {
    HOEFEM_PRECISION_F64 = 0,
    HOEFEM_PRECISION_F32 = 1               /* FP32 element tangents added into FP32 values */
} hoefem_precision;

typedef struct hoefem_model hoefem_model;
typedef struct hoefem_schwarz hoefem_schwarz;
typedef struct hoefem_sim hoefem_sim;

const char* hoefem_last_error(void);
int32_t hoefem_nodes_per_element(int32_t element_type);

/* From<YoungPoisson> for LameParameters (fenris/src/solid/materials.rs:27-38), including the
 * reference's lambda convention documented in errata/tex/errata.tex:24-38. */
void hoefem_lame_from_young_poisson(double young, double poisson, double* mu, double* lambda);

/* StableNeoHookeanMaterial::compute_stress_tensor / build_stable_neohookean_dp_df
 * (fenris/src/solid/materials.rs:299-322, :506-557) on nq deformation gradients.
 * F, P: [nq * 9] column-major; dPdF: [nq * 81]. Either output may be NULL. Device pointers. */
hoefem_status hoefem_material_eval(double mu, double lambda, int64_t nq, const double* F, double* P, double* dPdF);

/* StableNeoHookeanMaterial::compute_strain_energy_density (fenris/src/solid/materials.rs:269-297) on nq
 * deformation gradients: psi[q] for F[9 q .. 9 q + 8] (column-major). Not on the Newton-CG path (a
 * diagnostic). Device pointers. */
hoefem_status hoefem_material_energy(double mu, double lambda, int64_t nq, const double* F, double* psi);

/* The elastic models of simulation_toolbox/src/fem/fe_model.rs:35-45 built on Stable Neo-Hookean. */
typedef enum hoefem_material_kind {
    /* StableNeoHookeanMaterial (fenris/src/solid/materials.rs:229-377); the default */
    HOEFEM_STABLE_NEO_HOOKEAN = 0,
    /* InvertibleMaterial<StableNeoHookeanMaterial> (materials.rs:862-945): F is projected
     * (signed singular values raised to the threshold, project_deformation_gradient) before
     * both the stress and the tangent. The reference's default Material uses threshold 1e-6
     * (simulation_toolbox/src/fem/fe_model.rs:145-158). */
    HOEFEM_INVERTIBLE_STABLE_NEO_HOOKEAN = 1,
    /* ProjectedStableNeoHookeanMaterial (materials.rs:387-441): Stable Neo-Hookean stress; the
     * tangent dP/dF with its negative eigenvalues clamped to zero (analytic eigensystem,
     * materials.rs:561-687). */
    HOEFEM_PROJECTED_STABLE_NEO_HOOKEAN = 2
} hoefem_material_kind;

/* hoefem_material_eval for any hoefem_material_kind; inversion_threshold is used by
 * HOEFEM_INVERTIBLE_STABLE_NEO_HOOKEAN only. Device pointers. */
hoefem_status hoefem_material_eval_kind(int32_t kind, double inversion_threshold, double mu, double lambda,
                                        int64_t nq, const double* F, double* P, double* dPdF);

/* rotation_svd (fenris/src/util.rs:68-94) on nq 3x3 matrices: F = U diag(S) Vt with U, Vt
 * rotations (det +1), |S| descending and the sign of det F on S[2]. F, U, Vt: [nq * 9]
 * column-major; S: [nq * 3]. Device pointers. */
hoefem_status hoefem_rotation_svd(int64_t nq, const double* F, double* U, double* S, double* Vt);

/* reorder_mesh_par (fenris/src/reorder.rs:52-88), the reorder the reference's scenes apply to their
 * meshes before building models: a reverse Cuthill-McKee permutation of the vertices (neighbours
 * visited by increasing degree, ties by increasing index) and the elements stably sorted by their
 * smallest new vertex index. HOST arrays; connectivity [num_elements * nodes_per_element];
 * outputs map target to source: new_vertices[i] = vertices[vertex_permutation[i]], likewise for
 * elements. */
hoefem_status hoefem_mesh_reorder(int32_t nodes_per_element, int64_t num_vertices, int64_t num_elements,
                                  const int32_t* connectivity, int64_t* vertex_permutation,
                                  int64_t* element_permutation);

/* ElasticityModelParallel for NodalModel / EmbeddedModel (fenris/src/solid/mod.rs:124-172,
 * fenris/src/solid/impl_model.rs:111-164, fenris/src/embedding/mod.rs:506-639).
 * Creation builds the greedy element coloring (paradis/src/coloring.rs:48-72), the CSR
 * pattern (fenris/src/assembly.rs:177-249) and the device-side element data. */
hoefem_status hoefem_model_create(const hoefem_model_desc* desc, hoefem_model** out);
void hoefem_model_destroy(hoefem_model* model);
hoefem_status hoefem_model_set_stream(hoefem_model* model, void* hip_stream);
/* hoefem_pointer_mode of every later call on this model and on the Schwarz operators and simulations
 * built on it (default HOEFEM_POINTERS_DEVICE). hoefem_material_eval(_kind) and hoefem_rotation_svd
 * have no model and always take device pointers. */
hoefem_status hoefem_model_set_pointer_mode(hoefem_model* model, int32_t mode);
/* The elastic model used by this model's forces and stiffness (and by every simulation built
 * on it): the ElasticMaterialModel of the reference's Material (fe_model.rs:35-45, :131-158).
 * kind is a hoefem_material_kind; inversion_threshold applies to the invertible kind only. */
hoefem_status hoefem_model_set_material(hoefem_model* model, int32_t kind, double inversion_threshold);
int64_t hoefem_model_ndof(const hoefem_model* model);
int64_t hoefem_model_nnz(const hoefem_model* model);
/* host outputs: rowptr [ndof + 1], col [nnz] */
hoefem_status hoefem_model_pattern(const hoefem_model* model, int32_t* rowptr, int32_t* col);
/* pass 0 = interior elements, pass 1 = interface elements; offsets [ncolors + 1], labels = element ids */
int32_t hoefem_model_num_colors(const hoefem_model* model, int32_t pass);
hoefem_status hoefem_model_colors(const hoefem_model* model, int32_t pass, int32_t* offsets, int32_t* labels);

/* assemble_elastic_pseudo_forces_into_par: f += f_int(u) (device vectors of ndof). */
hoefem_status hoefem_model_forces_add(hoefem_model* model, double mu, double lambda, const double* u, double* f);
/* assemble_transformed_stiffness_into_par with NoTransformation: values += K(u) (device, nnz). */
hoefem_status hoefem_model_stiffness_add(hoefem_model* model, double mu, double lambda, const double* u,
                                         double* values);
/* The same with the value type chosen by precision (hoefem_precision): values is double[nnz] for
 * HOEFEM_PRECISION_F64 (identical to hoefem_model_stiffness_add) or float[nnz] for
 * HOEFEM_PRECISION_F32 (element tangents evaluated in FP32, the mixed-precision operator). */
hoefem_status hoefem_model_stiffness_add_prec(hoefem_model* model, double mu, double lambda, const double* u,
                                              void* values, int32_t precision);
/* ElasticityModel::assemble_mass + to_csr: values = M (device, nnz). */
hoefem_status hoefem_model_mass(hoefem_model* model, double density, double* values);
/* apply_homogeneous_dirichlet_bc_csr (fenris/src/assembly.rs:799-869) on a matrix with the
 * model pattern; nodes is a HOST array. */
hoefem_status hoefem_model_dirichlet_csr(hoefem_model* model, int64_t num_nodes, const int32_t* nodes,
                                         double* values);
/* y = beta * y + alpha * A x for a matrix with the model pattern (spmv_csr,
 * fenris/src/sparse.rs:1168-1199; MklCsrLinearOperator::apply, newton_cg.rs:595-619). */
hoefem_status hoefem_model_spmv(hoefem_model* model, const double* values, double beta, double* y, double alpha,
                                const double* x);

/* Schwarz preconditioner (simulation_toolbox/src/fem/schwarz_precond.rs:26-216, :250-342;
 * SchwarzPreconditioner::update / apply, newton_cg.rs:297-370). Subdomains are node sets
 * given as a NestedVec (offsets [num_subdomains + 1], nodes); untreated nodes get a
 * Jacobi diagonal. All inputs are HOST arrays. */
hoefem_status hoefem_schwarz_create(hoefem_model* model, int64_t num_subdomains, const int64_t* offsets,
                                    const int32_t* nodes, int64_t num_untreated, const int32_t* untreated,
                                    hoefem_schwarz** out);
void hoefem_schwarz_destroy(hoefem_schwarz* s);
int64_t hoefem_schwarz_nnz(const hoefem_schwarz* s);
hoefem_status hoefem_schwarz_pattern(const hoefem_schwarz* s, int32_t* rowptr, int32_t* col);
/* P <- Schwarz(J); J_values: device, model pattern. */
hoefem_status hoefem_schwarz_update(hoefem_schwarz* s, const double* J_values);
/* copies P's values (device, schwarz nnz) */
hoefem_status hoefem_schwarz_values(hoefem_schwarz* s, double* values);
/* y = P x (device) */
hoefem_status hoefem_schwarz_apply(hoefem_schwarz* s, const double* x, double* y);

/* ConjugateGradient::solve_with_guess (fenris/src/cg.rs:377-482) with the relative residual
 * criterion (:85-124). A has the model pattern (device values); the preconditioner is the
 * Schwarz operator, or the Jacobi diagonal of A when schwarz == NULL. b, x: device.
 * Returns HOEFEM_ERROR_CG_* for the SolveErrorKind cases; *iterations is always written. */
hoefem_status hoefem_pcg(hoefem_model* model, const double* A_values, hoefem_schwarz* schwarz, const double* b,
                         double* x, double tolerance, int64_t max_iterations, int64_t* iterations);

/* hoefem_pcg without allocation. No counterpart in the reference: for callers that cannot
 * allocate device or pinned memory during a solve (serving loops, harnesses that forbid it, code
 * that runs under stream capture nearby). Same kernels, arguments and results as hoefem_pcg; the
 * work buffers live in a fixed per-GPU array of 18.9 MB held until the process ends, so the model
 * may have at most 262144 DOFs (larger models: HOEFEM_ERROR_INVALID_ARGUMENT) and the model must use
 * device pointers. Runs on the model's stream and synchronizes it. Calls that overlap in time on
 * one GPU share that array and race; serialize them. FP64 operators only. */
hoefem_status hoefem_pcg_noalloc(hoefem_model* model, const double* A_values, hoefem_schwarz* schwarz,
                                 const double* b, double* x, double tolerance, int64_t max_iterations,
                                 int64_t* iterations);

/* SpMV tile autotuner (no counterpart in the reference, whose SpMV is MKL's inspector-executor
 * with set_mv_hint/optimize, newton_cg.rs:731-743): times the node-row-block kernel with 8, 16
 * and 32 lanes per node on this pattern and keeps the fastest for every later SpMV and CG
 * launch. candidate_us (optional, [3]) receives the measured average launch times. */
hoefem_status hoefem_model_autotune(hoefem_model* model, int32_t* lanes_per_node, double* candidate_us);
hoefem_status hoefem_schwarz_autotune(hoefem_schwarz* s, int32_t* lanes_per_node, double* candidate_us);
/* the same for a simulation's CG: every (lanes of the Jacobian kernel, lanes of the preconditioner
 * kernel) pair in {8, 16, 32}^2 runs the best of three 96-iteration solves of J = M with the Schwarz
 * operator of M; the fastest pair is kept (lanes_preconditioner = 0 without Schwarz). candidate_us
 * (optional, [9], index 3 * a + b) receives the time per CG iteration of each pair. J and P are
 * overwritten (the next step assembles them again). */
hoefem_status hoefem_sim_autotune(hoefem_sim* sim, int32_t* lanes_matrix, int32_t* lanes_preconditioner,
                                  double* candidate_us);

/* Kernel timing (no counterpart in the reference, which times phases with coarse_prof scopes,
 * simulation_toolbox/src/fem/newton_cg.rs:685-764): when enabled, every launch of the CG
 * system-matrix kernel (fused direction update + SpMV + p.Ap) is bracketed by HIP events on
 * the model stream. "full" launches performed a CG iteration; the remaining launches found
 * the solve already finished and returned immediately. */
hoefem_status hoefem_profile_enable(hoefem_model* model, int32_t enable);
hoefem_status hoefem_profile_read(hoefem_model* model, int64_t* full_launches, double* full_ms,
                                  int64_t* all_launches, double* all_ms);
/* Kernels of the simulation's assembly timed the same way while profiling is enabled (the
 * reference's coarse_prof scopes "assemble elastic forces" and the stiffness assembly inside
 * "solve jacobian combination", newton_cg.rs:460, :685-764). */
typedef enum hoefem_profiled_kernel {
    HOEFEM_PROF_ELEMENT_TANGENT = 0,       /* element stiffness matrices of the Newton Jacobian */
    HOEFEM_PROF_JACOBIAN_GATHER = 1,       /* their color-ordered gather into J = M + dt^2 K (+ dt D) */
    HOEFEM_PROF_ELEMENT_FORCES = 2,        /* element force vectors of eval_f */
    HOEFEM_PROF_FORCE_GATHER = 3           /* their gather into f = M g + f_int - D v */
} hoefem_profiled_kernel;
/* launches and total kernel time (ms) of one hoefem_profiled_kernel since profiling was enabled */
hoefem_status hoefem_profile_kernel(hoefem_model* model, int32_t kernel, int64_t* launches, double* total_ms);

/* NewtonCgIntegrator3d (simulation_toolbox/src/fem/newton_cg.rs:43-295) over
 * backward_euler_step (hamilton2/src/integrators/euler.rs:155-199) and
 * newton_line_search + BacktrackingLineSearch (hamilton2/src/newton.rs:76-246). */
typedef struct hoefem_sim_desc {
    double mu, lambda, density;
    double gravity;                        /* Gravity::Scalar on the last axis */
    int32_t gravity_enabled;
    int64_t num_dirichlet;
    const int32_t* dirichlet_nodes;        /* Homogeneous BC nodes (host) */
    int64_t num_subdomains;                /* 0: DiagonalPreconditioner */
    const int64_t* subdomain_offsets;
    const int32_t* subdomain_nodes;
    int64_t num_untreated;
    const int32_t* untreated_nodes;
    double newton_tolerance_factor;        /* reference: 1e-5 (newton_cg.rs:282-286) */
    int64_t max_newton_iter;               /* reference: 100 */
    double cg_tolerance;                   /* reference: 1e-5 (newton_cg.rs:761) */
    int64_t cg_max_iter;                   /* reference: 500 (newton_cg.rs:772) */
    int32_t mixed_precision;               /* 0: FP64 step (the reference's arithmetic). 1: element tangents,
                                              the Jacobian values, its SpMV and the Schwarz operator in FP32
                                              (FP32 accumulation); residual, forces, line search, CG dots and
                                              norms and the Schwarz eigensolves stay FP64 (Newton is the FP64
                                              refinement loop around the FP32 linear operator). 2: FP64
                                              tangents, Jacobian and Schwarz factorizations; the CG runs on
                                              FP32 copies of the Jacobian and Schwarz values inside an FP64
                                              iterative refinement of each linear solve (same relative
                                              tolerance and iteration limit as the plain CG). 3: as 1, but
                                              the element tangents are computed in FP64 and each Jacobian
                                              entry is summed in FP64 and rounded to FP32 once. */
    /* Damping D = mass_damping_coefficient M + stiffness_damping_coefficient K+(u), K+ the stiffness
       at the (sub)step's initial displacement with every element matrix projected onto the positive
       semi-definite cone (unless the material is ProjectedStableNeoHookean), rebuilt at the start of
       every (sub)step (compute_damping_matrix_into, simulation_toolbox/src/fem/system_assembly.rs:
       145-194; integrate_model, newton_cg.rs:256-265). f gains -D v and the Newton matrix
       M + dt D + dt^2 K. 0 leaves a term out (Material's None); both 0: no damping. */
    double mass_damping_coefficient;
    double stiffness_damping_coefficient;
} hoefem_sim_desc;

typedef struct hoefem_step_stats {
    int64_t newton_iterations;
    int64_t cg_iterations;
    int64_t substeps;
    int64_t residual_evaluations;
    double final_residual;
    double tolerance;
} hoefem_step_stats;

hoefem_status hoefem_sim_create(hoefem_model* model, const hoefem_sim_desc* desc, hoefem_sim** out);
void hoefem_sim_destroy(hoefem_sim* sim);
double hoefem_sim_representative_force(const hoefem_sim* sim);
/* mass matrix with Dirichlet conditions applied (device, nnz) */
hoefem_status hoefem_sim_mass_values(hoefem_sim* sim, double* values);
/* The damping matrix D of the current (sub)step on the model pattern (device output, nnz);
 * zeros when the simulation has no damping. */
hoefem_status hoefem_sim_damping_values(hoefem_sim* sim, double* values);
/* J = M + dt^2 K with Dirichlet conditions of the last Newton iteration (device, nnz) */
hoefem_status hoefem_sim_jacobian_values(hoefem_sim* sim, double* values);
/* F(w) of BackwardEulerImplicitFunction::eval_into at (u, v) (device vectors) */
hoefem_status hoefem_sim_residual(hoefem_sim* sim, double t, double dt, const double* u, const double* v,
                                  const double* w, double* F);
/* one time step: u, v device vectors updated in place */
hoefem_status hoefem_sim_step(hoefem_sim* sim, double t, double dt, double* u, double* v, hoefem_step_stats* stats);

/* Prescribed Dirichlet values: what DirichletBoundaryConditions::apply_displacement_bcs and
 * apply_velocity_bcs (simulation_toolbox/src/fem/bcs.rs:13-27) return for the sim's Dirichlet nodes,
 * three per node in the order of hoefem_sim_desc.dirichlet_nodes (host arrays; NULL: zeros, the
 * Homogeneous condition). Every step and dynamic-system call then sets the Dirichlet unknowns of u and v
 * to them (apply_dirichlet_bcs_unknowns, simulation_toolbox/src/fem/system_assembly.rs:78-104).
 * Time-independent conditions (ConstantDisplacement, ConstantUniformDisplacement) are set once;
 * a time-dependent one is set before each step with its values at the step's end time. */
hoefem_status hoefem_sim_set_dirichlet_values(hoefem_sim* sim, const double* u_values, const double* v_values);

/* ------------------------------------------------------------------ batched scenes
 * No counterpart in the reference, which steps one scene per process (integrate_model_with_substeps,
 * simulation_toolbox/src/fem/newton_cg.rs:146-228). A batch steps independent sims (each its own
 * model, material and boundary conditions) together: their Newton loops run in lockstep, their
 * residual evaluations, assemblies and line-search trials are issued before one synchronization per
 * round, and each round's linear solves run as one batched PCG whose scenes stop on the device as
 * they converge. Every scene's u and v equal what hoefem_sim_step gives it, bit for bit (a scene
 * whose single-step attempt fails continues alone with hoefem_sim_step's substepping). The sims
 * must outlive the batch, take device pointers, have no learned correction attached and share one
 * precision mode (FP64, or mixed_precision 1); their models may use different streams. */
typedef struct hoefem_batch hoefem_batch;
hoefem_status hoefem_batch_create(hoefem_sim* const* sims, int32_t count, hoefem_batch** out);
void hoefem_batch_destroy(hoefem_batch* batch);
/* One time step of every scene: dt[i], u[i], v[i] (device vectors, updated in place) of sim i.
 * Optional outputs per scene: stats; latency_ms, the time from the call's start until the scene's
 * Newton loop ended; status, the scene's hoefem_sim_step status. Returns the first scene error. */
hoefem_status hoefem_batch_step(hoefem_batch* batch, const double* dt, double* const* u, double* const* v,
                                hoefem_step_stats* stats, double* latency_ms, int32_t* status);

/* ------------------------------------------------------------------ dynamic system
 * The DynamicSystem / DifferentiableDynamicSystem surface of NewtonCgDynamicSystem
 * (hamilton2/src/dynamic_system/mutable.rs:14-76, simulation_toolbox/src/fem/newton_cg.rs:413-820), so
 * that hamilton2's own backward_euler_step (hamilton2/src/integrators/euler.rs:155-199) and
 * newton_line_search can drive the simulation: hoefem_sim_step is that composition run inside the
 * library. The vector arithmetic of hoefem_sim_step rounds every operation separately (no fused
 * multiply-add), so the same step composed by the caller from these calls, with the caller's host
 * IEEE arithmetic for the vector updates, reproduces hoefem_sim_step's iterates bit for bit. Option<T>
 * arguments are pointers, NULL for None. */

/* compute_damping_matrix_into (simulation_toolbox/src/fem/system_assembly.rs:145-194) at the step's
 * initial displacement u, as integrate_model does before each backward_euler_step (newton_cg.rs:
 * 256-265); nothing when the simulation has no damping. */
hoefem_status hoefem_sim_update_damping(hoefem_sim* sim, const double* u);
/* DynamicSystem::eval_f (newton_cg.rs:441-478): f = M g + f_int(u) - D v, Dirichlet rows zeroed, with
 * the Dirichlet values applied to copies of u and v. Overwrites f. */
hoefem_status hoefem_sim_eval_f(hoefem_sim* sim, double t, const double* u, const double* v, double* f);
/* DynamicSystem::apply_mass_matrix (newton_cg.rs:428-431): y += M x (M with Dirichlet conditions). */
hoefem_status hoefem_sim_apply_mass_add(hoefem_sim* sim, double* y, const double* x);
/* DifferentiableDynamicSystem::set_state (newton_cg.rs:661-671): stores (u, v) with the Dirichlet values
 * applied; clears the coefficients of init_solve_jacobian_combination. */
hoefem_status hoefem_sim_set_state(hoefem_sim* sim, double t, const double* u, const double* v);
/* init_solve_jacobian_combination (newton_cg.rs:673-679): H = M + alpha df/du + beta df/dv
 * = M - alpha K - beta D at the stored state. */
hoefem_status hoefem_sim_init_solve_jacobian_combination(hoefem_sim* sim, const double* alpha, const double* beta);
/* solve_jacobian_combination (newton_cg.rs:680-802): assembles H with the Dirichlet conditions and
 * updates the preconditioner when the state or the coefficients changed, then PCG from a zero guess
 * (the simulation's CG settings; MaxIterationsReached and IndefiniteOperator are tolerated, any other
 * CG status is returned). cg_iterations (optional) receives the iteration count. */
hoefem_status hoefem_sim_solve_jacobian_combination(hoefem_sim* sim, const double* rhs, double* sol,
                                                     int64_t* cg_iterations);
/* init_apply_jacobian_combination / apply_jacobian_combination (mutable.rs:44-58; unimplemented in
 * NewtonCgDynamicSystem, built here): y += (alpha df/du + beta df/dv + gamma M) x at the stored state,
 * with the Dirichlet treatment of compute_jacobian_combination_into (so that with gamma = 1 it is the
 * operator solve_jacobian_combination inverts). */
hoefem_status hoefem_sim_init_apply_jacobian_combination(hoefem_sim* sim, const double* alpha, const double* beta,
                                                         const double* gamma);
hoefem_status hoefem_sim_apply_jacobian_combination(hoefem_sim* sim, double* y, const double* x);

/* ------------------------------------------------------------------ learned coarse correction
 * No counterpart in the reference: an extra correction on top of its Schwarz preconditioner
 * (SystemPreconditioner::apply, simulation_toolbox/src/fem/newton_cg.rs:397-401), used by the
 * simulation's linear solves when enabled.
 *
 * The nodes are split into aggregates; each aggregate carries 64 orthonormal modes taken from
 * error snapshots x* - x_k of the plain Schwarz-CG's own solves. The correction is linear,
 * symmetric and positive semidefinite in the residual, so the preconditioner stays SPD and CG
 * keeps its Krylov properties:
 *     z = P r + sum_a Phi_a diag(gamma_a) Phi_a^T r,   gamma_a = kappa softplus(MLP(f_a)),
 * where one network shared by all aggregates (64 -> 128 -> 128 -> 64, GELU, BF16 weights, FP32
 * accumulation) maps the aggregate's residual projection profile f_a (RMS of Phi_a^T r over
 * harvested residuals, normalized) to its 64 mode gains. Training minimizes
 * -(r^T z)^2 / (z^T J z) (the energy left after the best step along z) over harvested residuals
 * (Adam). A learned solve runs the flexible (Polak-Ribiere) PCG under the plain path's rules
 * (the CG iteration limit is tolerated; the Newton residual tolerance is unchanged); on a
 * breakdown the plain Schwarz-CG solves the same system from zero.
 *
 * Lifetime: the correction refers to its simulation; destroying the simulation first detaches
 * it (training then fails with HOEFEM_ERROR_INVALID_ARGUMENT). */
typedef struct hoefem_learned hoefem_learned;

typedef struct {
    int32_t num_aggregates; /* target count (greedy growth may end a few above it) */
    uint64_t seed;          /* weight initialization */
} hoefem_learned_desc;

typedef struct {
    int32_t num_aggregates, max_aggregate_dofs;
    int64_t harvested_pairs; /* (residual, error) snapshot pairs taken so far */
    int64_t stored_pairs;    /* pairs held in the ring */
    int32_t trained, enabled;
    int64_t learned_solves, fallbacks, learned_iterations;
    int32_t last_status; /* of the last learned solve: HOEFEM_OK or the CG status that caused the fallback */
} hoefem_learned_stats;

/* number of network parameters (FP32 master copy): W1[128][64] b1[128] W2[128][128] b2[128] W3[64][128] b3[64] */
int64_t hoefem_learned_param_count(void);
hoefem_status hoefem_learned_create(hoefem_sim* sim, const hoefem_learned_desc* desc, hoefem_learned** out);
hoefem_status hoefem_learned_destroy(hoefem_learned* l);
/* while enabled, every plain Schwarz-CG solve of the simulation keeps (x_k, r_k) every `every`
 * iterations and stores (x* - x_k, r_k) in a ring of `capacity` (>= 64) pairs */
hoefem_status hoefem_learned_harvest(hoefem_learned* l, int32_t enable, int32_t capacity, int32_t every);
hoefem_status hoefem_learned_info(hoefem_learned* l, hoefem_learned_stats* out);
/* builds the aggregate modes from the stored error snapshots (first call) and runs `steps` Adam
 * steps on batches of `batch` stored residuals against the simulation's current Jacobian and
 * Schwarz operator; losses are averaged over the batch (initial: before the first step) */
hoefem_status hoefem_learned_train(hoefem_learned* l, int32_t steps, double learning_rate, int32_t batch,
                                   double* initial_loss, double* final_loss);
hoefem_status hoefem_learned_enable(hoefem_learned* l, int32_t enable);
/* host arrays of hoefem_learned_param_count() floats */
hoefem_status hoefem_learned_get_params(hoefem_learned* l, float* params);
hoefem_status hoefem_learned_set_params(hoefem_learned* l, const float* params);
/* network on n device samples: y[n][64] = MLP(x[n][64]); grad (device, param_count floats) =
 * d(sum_s dy[s] . MLP(x[s])) / d params */
hoefem_status hoefem_learned_mlp_forward(hoefem_learned* l, int64_t n, const float* x, float* y);
hoefem_status hoefem_learned_mlp_grad(hoefem_learned* l, int64_t n, const float* x, const float* dy, float* grad);

#ifdef __cplusplus
}
#endif
#endif
