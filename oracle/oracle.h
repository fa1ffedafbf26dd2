/*
 * CPU FP64 oracle for the implicit elastodynamics step (test infrastructure only).
 *
 * This library restates, in plain C++ on the CPU, the reference algorithms of the
 * hot path: Stable Neo-Hookean stress/tangent, higher-order element bases, colored
 * force/stiffness assembly into a fixed CSR pattern, homogeneous Dirichlet
 * conditions, the additive Schwarz preconditioner, preconditioned CG, Newton with
 * backtracking line search and Backward Euler with substepping.
 *
 * It verifies the HIP library (tests, smoke test, CPU baseline of the benchmark).
 * Product code never links or calls it.
 *
 * Conventions: 3x3 matrices are column-major double[9] (m[r + 3*c]); dP/dF is a
 * column-major 9x9 matrix in the vec() ordering vec(A)[3*c + r] = A(r, c), exactly
 * like fenris/src/solid/materials.rs:745-761. CSR matrices use int32 offsets and
 * column indices (the reference's MKL boundary type, simulation_toolbox/src/fem/
 * newton_cg.rs:621-658).
 */
#ifndef HOEFEM_ORACLE_H
#define HOEFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types; node orderings follow fenris/src/element.rs */
enum {
    ORC_TET4 = 0,
    ORC_TET10 = 1,
    ORC_TET20 = 2,
    ORC_HEX8 = 3,
    ORC_HEX20 = 4,
    ORC_HEX27 = 5
};

/* elastic model kinds (simulation_toolbox/src/fem/fe_model.rs:35-45) */
enum {
    ORC_MAT_SNH = 0,            /* StableNeoHookeanMaterial */
    ORC_MAT_INVERTIBLE_SNH = 1, /* InvertibleMaterial<StableNeoHookeanMaterial> */
    ORC_MAT_PROJECTED_SNH = 2   /* ProjectedStableNeoHookeanMaterial */
};

/* status codes (same numbering as the product library's hoefem_status) */
enum {
    ORC_OK = 0,
    ORC_ERR_INVALID = 1,
    ORC_ERR_SINGULAR = 3,
    ORC_ERR_PATTERN = 4,
    ORC_ERR_CG_MAX_ITER = 5,
    ORC_ERR_CG_INDEFINITE_OPERATOR = 6,
    ORC_ERR_CG_INDEFINITE_PRECONDITIONER = 7,
    ORC_ERR_NEWTON_MAX_ITER = 8,
    ORC_ERR_LINE_SEARCH = 9,
    ORC_ERR_SUBSTEPPING = 10,
    ORC_ERR_NOT_FINITE = 11,
    ORC_ERR_JACOBIAN = 12
};

const char* orc_last_error(void);
int orc_nodes_per_element(int element_type);
int orc_geometry_nodes(int element_type);

/* ---------------- materials (fenris/src/solid/materials.rs) ---------------- */
void orc_lame_from_young_poisson(double young, double poisson, double* mu, double* lambda);
void orc_snh_reparametrize(double mu, double lambda, double* mu_out, double* lambda_out);
int orc_snh_energy(double mu, double lambda, int64_t nq, const double* F, double* psi);
int orc_snh_stress(double mu, double lambda, int64_t nq, const double* F, double* P);
int orc_snh_dpdf(double mu, double lambda, int64_t nq, const double* F, double* dpdf);
int orc_snh_contract_with(double mu, double lambda, const double* F, const double* a, const double* b,
                          double* out);
int orc_snh_contract_multiple(double mu, double lambda, const double* F, int n, const double* a, double* out);

/* ---------------- elements (fenris/src/element.rs) ---------------- */
/* SVD-based variants of the material (fenris/src/solid/materials.rs:387-441, :862-945,
 * fenris/src/util.rs:68-94). svd3: sigma >= 0, U and V orthogonal (one-sided Jacobi);
 * rotation_svd: U, Vt rotations, signed S. kind: ORC_MAT_*; threshold for the invertible kind. */
int orc_svd3(int64_t nq, const double* F, double* U, double* S, double* V);
int orc_rotation_svd(int64_t nq, const double* F, double* U, double* S, double* Vt);
int orc_project_deformation_gradient(double threshold, int64_t nq, const double* F, double* Fout);
int orc_snh_dpdf_eigen(double mu, double lambda, const double* F, double* evals, double* evecs);
int orc_material_stress(int kind, double threshold, double mu, double lambda, int64_t nq, const double* F, double* P);
int orc_material_dpdf(int kind, double threshold, double mu, double lambda, int64_t nq, const double* F,
                      double* dpdf);
int orc_material_contract_multiple(int kind, double threshold, double mu, double lambda, const double* F, int n,
                                   const double* a, double* out);

int orc_element_basis(int element_type, int64_t npts, const double* xi, double* phi);
int orc_element_gradients(int element_type, int64_t npts, const double* xi, double* grad);
int orc_element_reference_jacobian(int element_type, const double* nodes, const double* xi, double* J);

/* ---------------- model ---------------- */
typedef struct orc_quadrature {
    int32_t interior_nq;
    const double* interior_w;      /* [interior_nq] */
    const double* interior_xi;     /* [interior_nq*3] */
    const int64_t* interface_offsets; /* [num_interface+1] or NULL */
    const double* interface_w;
    const double* interface_xi;
} orc_quadrature;

typedef struct orc_model_desc {
    int32_t element_type;
    int64_t num_vertices;
    const double* vertices;        /* [num_vertices*3] */
    int64_t num_elements;
    const int32_t* connectivity;   /* [num_elements*nodes_per_element] */
    int64_t num_interior;          /* elements [0,num_interior) form the first coloring pass */
    orc_quadrature mass_quadrature;
    orc_quadrature stiffness_quadrature;
    orc_quadrature elliptic_quadrature;
    double mass_regularization_factor;
} orc_model_desc;

typedef struct orc_model orc_model;

int orc_model_create(const orc_model_desc* desc, orc_model** out);
void orc_model_destroy(orc_model* m);
/* the elastic model used by forces / stiffness / simulations of this model (ORC_MAT_*) */
int orc_model_set_material(orc_model* m, int kind, double threshold);
int64_t orc_model_ndof(const orc_model* m);
int orc_model_num_colors(const orc_model* m, int pass);
/* colors of a pass as NestedVec: offsets[ncolors+1], labels (global element indices) */
int orc_model_colors(const orc_model* m, int pass, int32_t* offsets, int32_t* labels);
int64_t orc_model_nnz(const orc_model* m);
int orc_model_pattern(const orc_model* m, int32_t* rowptr, int32_t* col);
int orc_model_assemble_mass(const orc_model* m, double density, double* values);
int orc_model_forces_add(const orc_model* m, double mu, double lambda, const double* u, double* f);
int orc_model_stiffness_add(const orc_model* m, double mu, double lambda, const double* u, double* values);
/* the same with every element matrix projected onto the positive semi-definite cone */
int orc_model_stiffness_add_projected(const orc_model* m, double mu, double lambda, const double* u, double* values);
/* cuthill_mckee on a square CSR pattern (perm: target -> source) and reorder_mesh_par
 * (fenris/src/reorder.rs:52-88, :169-243) on an element-major connectivity */
int orc_cuthill_mckee(int64_t n, const int32_t* rowptr, const int32_t* col, int64_t* perm);
int orc_mesh_reorder(int nodes_per_element, int64_t num_vertices, int64_t num_elements, const int32_t* conn,
                     int64_t* vertex_perm, int64_t* element_perm);
/* DefaultSemidefiniteProjection of one symmetric dim x dim matrix (column-major) */
int orc_psd_project(int dim, const double* a, double* out);

/* paradis/src/coloring.rs:48-72 on an arbitrary NestedVec of node sets */
int orc_greedy_coloring(int64_t nsets, const int64_t* offsets, const int32_t* nodes, int32_t* color_of_set,
                        int32_t* num_colors);

/* ---------------- sparse + boundary conditions ---------------- */
int orc_coo_pattern(int64_t n, int64_t nsets, const int64_t* offsets, const int32_t* nodes, int sdim,
                    int64_t* nnz_out, int32_t* rowptr, int32_t* col);
void orc_spmv(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals, double beta, double* y,
              double alpha, const double* x);
void orc_dirichlet_csr(int64_t n, const int32_t* rowptr, const int32_t* col, double* vals, int64_t nnodes,
                       const int32_t* nodes, int sdim);
void orc_dirichlet_rhs(double* rhs, int64_t nnodes, const int32_t* nodes);

/* ---------------- Schwarz preconditioner (simulation_toolbox/src/fem/schwarz_precond.rs) ---- */
typedef struct orc_schwarz orc_schwarz;
int orc_schwarz_create(int64_t num_nodes, int64_t nsub, const int64_t* offsets, const int32_t* nodes,
                       int64_t nuntreated, const int32_t* untreated, orc_schwarz** out);
void orc_schwarz_destroy(orc_schwarz* s);
int64_t orc_schwarz_nnz(const orc_schwarz* s);
int orc_schwarz_pattern(const orc_schwarz* s, int32_t* rowptr, int32_t* col);
int orc_schwarz_update(orc_schwarz* s, const int32_t* rowptr, const int32_t* col, const double* J);
int orc_schwarz_values(const orc_schwarz* s, double* values);
int orc_invert_block(int m, const double* block, double* out);
int orc_symmetric_eigen(int m, const double* a, double* evals, double* evecs);

/* ---------------- CG (fenris/src/cg.rs:377-482) ---------------- */
int orc_pcg_csr(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals,
                const int32_t* prowptr, const int32_t* pcol, const double* pvals, const double* b, double* x,
                double tol, int64_t max_iter, int64_t* iterations);

/* ---------------- Newton + Backward Euler with callbacks (hamilton2) ---------------- */
typedef void (*orc_eval_fn)(void* ctx, const double* x, double* f);
typedef int (*orc_solve_fn)(void* ctx, const double* x, const double* rhs, double* sol);
int orc_newton(int64_t n, double* x, void* ctx, orc_eval_fn eval, orc_solve_fn solve, double tolerance,
               int64_t max_iter, int line_search, int64_t* iterations);
/* backward_euler_step for a dynamic system given by callbacks (MockDynamicSystem-style tests) */
typedef int (*orc_dyn_eval_fn)(void* ctx, double t, const double* u, const double* v, double* f);
typedef void (*orc_dyn_mass_fn)(void* ctx, double* y, const double* x);
typedef int (*orc_dyn_solve_fn)(void* ctx, double t, const double* u, const double* v, double alpha, double beta,
                                const double* rhs, double* sol);
int orc_backward_euler_step(int64_t n, double* u, double* v, double t0, double dt, void* ctx, orc_dyn_eval_fn eval_f,
                            orc_dyn_mass_fn apply_mass, orc_dyn_solve_fn solve_jac, double tolerance,
                            int64_t max_newton_iter, int64_t* iterations);

/* integrate_model_with_substeps (simulation_toolbox/src/fem/newton_cg.rs:146-228) over
 * orc_backward_euler_step's dynamic system: on failure the step is retried from the initial state
 * with 5x as many substeps, up to (excluding) 1000. *substeps receives the count that succeeded. */
int orc_integrate_with_substeps(int64_t n, double* u, double* v, double t0, double dt, void* ctx,
                                orc_dyn_eval_fn eval_f, orc_dyn_mass_fn apply_mass, orc_dyn_solve_fn solve_jac,
                                double tolerance, int64_t max_newton_iter, int64_t* substeps,
                                int64_t* newton_iterations);

/* ---------------- the FE dynamic system (simulation_toolbox/src/fem/newton_cg.rs) ---------- */
typedef struct orc_sim_desc {
    double mu, lambda, density;
    double gravity;                /* scalar gravity on the last axis (Gravity::Scalar) */
    int32_t gravity_enabled;
    int64_t num_dirichlet;
    const int32_t* dirichlet_nodes;
    /* Schwarz subdomains; nsub == 0 selects the diagonal (Jacobi) preconditioner */
    int64_t num_subdomains;
    const int64_t* subdomain_offsets;
    const int32_t* subdomain_nodes;
    int64_t num_untreated;
    const int32_t* untreated_nodes;
    double newton_tolerance_factor; /* 1e-5 in the reference */
    int64_t max_newton_iter;        /* 100 */
    double cg_tolerance;            /* 1e-5 */
    int64_t cg_max_iter;            /* 500 */
    int32_t num_threads;            /* 0: OpenMP default */
    /* Material::mass_damping_coefficient / stiffness_damping_coefficient (0 = None):
     * D = a_M M + a_K K+(u_n) (simulation_toolbox/src/fem/system_assembly.rs:145-194) */
    double mass_damping_coefficient;
    double stiffness_damping_coefficient;
} orc_sim_desc;

typedef struct orc_step_stats {
    int64_t newton_iterations;
    int64_t cg_iterations;
    int64_t substeps;
    int64_t line_search_evals;
    double final_residual;
    double tolerance;
    double t_forces, t_stiffness, t_jacobian, t_precond, t_cg;
} orc_step_stats;

typedef struct orc_sim orc_sim;
int orc_sim_create(const orc_model* model, const orc_sim_desc* desc, orc_sim** out);
void orc_sim_destroy(orc_sim* s);
double orc_sim_representative_force(const orc_sim* s);
int orc_sim_set_dirichlet_values(orc_sim* s, const double* u_values, const double* v_values);
int orc_sim_mass_values(const orc_sim* s, double* values);
int orc_sim_step(orc_sim* s, double t, double dt, double* u, double* v, orc_step_stats* stats);
/* the Jacobian combination of the last Newton iteration (J = M + dt^2 K with BCs) */
int orc_sim_jacobian_values(const orc_sim* s, double* values);
/* the damping matrix of the last (sub)step (zeros without damping) */
int orc_sim_damping_values(const orc_sim* s, double* values);
int orc_sim_residual(orc_sim* s, double t, double dt, const double* u, const double* v, const double* w,
                     double* F);

/* The per-aggregate network of the learned coarse correction (no reference counterpart):
 * 64 -> 128 -> 128 -> 64, GELU (erf form) after the two hidden layers. Parameter layout
 * W1[128][64] b1[128] W2[128][128] b2[128] W3[64][128] b3[64] (33088 floats). FP64 arithmetic;
 * with emulate_bf16 the inputs, weights and hidden activations are rounded to BF16
 * (round-to-nearest-even) before they enter a product, as the GPU kernels do. */
int orc_mlp_forward(const float* params, int64_t n, const float* x, int emulate_bf16, double* y);
/* grad = d(sum_s dy[s] . y[s]) / d params; with emulate_bf16 the derivative treats the BF16
 * rounding of the weights as the identity (gradients land on the FP32 master copy) */
int orc_mlp_grad(const float* params, int64_t n, const float* x, const float* dy, int emulate_bf16, double* grad);

#ifdef __cplusplus
}
#endif
#endif
