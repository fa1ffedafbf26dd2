// C ABI of the hoefem library: model setup, assembly entry points, Schwarz, PCG and the
// Backward-Euler Newton driver (all device work is enqueued on the model's stream).
#include <algorithm>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstring>

#include "fem_device.hpp"
#include "internal.hpp"

namespace hoefem {

// solver.hip
void launch_state(Model& m, const double* up, const double* w, double dt, const uint8_t* dirmask, const double* bcu,
                  double* uu);
void launch_bc_copy(Model& m, const double* src, const uint8_t* dirmask, const double* bc, double* dst);
void launch_node_force(Model& m, const double* gforce, const double* dv, const uint8_t* dirmask, double* f);
void launch_scale(Model& m, const double* src, double sc, double* dst);
void launch_mass_add(Model& m, const double* Mvals, const double* x, double* y, CgWork& cw, double* norm2);
template <class T>
void launch_comb(Model& m, int64_t n, T* out, double cd, const double* D, double cm, const double* M);
void launch_state_v(Model& m, const double* vn, const double* w, const uint8_t* dirmask, const double* bcv,
                    double* vv);
void launch_axpy_neg(Model& m, double* w, double delta, const double* sol);
void launch_be_begin(Model& m, const double* u, const double* v, double dt, double* up, double* vn);
void launch_be_end(Model& m, double* u, double* v, const double* w, double dt, const uint8_t* dirmask,
                   const double* bcu, const double* bcv);
void launch_set_dirichlet(Model& m, double* a, double* b, const uint8_t* dirmask, const double* bca,
                          const double* bcb);
void launch_finite_check(Model& m, const double* a, const double* b);
void launch_f32_to_f64(Model& m, const float* a, double* b, int64_t n);
void launch_f64_to_f32(Model& m, const double* a, float* b, int64_t n);
void launch_fill(Model& m, double* a, int64_t n, double value);
void launch_add(Model& m, double* a, const double* b, int64_t n);
double device_dot(Model& m, CgWork& w, const double* a, const double* b, double* d_out, double* h_out);

thread_local std::string g_last_error;

static void upload_quadrature(Model& m, const hoefem_quadrature& q, DeviceQuadrature& dq) {
    const int64_t nif = m.ne - m.ninterior;
    if (m.ninterior > 0 && (q.interior_num_points <= 0 || !q.interior_weights || !q.interior_points))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "missing interior quadrature");
    if (nif > 0 && (!q.interface_offsets || !q.interface_weights || !q.interface_points))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "missing interface quadrature");
    std::vector<int32_t> off(m.ne + 1, 0);
    for (int64_t e = 0; e < m.ne; ++e) {
        const int64_t cnt = e < m.ninterior ? q.interior_num_points
                                            : q.interface_offsets[e - m.ninterior + 1] - q.interface_offsets[e - m.ninterior];
        if (cnt < 0) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad interface quadrature offsets");
        if ((int64_t)off[e] + cnt > INT32_MAX) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "too many quadrature points");
        off[e + 1] = off[e] + (int32_t)cnt;
        dq.max_points = std::max<int>(dq.max_points, (int)cnt);
    }
    std::vector<double> w(off[m.ne]), xi(3 * (size_t)off[m.ne]);
    for (int64_t e = 0; e < m.ne; ++e) {
        const double *sw, *sx;
        if (e < m.ninterior) {
            sw = q.interior_weights;
            sx = q.interior_points;
        } else {
            const int64_t k = q.interface_offsets[e - m.ninterior];
            sw = q.interface_weights + k;
            sx = q.interface_points + 3 * k;
        }
        const int cnt = off[e + 1] - off[e];
        std::copy(sw, sw + cnt, w.begin() + off[e]);
        std::copy(sx, sx + 3 * cnt, xi.begin() + 3 * (size_t)off[e]);
    }
    dq.off.upload(off, m.stream);
    dq.w.upload(w, m.stream);
    dq.xi.upload(xi, m.stream);
}

static void model_create(Model& mm, const hoefem_model_desc* d) {
    if (!d) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null descriptor");
    const int n = nodes_per_element(d->element_type);
    if (n < 0) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unknown element type");
    if (d->num_vertices < 0 || d->num_elements < 0 || d->num_interior < 0 || d->num_interior > d->num_elements)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad mesh sizes");
    if ((d->num_vertices && !d->vertices) || (d->num_elements && !d->connectivity))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null mesh arrays");
    Model* m = &mm;
    m->type = d->element_type;
    m->n = n;
    m->gn = geometry_nodes(d->element_type);
    m->nv = d->num_vertices;
    m->ne = d->num_elements;
    m->ninterior = d->num_interior;
    m->mass_reg = d->mass_regularization_factor;
    m->vertices.assign(d->vertices, d->vertices + 3 * m->nv);
    m->conn.assign(d->connectivity, d->connectivity + (size_t)n * m->ne);
    for (int32_t c : m->conn)
        if (c < 0 || c >= m->nv) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "connectivity out of range");
    // colors of the interior and interface passes (fenris/src/model.rs:90-102,
    // fenris/src/embedding/mod.rs:262-270)
    const int64_t bounds[3] = {0, m->ninterior, m->ne};
    for (int p = 0; p < 2; ++p) {
        const int64_t cnt = bounds[p + 1] - bounds[p];
        std::vector<int64_t> off(cnt + 1);
        for (int64_t i = 0; i <= cnt; ++i) off[i] = i * n;
        auto cols = greedy_coloring(cnt, off.data(), m->conn.data() + bounds[p] * n);
        for (auto& c : cols)
            for (auto& l : c) l += (int32_t)bounds[p];
        m->colors[p] = std::move(cols);
    }
    std::vector<int64_t> eoff(m->ne + 1);
    for (int64_t i = 0; i <= m->ne; ++i) eoff[i] = i * n;
    build_node_pattern(m->nv, m->ne, eoff.data(), m->conn.data(), {}, m->pattern);
    m->scatter.build(m->pattern, m->ne, eoff.data(), m->conn.data(), m->colors, 2);
    // device data
    m->pattern.upload(m->stream);
    m->scatter.upload(m->stream);
    m->d_vertices.upload(m->vertices, m->stream);
    std::vector<int32_t> soa((size_t)n * m->ne);
    for (int64_t e = 0; e < m->ne; ++e)
        for (int k = 0; k < n; ++k) soa[(size_t)k * m->ne + e] = m->conn[(size_t)e * n + k];
    m->d_conn.upload(soa, m->stream);
    upload_quadrature(*m, d->mass_quadrature, m->q_mass);
    upload_quadrature(*m, d->stiffness_quadrature, m->q_stiff);
    upload_quadrature(*m, d->elliptic_quadrature, m->q_ell);
    if (m->mass_reg > 0.0) {
        const auto& q = d->mass_quadrature;
        if (q.interior_num_points <= 0) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "regularization needs an interior rule");
        m->q_interior_mass.w.upload(q.interior_weights, q.interior_num_points, m->stream);
        m->q_interior_mass.xi.upload(q.interior_points, 3 * (size_t)q.interior_num_points, m->stream);
        m->q_interior_mass.max_points = q.interior_num_points;
    }
    m->d_err.alloc(1);
    HIP_CHECK(hipMemsetAsync(m->d_err.get(), 0, sizeof(int), m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
}

static void raise_kernel_errors(Model& m) {
    const int e = read_and_clear_err(m);
    if (e & KERR_SINGULAR) throw Error(HOEFEM_ERROR_SINGULAR, "singular deformation gradient or element Jacobian");
    if (e & KERR_ZERO_DIAG) throw Error(HOEFEM_ERROR_SINGULAR, "zero diagonal entry");
    if (e & KERR_NONFINITE) throw Error(HOEFEM_ERROR_NOT_FINITE, "non-finite values");
    if (e & KERR_SCHWARZ)
        throw Error(HOEFEM_ERROR_SINGULAR, "a Schwarz block above 96 x 96 is not certified positive definite "
                                           "(|B|_F |B^-1|_F >= 5e11 or a non-positive pivot)");
    if (e) throw Error(HOEFEM_ERROR_DEVICE, "kernel error flag set");
}

// ------------------------------------------------------------------ simulation
struct Sim {
    Model* model = nullptr;
    hoefem_sim_desc d{};
    double mu_r = 0, lambda_r = 0;
    std::vector<int32_t> dir;
    std::unique_ptr<Schwarz> schwarz;
    DevBuf<uint8_t> d_dirmask;
    DevBuf<int32_t> d_bc_nodes;
    int64_t n_bc = 0;
    DevBuf<double> d_M, d_J, d_gforce, d_up, d_vn, d_w, d_F, d_sol, d_uu, d_us, d_vs, d_scale, d_norm2;
    bool mixed = false;   // FP32 CG operator (J and Schwarz values)
    int mixed_mode = 0;   // 1: FP32 element tangents, the Newton loop refines; 2: FP64 tangents and J, FP32
                          //    copy for the CG, FP64 iterative refinement of every linear solve; 3: FP64
                          //    tangents rounded into the FP32 J at the gather, the Newton loop refines
    // J is assembled straight into FP32 values (modes 1 and 3): the CG runs on d_J32 and d_J is unused
    bool fp32_jacobian() const { return mixed_mode == 1 || mixed_mode == 3; }
    DevBuf<float> d_J32;
    DevBuf<double> d_res, d_dx;  // iterative refinement (mode 2)
    int64_t refinements = 0;
    PinnedBuf<double> h_norm2{1};
    CgWork cg;
    Learned* learned = nullptr;                 // optional learned coarse correction (not owned)
    struct hoefem_learned* learned_handle = nullptr;
    double rep_force = 0.0;
    int64_t cg_total = 0, evals = 0;
    int last_err_flags = 0;
    bool damping = false;           // a damping matrix D is in use
    DevBuf<double> d_D, d_vv, d_dv;  // D (model pattern), v = v_n + w, D v
    DevBuf<double> d_f;              // eval_f output of the Newton residual
    // dynamic-system state of the hamilton2 entry points (NewtonCgDynamicSystem, newton_cg.rs:413-820)
    DevBuf<double> d_su, d_sv;       // state of set_state, Dirichlet DOFs zeroed
    bool state_set = false;
    bool solve_outdated = true;      // J and the preconditioner need rebuilding for solve
    bool has_alpha = false, has_beta = false;
    double alpha = 0.0, beta = 0.0;  // init_solve_jacobian_combination
    bool apply_outdated = true, apply_ready = false;
    bool ap_has[3] = {false, false, false};
    double ap_coef[3] = {0.0, 0.0, 0.0};  // init_apply_jacobian_combination (alpha, beta, gamma)
    DevBuf<double> d_H;              // the combination applied by apply_jacobian_combination
    // Dirichlet values (DOF vectors, zero off the Dirichlet DOFs); empty: homogeneous conditions
    DevBuf<double> d_bcu, d_bcv;
    bool has_bcu = false, has_bcv = false;
    const double* bcu() const { return has_bcu ? d_bcu.get() : nullptr; }
    const double* bcv() const { return has_bcv ? d_bcv.get() : nullptr; }
};

static void build_bc(Model& m, const std::vector<int32_t>& dir, DevBuf<uint8_t>& dmask, DevBuf<int32_t>& dbc,
                     int64_t& nbc) {
    std::vector<uint8_t> mask(m.nv, 0);
    for (int32_t v : dir) {
        if (v < 0 || v >= m.nv) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "Dirichlet node out of range");
        mask[v] = 1;
    }
    // rows to touch: Dirichlet nodes and their pattern neighbours
    std::vector<uint8_t> touch(m.nv, 0);
    for (int32_t v : dir)
        for (int32_t k = m.pattern.nptr[v]; k < m.pattern.nptr[v + 1]; ++k) touch[m.pattern.ncol[k]] = 1;
    for (int32_t v : dir) touch[v] = 1;
    std::vector<int32_t> bc;
    for (int64_t v = 0; v < m.nv; ++v)
        if (touch[v]) bc.push_back((int32_t)v);
    nbc = (int64_t)bc.size();
    dmask.upload(mask, m.stream);
    dbc.upload(bc, m.stream);
}

static void sim_create(Sim& sref, Model& m, const hoefem_sim_desc* d) {
    Sim* s = &sref;
    s->model = &m;
    s->d = *d;
    dev::snh_reparam(d->mu, d->lambda, &s->mu_r, &s->lambda_r);
    s->dir.assign(d->dirichlet_nodes, d->dirichlet_nodes + d->num_dirichlet);
    build_bc(m, s->dir, s->d_dirmask, s->d_bc_nodes, s->n_bc);
    const int64_t n = 3 * m.nv, nnz = m.pattern.nnz();
    for (DevBuf<double>* b : {&s->d_gforce, &s->d_up, &s->d_vn, &s->d_w, &s->d_F, &s->d_sol, &s->d_uu, &s->d_us,
                              &s->d_vs, &s->d_f, &s->d_su, &s->d_sv})
        b->ensure(n);
    if (d->mixed_precision < 0 || d->mixed_precision > 3)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "mixed_precision must be 0, 1, 2 or 3");
    if (!(d->mass_damping_coefficient >= 0.0 && std::isfinite(d->mass_damping_coefficient)) ||
        !(d->stiffness_damping_coefficient >= 0.0 && std::isfinite(d->stiffness_damping_coefficient)))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "damping coefficients must be finite and >= 0");
    s->damping = d->mass_damping_coefficient > 0.0 || d->stiffness_damping_coefficient > 0.0;
    s->d_vv.ensure(n);
    if (s->damping) {
        s->d_D.ensure(nnz);
        s->d_dv.ensure(n);
        HIP_CHECK(hipMemsetAsync(s->d_D.get(), 0, sizeof(double) * nnz, m.stream));
    }
    s->mixed = d->mixed_precision != 0;
    s->mixed_mode = d->mixed_precision;
    s->d_M.ensure(nnz);
    s->d_J.ensure(nnz);
    if (s->mixed) {
        s->d_J32.ensure(nnz);
        s->d_res.ensure(n);
        s->d_dx.ensure(n);
    }
    s->d_scale.ensure(1);
    s->d_norm2.ensure(1);
    // mass matrix with Dirichlet conditions (fe_model.rs:490-515)
    launch_element_mass(m, d->density);
    launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), s->d_M.get(), 0, 0.0, nullptr, true, m.stream);
    launch_dirichlet(m.pattern, s->d_M.get(), s->d_dirmask.get(), s->d_bc_nodes.get(), s->n_bc, s->d_scale.get(), m.stream);
    // representative force |M a_rep|, a_rep = sqrt(100/3) (fe_model.rs:423-431), and M g
    std::vector<double> arep(n, std::sqrt((10.0 * 10.0) / 3.0)), gd(n, 0.0);
    if (d->gravity_enabled)
        for (int64_t i = 0; i < n; ++i)
            if ((i % 3) + 1 == 3) gd[i] = d->gravity;
    s->d_up.upload(arep, m.stream);
    launch_spmv(m.pattern, s->d_M.get(), 0.0, s->d_F.get(), 1.0, s->d_up.get(), m.stream);
    std::vector<double> Ma(n);
    HIP_CHECK(hipMemcpyAsync(Ma.data(), s->d_F.get(), sizeof(double) * n, hipMemcpyDeviceToHost, m.stream));
    s->d_vn.upload(gd, m.stream);
    launch_spmv(m.pattern, s->d_M.get(), 0.0, s->d_gforce.get(), 1.0, s->d_vn.get(), m.stream);
    HIP_CHECK(hipStreamSynchronize(m.stream));
    double ss = 0.0;
    for (double v : Ma) ss += v * v;
    s->rep_force = std::sqrt(ss);
    if (d->num_subdomains > 0) {
        s->schwarz.reset(new Schwarz());
        s->schwarz->fp32 = s->mixed;
        schwarz_setup(*s->schwarz, m, d->num_subdomains, d->subdomain_offsets, d->subdomain_nodes, d->num_untreated,
                      d->untreated_nodes);
    }
    raise_kernel_errors(m);
}

// f = M g + f_int(uu) - D vv with Dirichlet rows zeroed, uu and vv already carrying the Dirichlet
// values (NewtonCgDynamicSystem::eval_f, newton_cg.rs:441-478)
static void sim_eval_f(Sim& s, const double* uu, const double* vv, double* f) {
    Model& m = *s.model;
    m.prof_tag = PROF_ELEMENT_FORCES;
    launch_element_forces(m, s.mu_r, s.lambda_r, uu);
    m.prof_tag = -1;
    if (s.damping) launch_spmv(m.pattern, s.d_D.get(), 0.0, s.d_dv.get(), 1.0, vv, m.stream);
    m.prof_tag = PROF_FORCE_GATHER;
    launch_node_force(m, s.d_gforce.get(), s.damping ? s.d_dv.get() : nullptr, s.d_dirmask.get(), f);
    m.prof_tag = -1;
}

// F(w) = -dt f(u_p + dt w, v_n + w) + M w and |F|^2 for the current (u_p, v_n) held in the sim
// (BackwardEulerImplicitFunction::eval_into, hamilton2/src/integrators/euler.rs:97-114), composed of
// the same kernels as the dynamic-system entry points eval_f and apply_mass_add
// (sim_eval_launch enqueues it with the readback of |F|^2 and the error flags; sim_eval_finish reads
// them once the model stream has been synchronized)
static void sim_eval_launch(Sim& s, double dt, const double* w) {
    Model& m = *s.model;
    launch_state(m, s.d_up.get(), w, dt, s.d_dirmask.get(), s.bcu(), s.d_uu.get());
    if (s.damping) launch_state_v(m, s.d_vn.get(), w, s.d_dirmask.get(), s.bcv(), s.d_vv.get());
    sim_eval_f(s, s.d_uu.get(), s.d_vv.get(), s.d_f.get());
    launch_scale(m, s.d_f.get(), dt * -1.0, s.d_F.get());
    launch_mass_add(m, s.d_M.get(), w, s.d_F.get(), s.cg, s.d_norm2.get());
    HIP_CHECK(hipMemcpyAsync(m.h_err.p, m.d_err.get(), sizeof(int), hipMemcpyDeviceToHost, m.stream));
    HIP_CHECK(hipMemcpyAsync(s.h_norm2.p, s.d_norm2.get(), sizeof(double), hipMemcpyDeviceToHost, m.stream));
}

static double sim_eval_finish(Sim& s) {
    Model& m = *s.model;
    s.evals++;
    if (*m.h_err.p) {
        s.last_err_flags |= *m.h_err.p;
        HIP_CHECK(hipMemsetAsync(m.d_err.get(), 0, sizeof(int), m.stream));
        return NAN;  // a singular F makes the residual unusable; the line search then fails
    }
    return *s.h_norm2.p;
}

static double sim_eval(Sim& s, double dt, const double* w) {
    sim_eval_launch(s, dt, w);
    HIP_CHECK(hipStreamSynchronize(s.model->stream));
    return sim_eval_finish(s);
}

// compute_jacobian_combination_into (simulation_toolbox/src/fem/system_assembly.rs:196-235) at the
// state uu (Dirichlet values applied): J = M - alpha K(uu) - beta D, a term left out when its
// coefficient is None, then the Dirichlet conditions and SchwarzPreconditioner::update
// (newton_cg.rs:700-728). K is the unprojected stiffness. Mixed precision: mode 1 builds J in FP32
// from FP32 element tangents; mode 2 builds it in FP64 and keeps an FP32 copy for the CG.
// (sim_assemble_jacobian_launch enqueues it with the readback of the error flags, and
// sim_assemble_jacobian_finish reads them once the model stream has been synchronized)
static void sim_assemble_jacobian_launch(Sim& s, bool has_a, double a, bool has_b, double b, const double* uu) {
    Model& m = *s.model;
    const int64_t nnz = m.pattern.nnz();
    const double* D = (has_b && s.damping) ? s.d_D.get() : nullptr;
    const double cd = has_b ? -b : 0.0;
    KernelProf* gprof = m.prof ? &m.kprof[PROF_JACOBIAN_GATHER] : nullptr;
    if (has_a) {
        m.prof_tag = PROF_TANGENT;
        launch_element_stiffness(m, s.mu_r, s.lambda_r, uu, s.mixed_mode == 1);
        m.prof_tag = -1;
    }
    if (s.fp32_jacobian()) {
        if (has_a && s.mixed_mode == 1)
            launch_block_gather(m.pattern, m.scatter, m.d_ke32.get(), s.d_J32.get(), 1, -a, s.d_M.get(), true,
                                m.stream, D, cd, 1.0, gprof);
        else if (has_a)  // mode 3: each entry is summed in FP64 and rounded once
            launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), s.d_J32.get(), 1, -a, s.d_M.get(), true,
                                m.stream, D, cd, 1.0, gprof);
        else
            launch_comb(m, nnz, s.d_J32.get(), cd, D, 1.0, s.d_M.get());
        launch_dirichlet(m.pattern, s.d_J32.get(), s.d_dirmask.get(), s.d_bc_nodes.get(), s.n_bc, s.d_scale.get(),
                         m.stream);
        if (s.schwarz) schwarz_update(*s.schwarz, s.d_J32.get());
    } else {
        if (has_a)
            launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), s.d_J.get(), 1, -a, s.d_M.get(), true, m.stream,
                                D, cd, 1.0, gprof);
        else
            launch_comb(m, nnz, s.d_J.get(), cd, D, 1.0, s.d_M.get());
        launch_dirichlet(m.pattern, s.d_J.get(), s.d_dirmask.get(), s.d_bc_nodes.get(), s.n_bc, s.d_scale.get(),
                         m.stream);
        // mode 2: the FP64 Jacobian (refinement residual, Schwarz factorizations) and an FP32 copy of
        // its values for the CG operator; P is stored in FP32
        if (s.mixed_mode == 2) launch_f64_to_f32(m, s.d_J.get(), s.d_J32.get(), nnz);
        if (s.schwarz) schwarz_update(*s.schwarz, s.d_J.get());
    }
    HIP_CHECK(hipMemcpyAsync(m.h_err.p, m.d_err.get(), sizeof(int), hipMemcpyDeviceToHost, m.stream));
}

static hoefem_status sim_assemble_jacobian_finish(Sim& s) {
    Model& m = *s.model;
    const int e = *m.h_err.p;
    if (e) {
        HIP_CHECK(hipMemsetAsync(m.d_err.get(), 0, sizeof(int), m.stream));
        s.last_err_flags |= e;
        return HOEFEM_ERROR_SINGULAR;
    }
    return HOEFEM_OK;
}

static hoefem_status sim_assemble_jacobian(Sim& s, bool has_a, double a, bool has_b, double b, const double* uu) {
    sim_assemble_jacobian_launch(s, has_a, a, has_b, b, uu);
    HIP_CHECK(hipStreamSynchronize(s.model->stream));
    return sim_assemble_jacobian_finish(s);
}

// The CG part of NewtonCgDynamicSystem::solve_jacobian_combination (newton_cg.rs:730-801): PCG from
// a zero guess on the assembled J with the updated preconditioner. MaxIterationsReached and
// IndefiniteOperator are tolerated (newton_cg.rs:776-797); any other CG status is returned.
static hoefem_status sim_cg(Sim& s, const double* rhs, double* sol, int64_t* iterations) {
    Model& m = *s.model;
    HIP_CHECK(hipMemsetAsync(sol, 0, sizeof(double) * 3 * m.nv, m.stream));
    int64_t it = 0;
    *iterations = 0;
    Learned* L = s.learned;
    if (L && L->enabled && L->trained && s.schwarz) {
        // learned flexible PCG. Its result is used under the plain path's rules (reaching the
        // iteration limit is tolerated, NewtonCgDynamicSystem::solve_jacobian_combination,
        // simulation_toolbox/src/fem/newton_cg.rs:680-802), and the Newton loop still has to
        // reach the same residual tolerance; on a breakdown (an indefinite or non-finite
        // preconditioned residual) the plain Schwarz-CG solves the system from zero.
        const LearnedView lv = learned_view(*L);
        const hoefem_status lst =
            s.mixed ? pcg_solve_learned(m, s.cg, s.d_J32.get(), s.schwarz.get(), lv, rhs, sol, s.d.cg_tolerance,
                                        s.d.cg_max_iter, &it)
                    : pcg_solve_learned(m, s.cg, s.d_J.get(), s.schwarz.get(), lv, rhs, sol, s.d.cg_tolerance,
                                        s.d.cg_max_iter, &it);
        L->solves += 1;
        L->iterations += it;
        L->last_status = (int)lst;
        s.cg_total += it;
        *iterations += it;
        if (lst == HOEFEM_OK || lst == HOEFEM_ERROR_CG_MAX_ITERATIONS) return HOEFEM_OK;
        L->fallbacks += 1;
        HIP_CHECK(hipMemsetAsync(sol, 0, sizeof(double) * 3 * m.nv, m.stream));
        it = 0;
    }
    hoefem_status st;
    if (L && L->harvesting) {
        CgHarvest h;
        h.xs = L->d_snap_x.get();
        h.rs = L->d_snap_r.get();
        h.max = L->snap_max;
        h.every = L->every;
        st = s.mixed ? pcg_solve_harvest(m, s.cg, s.d_J32.get(), s.schwarz.get(), rhs, sol, s.d.cg_tolerance,
                                         s.d.cg_max_iter, &it, &h)
                     : pcg_solve_harvest(m, s.cg, s.d_J.get(), s.schwarz.get(), rhs, sol, s.d.cg_tolerance,
                                         s.d.cg_max_iter, &it, &h);
        learned_harvest_solve(*L, sol, h.taken);
    } else if (s.mixed_mode == 2) {
        // iterative refinement: the CG on the FP32 operator solves for a correction, the residual
        // is recomputed in FP64 (rhs - J sol) until it meets the same relative tolerance as the
        // plain solve; the CG iterations of all passes share the plain solve's iteration limit
        const int64_t n = 3 * m.nv;
        const double bnorm = std::sqrt(device_dot(m, s.cg, rhs, rhs, s.d_norm2.get(), s.h_norm2.p));
        HIP_CHECK(hipMemcpyAsync(s.d_res.get(), rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
        st = HOEFEM_OK;
        for (int pass = 0; pass < 8; ++pass) {
            int64_t ip = 0;
            HIP_CHECK(hipMemsetAsync(s.d_dx.get(), 0, sizeof(double) * n, m.stream));
            const int64_t budget = s.d.cg_max_iter >= 0 ? std::max<int64_t>(1, s.d.cg_max_iter - it) : -1;
            st = pcg_solve(m, s.cg, s.d_J32.get(), s.schwarz.get(), s.d_res.get(), s.d_dx.get(), s.d.cg_tolerance,
                           budget, &ip);
            it += ip;
            launch_add(m, sol, s.d_dx.get(), n);
            if (st != HOEFEM_OK && st != HOEFEM_ERROR_CG_MAX_ITERATIONS) break;
            HIP_CHECK(hipMemcpyAsync(s.d_res.get(), rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
            launch_spmv(m.pattern, s.d_J.get(), 1.0, s.d_res.get(), -1.0, sol, m.stream);
            s.refinements += 1;
            const double rn = std::sqrt(device_dot(m, s.cg, s.d_res.get(), s.d_res.get(), s.d_norm2.get(), s.h_norm2.p));
            if (rn <= s.d.cg_tolerance * bnorm || bnorm == 0.0) {
                st = HOEFEM_OK;
                break;
            }
            if (s.d.cg_max_iter >= 0 && it >= s.d.cg_max_iter) {
                st = HOEFEM_ERROR_CG_MAX_ITERATIONS;
                break;
            }
        }
    } else if (s.fp32_jacobian()) {
        st = pcg_solve(m, s.cg, s.d_J32.get(), s.schwarz.get(), rhs, sol, s.d.cg_tolerance, s.d.cg_max_iter, &it);
    } else {
        st = pcg_solve(m, s.cg, s.d_J.get(), s.schwarz.get(), rhs, sol, s.d.cg_tolerance, s.d.cg_max_iter, &it);
    }
    s.cg_total += it;
    *iterations += it;
    if (st == HOEFEM_ERROR_CG_MAX_ITERATIONS || st == HOEFEM_ERROR_CG_INDEFINITE_OPERATOR) st = HOEFEM_OK;
    return st;
}

// NewtonCgDynamicSystem::solve_jacobian_combination (newton_cg.rs:680-802) inside the Newton loop:
// state u_p + dt w, (alpha, beta) = (-dt^2, -dt) (BackwardEulerImplicitFunction::solve_jacobian_system,
// euler.rs:122-140), then the CG.
static hoefem_status sim_solve(Sim& s, double dt, const double* w, const double* rhs, double* sol) {
    Model& m = *s.model;
    launch_state(m, s.d_up.get(), w, dt, s.d_dirmask.get(), s.bcu(), s.d_uu.get());
    if (sim_assemble_jacobian(s, true, -dt * dt, true, -dt, s.d_uu.get()) != HOEFEM_OK) return HOEFEM_ERROR_SINGULAR;
    int64_t it = 0;
    return sim_cg(s, rhs, sol, &it);
}

// backward_euler_step (hamilton2/src/integrators/euler.rs:155-199) with newton_line_search and
// BacktrackingLineSearch (hamilton2/src/newton.rs:76-131, :171-246). u, v: device state.
// compute_damping_matrix_into (simulation_toolbox/src/fem/system_assembly.rs:145-194) at the
// (sub)step's initial displacement: D = 0 + a_K K+(u) + a_M M, K+ without boundary conditions and
// with every element matrix projected onto the positive semi-definite cone unless the material
// is positive semi-definite by construction (ProjectedStableNeoHookean).
static hoefem_status sim_damping(Sim& s, const double* u) {
    Model& m = *s.model;
    const double aK = s.d.stiffness_damping_coefficient, aM = s.d.mass_damping_coefficient;
    const int64_t nnz = m.pattern.nnz();
    if (aK > 0.0) {
        launch_element_stiffness(m, s.mu_r, s.lambda_r, u, false);
        if (m.material != dev::MAT_PROJECTED_SNH) launch_element_psd_projection(m);
        launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), s.d_D.get(), 1, aK, aM > 0.0 ? s.d_M.get() : nullptr,
                            true, m.stream, nullptr, 0.0, aM);
    } else {
        launch_lincomb(m.stream, nnz, s.d_D.get(), aM, s.d_M.get());
    }
    const int e = read_and_clear_err(m);
    if (e) {
        s.last_err_flags |= e;
        return HOEFEM_ERROR_SINGULAR;
    }
    return HOEFEM_OK;
}

static hoefem_status sim_be_step(Sim& s, double dt, double* u, double* v, int64_t* newton_iters, double* final_res,
                                 double* tol_out) {
    Model& m = *s.model;
    if (s.damping) {
        const hoefem_status ds = sim_damping(s, u);
        if (ds != HOEFEM_OK) return ds;
    }
    launch_be_begin(m, u, v, dt, s.d_up.get(), s.d_vn.get());
    HIP_CHECK(hipMemsetAsync(s.d_w.get(), 0, sizeof(double) * 3 * m.nv, m.stream));
    const double tol = dt * (s.d.newton_tolerance_factor * s.rep_force);
    *tol_out = tol;
    double n2 = sim_eval(s, dt, s.d_w.get());
    int64_t it = 0;
    *newton_iters = 0;
    while (true) {
        const double fnorm = std::sqrt(n2);
        *final_res = fnorm;
        if (!(fnorm > tol)) break;
        if (s.d.max_newton_iter >= 0 && it == s.d.max_newton_iter) {
            *newton_iters = it;
            return HOEFEM_ERROR_NEWTON_MAX_ITERATIONS;
        }
        // solve J (-dx) = F
        hoefem_status st = sim_solve(s, dt, s.d_w.get(), s.d_F.get(), s.d_sol.get());
        if (st != HOEFEM_OK) {
            *newton_iters = it;
            return HOEFEM_ERROR_JACOBIAN;
        }
        // backtracking on g = 0.5 |F|^2 along p = -sol
        const double c = 1e-4, alpha_min = 1e-6;
        const double g_initial = 0.5 * n2;
        double alpha_prev = 0.0, alpha = 1.0, tail = 0.25;
        const double initial[3] = {1.0, 0.75, 0.5};
        int idx = 0;
        while (true) {
            launch_axpy_neg(m, s.d_w.get(), alpha - alpha_prev, s.d_sol.get());
            n2 = sim_eval(s, dt, s.d_w.get());
            const double g = 0.5 * n2;
            if (g <= (1.0 - c * alpha) * g_initial) break;
            if (alpha < alpha_min) {
                *newton_iters = it;
                return HOEFEM_ERROR_LINE_SEARCH;
            }
            alpha_prev = alpha;
            ++idx;
            if (idx < 3) {
                alpha = initial[idx];
            } else {
                alpha = tail;
                tail *= 0.25;
            }
        }
        ++it;
    }
    *newton_iters = it;
    launch_be_end(m, u, v, s.d_w.get(), dt, s.d_dirmask.get(), s.bcu(), s.bcv());
    return HOEFEM_OK;
}

// integrate_model_with_substeps (simulation_toolbox/src/fem/newton_cg.rs:146-228), in parts: the
// step's start (counters, u and v with the Dirichlet DOFs zeroed copied to the working state), the
// attempts from `num_substeps` substeps on, and the end (finite check, working state back to u, v).
static void sim_step_begin(Sim& s, double* u, double* v, hoefem_step_stats& st) {
    Model& m = *s.model;
    const int64_t n = 3 * m.nv;
    st = hoefem_step_stats{};
    s.cg_total = 0;
    s.evals = 0;
    s.last_err_flags = 0;
    launch_set_dirichlet(m, u, v, s.d_dirmask.get(), s.bcu(), s.bcv());
    HIP_CHECK(hipMemcpyAsync(s.d_us.get(), u, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
    HIP_CHECK(hipMemcpyAsync(s.d_vs.get(), v, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
}

// the substep count after an attempt with num_substeps failed (last: its error); throws when it
// reaches the limit, else restores the working state for the next attempt
static int sim_next_substeps(Sim& s, const double* u, const double* v, hoefem_step_stats& st, int num_substeps,
                             hoefem_status last) {
    Model& m = *s.model;
    const int64_t n = 3 * m.nv;
    const int max_substeps = 1000, factor = 5;
    num_substeps *= factor;
    if (num_substeps >= max_substeps) {
        st.cg_iterations = s.cg_total;
        st.residual_evaluations = s.evals;
        throw Error(HOEFEM_ERROR_SUBSTEPPING, "Substepping failed; last error " + std::to_string((int)last));
    }
    HIP_CHECK(hipMemcpyAsync(s.d_us.get(), u, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
    HIP_CHECK(hipMemcpyAsync(s.d_vs.get(), v, sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
    return num_substeps;
}

static void sim_substep_loop(Sim& s, double dt, const double* u, const double* v, hoefem_step_stats& st,
                             int num_substeps) {
    while (true) {
        const double sdt = dt / (double)num_substeps;
        hoefem_status last = HOEFEM_OK;
        for (int k = 0; k < num_substeps && last == HOEFEM_OK; ++k) {
            int64_t it = 0;
            last = sim_be_step(s, sdt, s.d_us.get(), s.d_vs.get(), &it, &st.final_residual, &st.tolerance);
            st.newton_iterations += it;
        }
        if (last == HOEFEM_OK) break;
        num_substeps = sim_next_substeps(s, u, v, st, num_substeps, last);
    }
    st.substeps = num_substeps;
}

static void sim_step_end_launch(Sim& s, hoefem_step_stats& st) {
    Model& m = *s.model;
    st.cg_iterations = s.cg_total;
    st.residual_evaluations = s.evals;
    launch_finite_check(m, s.d_us.get(), s.d_vs.get());
    HIP_CHECK(hipMemcpyAsync(m.h_err.p, m.d_err.get(), sizeof(int), hipMemcpyDeviceToHost, m.stream));
}

// after the model stream has been synchronized: the finite check's verdict, then the copy back
static void sim_step_end_finish(Sim& s, double* u, double* v) {
    Model& m = *s.model;
    const int64_t n = 3 * m.nv;
    const int e = *m.h_err.p;
    if (e) HIP_CHECK(hipMemsetAsync(m.d_err.get(), 0, sizeof(int), m.stream));
    if (e & KERR_NONFINITE) throw Error(HOEFEM_ERROR_NOT_FINITE, "non-finite state after step");
    HIP_CHECK(hipMemcpyAsync(u, s.d_us.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
    HIP_CHECK(hipMemcpyAsync(v, s.d_vs.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
}

static hoefem_status sim_step(Sim& s, double t, double dt, double* u, double* v, hoefem_step_stats* stats) {
    Model& m = *s.model;
    hoefem_step_stats local{};
    hoefem_step_stats& st = stats ? *stats : local;
    (void)t;
    sim_step_begin(s, u, v, st);
    sim_substep_loop(s, dt, u, v, st, 1);
    sim_step_end_launch(s, st);
    HIP_CHECK(hipStreamSynchronize(m.stream));
    sim_step_end_finish(s, u, v);
    HIP_CHECK(hipStreamSynchronize(m.stream));
    return HOEFEM_OK;
}

// ------------------------------------------------------------------ batched scenes
// Independent scenes advanced by one time step together. Every scene runs the kernels of sim_step
// in the same order with the same arguments, on its own model stream; the batch changes only how
// they are issued. One host thread drives every scene's Newton loop as a state machine: it enqueues a
// scene's next piece of work (residual evaluation, Jacobian assembly, line-search trial), records an
// event and moves on, and takes the scene's next decision (Newton test, line-search acceptance, the
// same ones sim_be_step takes) when that event has completed. A scene whose Jacobian is ready joins
// the CG pool, whose batched launches carry every running solve and which reports each solve as it
// ends on the device. Scenes therefore never wait for each other, and each scene's u and v are the
// ones sim_step gives it, bit for bit. A scene whose single-step attempt fails continues alone with
// the substepping of sim_step once the others are done.
// Scenes go to CG pools by size (DOFs below 2e4, below 8e4, larger), each on its own stream, so the
// iterations of small systems do not wait for the launches the large ones lengthen.
constexpr int BATCH_POOLS = 3;
constexpr int64_t BATCH_POOL_DOFS[BATCH_POOLS - 1] = {20000, 80000};

struct Batch {
    std::vector<Sim*> sims;
    CgBatch cg[BATCH_POOLS];
    std::vector<hipEvent_t> ev;  // one per scene
    ~Batch() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

namespace {
enum BatchPhase { BP_EVAL, BP_JACOBIAN, BP_CG, BP_TRIAL, BP_DONE, BP_FAILED, BP_ERROR };
struct BatchScene {
    Sim* s = nullptr;
    double dt = 0.0;
    double* u = nullptr;
    double* v = nullptr;
    hoefem_step_stats st{};
    BatchPhase phase = BP_EVAL;
    bool waiting = false;  // its event is pending
    hoefem_status status = HOEFEM_OK;  // BP_FAILED: the attempt's error; BP_ERROR: the step's
    std::string message;
    double n2 = 0.0, tol = 0.0, g_initial = 0.0, alpha = 1.0, alpha_prev = 0.0, tail = 0.25;
    int idx = 0;
    int64_t it = 0;
    int cg_id = -1;
    double latency_ms = 0.0;
};
}  // namespace

static void batch_step(Batch& B, const double* dt, double* const* u, double* const* v, hoefem_step_stats* stats,
                       double* latency_ms, int32_t* status) {
    const size_t ns = B.sims.size();
    std::vector<hipStream_t> streams;
    for (Sim* s : B.sims) {
        if (s->learned) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "a batched scene has a learned correction attached");
        if (std::find(streams.begin(), streams.end(), s->model->stream) == streams.end())
            streams.push_back(s->model->stream);
    }
    while (B.ev.size() < ns) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        B.ev.push_back(e);
    }
    const bool fp32 = B.sims[0]->fp32_jacobian();
    std::vector<int> pool_of(ns);
    int pool_size[BATCH_POOLS] = {};
    for (size_t i = 0; i < ns; ++i) {
        const int64_t ndof = 3 * B.sims[i]->model->nv;
        int c = 0;
        while (c < BATCH_POOLS - 1 && ndof >= BATCH_POOL_DOFS[c]) ++c;
        pool_of[i] = c;
        pool_size[c] += 1;
    }
    std::unique_ptr<CgPool> pools[BATCH_POOLS];
    for (int c = 0; c < BATCH_POOLS; ++c)
        if (pool_size[c]) pools[c] = make_cg_pool(B.cg[c], fp32, pool_size[c]);
    std::vector<BatchScene> q(ns);
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed_ms = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    auto mark = [&](size_t i) {
        HIP_CHECK(hipEventRecord(B.ev[i], q[i].s->model->stream));
        q[i].waiting = true;
    };
    auto fail = [&](BatchScene& b, hoefem_status e) {
        b.status = e;
        b.phase = BP_FAILED;
        b.st.newton_iterations += b.it;
    };
    auto error = [&](BatchScene& b, const Error& e) {
        b.status = e.code;
        b.message = e.what();
        b.phase = BP_ERROR;
        b.latency_ms = elapsed_ms();
    };
    // the Newton test of a scene whose last step was accepted (newton.rs:76-131), then its next work
    auto newton = [&](size_t i) {
        BatchScene& b = q[i];
        Sim& s = *b.s;
        const double fnorm = std::sqrt(b.n2);
        b.st.final_residual = fnorm;
        if (!(fnorm > b.tol)) {
            b.phase = BP_DONE;
            b.st.newton_iterations += b.it;
            b.st.substeps = 1;
            launch_be_end(*s.model, s.d_us.get(), s.d_vs.get(), s.d_w.get(), b.dt, s.d_dirmask.get(), s.bcu(),
                          s.bcv());
            b.latency_ms = elapsed_ms();
            return;
        }
        if (s.d.max_newton_iter >= 0 && b.it == s.d.max_newton_iter) {
            fail(b, HOEFEM_ERROR_NEWTON_MAX_ITERATIONS);
            return;
        }
        // J (-dx) = F (sim_solve): the Jacobian (and the Jacobi diagonal), then the pool's CG
        launch_state(*s.model, s.d_up.get(), s.d_w.get(), b.dt, s.d_dirmask.get(), s.bcu(), s.d_uu.get());
        sim_assemble_jacobian_launch(s, true, -b.dt * b.dt, true, -b.dt, s.d_uu.get());
        if (!s.schwarz) {
            launch_jacobi_diag(*s.model, s.cg, fp32 ? (const void*)s.d_J32.get() : (const void*)s.d_J.get(), fp32);
            HIP_CHECK(hipMemcpyAsync(s.model->h_err.p, s.model->d_err.get(), sizeof(int), hipMemcpyDeviceToHost,
                                     s.model->stream));
        }
        b.phase = BP_JACOBIAN;
        mark(i);
    };
    // the next line-search trial of a scene (BacktrackingLineSearch, newton.rs:171-246)
    auto trial = [&](size_t i) {
        BatchScene& b = q[i];
        Sim& s = *b.s;
        launch_axpy_neg(*s.model, s.d_w.get(), b.alpha - b.alpha_prev, s.d_sol.get());
        sim_eval_launch(s, b.dt, s.d_w.get());
        b.phase = BP_TRIAL;
        mark(i);
    };
    // start of sim_step and of sim_be_step for one substep of dt
    for (size_t i = 0; i < ns; ++i) {
        BatchScene& b = q[i];
        b.s = B.sims[i];
        b.dt = dt[i];
        b.u = u[i];
        b.v = v[i];
        Sim& s = *b.s;
        Model& m = *s.model;
        if (!b.u || !b.v) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null state vector");
        sim_step_begin(s, b.u, b.v, b.st);
        if (s.damping) {
            const hoefem_status ds = sim_damping(s, s.d_us.get());
            if (ds != HOEFEM_OK) {
                fail(b, ds);
                continue;
            }
        }
        launch_be_begin(m, s.d_us.get(), s.d_vs.get(), b.dt, s.d_up.get(), s.d_vn.get());
        HIP_CHECK(hipMemsetAsync(s.d_w.get(), 0, sizeof(double) * 3 * m.nv, m.stream));
        b.tol = b.dt * (s.d.newton_tolerance_factor * s.rep_force);
        b.st.tolerance = b.tol;
        sim_eval_launch(s, b.dt, s.d_w.get());
        b.phase = BP_EVAL;
        mark(i);
    }
    std::vector<int> ended;
    std::vector<int> owner[BATCH_POOLS];  // scene of every pool system
    const double c = 1e-4, alpha_min = 1e-6;
    const double initial[3] = {1.0, 0.75, 0.5};
    while (true) {
        for (size_t i = 0; i < ns; ++i) {
            BatchScene& b = q[i];
            if (!b.waiting) continue;
            const hipError_t qe = hipEventQuery(B.ev[i]);
            if (qe == hipErrorNotReady) continue;
            HIP_CHECK(qe);
            b.waiting = false;
            Sim& s = *b.s;
            if (b.phase == BP_EVAL) {
                b.n2 = sim_eval_finish(s);
                newton(i);
            } else if (b.phase == BP_JACOBIAN) {
                const int e = *s.model->h_err.p;
                if (e) {
                    HIP_CHECK(hipMemsetAsync(s.model->d_err.get(), 0, sizeof(int), s.model->stream));
                    s.last_err_flags |= e;
                }
                // without Schwarz, a zero diagonal comes from the Jacobi setup, which pcg_solve reports
                // by an exception; any other flag (a Schwarz zero diagonal included) fails the assembly
                const int jac_err = s.schwarz ? 0 : (e & KERR_ZERO_DIAG);
                if (e & ~jac_err) {
                    fail(b, HOEFEM_ERROR_JACOBIAN);
                } else if (jac_err) {
                    error(b, Error(HOEFEM_ERROR_SINGULAR, "Diagonal matrix is not invertible."));
                } else {
                    CgBatchSystem sys;
                    sys.m = s.model;
                    sys.w = &s.cg;
                    sys.A = fp32 ? (const void*)s.d_J32.get() : (const void*)s.d_J.get();
                    sys.P = s.schwarz.get();
                    sys.b = s.d_F.get();
                    sys.x = s.d_sol.get();
                    sys.tol = s.d.cg_tolerance;
                    sys.max_iter = s.d.cg_max_iter;
                    const int pc = pool_of[i];
                    b.cg_id = pools[pc]->join(sys);
                    if ((int)owner[pc].size() <= b.cg_id) owner[pc].resize(b.cg_id + 1, -1);
                    owner[pc][b.cg_id] = (int)i;
                    b.phase = BP_CG;
                }
            } else if (b.phase == BP_TRIAL) {
                b.n2 = sim_eval_finish(s);
                const double g = 0.5 * b.n2;
                if (g <= (1.0 - c * b.alpha) * b.g_initial) {
                    ++b.it;
                    newton(i);
                } else if (b.alpha < alpha_min) {
                    fail(b, HOEFEM_ERROR_LINE_SEARCH);
                } else {
                    b.alpha_prev = b.alpha;
                    ++b.idx;
                    if (b.idx < 3) {
                        b.alpha = initial[b.idx];
                    } else {
                        b.alpha = b.tail;
                        b.tail *= 0.25;
                    }
                    trial(i);
                }
            }
        }
        bool any_ended = false;
        for (int pc = 0; pc < BATCH_POOLS; ++pc) {
            if (!pools[pc]) continue;
            ended.clear();
            pools[pc]->progress(ended);
            any_ended |= !ended.empty();
            for (int id : ended) {
                const size_t i = (size_t)owner[pc][id];
                BatchScene& b = q[i];
                const CgBatchSystem& r = pools[pc]->system(id);
                b.s->cg_total += r.iterations;
                hoefem_status cs = r.status;
                if (cs == HOEFEM_ERROR_CG_MAX_ITERATIONS || cs == HOEFEM_ERROR_CG_INDEFINITE_OPERATOR) cs = HOEFEM_OK;
                if (cs != HOEFEM_OK) {
                    fail(b, HOEFEM_ERROR_JACOBIAN);
                    continue;
                }
                b.g_initial = 0.5 * b.n2;
                b.alpha_prev = 0.0;
                b.alpha = 1.0;
                b.tail = 0.25;
                b.idx = 0;
                trial(i);
            }
        }
        // the step goes on while any scene is still in its Newton loop (evaluating, assembling, in the
        // pool or trying a step length)
        bool active = false, host_pending = false;
        for (const BatchScene& b : q) {
            active |= b.phase == BP_EVAL || b.phase == BP_JACOBIAN || b.phase == BP_CG || b.phase == BP_TRIAL;
            host_pending |= b.waiting;
        }
        if (!active) break;
        // nothing for the host to do until the GPU moves on
        if (!any_ended && !host_pending) std::this_thread::yield();
    }
    // scenes whose attempt failed: the rest of sim_step's substepping, one scene at a time
    for (BatchScene& b : q) {
        if (b.phase != BP_FAILED) continue;
        Sim& s = *b.s;
        try {
            const int next = sim_next_substeps(s, b.u, b.v, b.st, 1, b.status);
            sim_substep_loop(s, b.dt, b.u, b.v, b.st, next);
            b.phase = BP_DONE;
            b.latency_ms = elapsed_ms();
        } catch (const Error& e) {
            error(b, e);
        }
    }
    for (BatchScene& b : q)
        if (b.phase == BP_DONE) sim_step_end_launch(*b.s, b.st);
    for (hipStream_t st : streams) HIP_CHECK(hipStreamSynchronize(st));
    for (BatchScene& b : q) {
        if (b.phase != BP_DONE) continue;
        try {
            sim_step_end_finish(*b.s, b.u, b.v);
        } catch (const Error& e) {
            error(b, e);
        }
    }
    for (hipStream_t st : streams) HIP_CHECK(hipStreamSynchronize(st));
    const BatchScene* first_error = nullptr;
    for (size_t i = 0; i < ns; ++i) {
        BatchScene& b = q[i];
        if (b.phase == BP_DONE) b.status = HOEFEM_OK;
        if (stats) stats[i] = b.st;
        if (latency_ms) latency_ms[i] = b.latency_ms;
        if (status) status[i] = (int32_t)b.status;
        if (b.phase == BP_ERROR && !first_error) first_error = &b;
    }
    if (first_error)
        throw Error(first_error->status, "scene " + std::to_string(first_error - q.data()) + ": " + first_error->message);
}

}  // namespace hoefem

// ====================================================================== C ABI
using namespace hoefem;

struct hoefem_model {
    Model m;
};
struct hoefem_schwarz {
    Schwarz s;
};
struct hoefem_sim {
    Sim s;
};
struct hoefem_learned {
    Learned L;
    Sim* sim = nullptr;
};

// The caller's vectors of one call. In HOEFEM_POINTERS_HOST mode they live in host memory and are
// staged through model-owned device buffers: inputs uploaded on the model stream, outputs copied back
// by finish(), which also waits for the call's work. In device mode the pointers pass through.
struct CallIo {
    Model& m;
    bool host;
    int slot = 0;
    struct Out {
        void* user;
        void* dev;
        size_t bytes;
    };
    std::vector<Out> outs;
    explicit CallIo(Model& mm) : m(mm), host(mm.pointer_mode == HOEFEM_POINTERS_HOST) {}
    void* stage(size_t bytes) {
        if (slot >= (int)(sizeof(m.io_stage) / sizeof(m.io_stage[0])))
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "too many host vectors in one call");
        DevBuf<char>& b = m.io_stage[slot++];
        b.ensure(std::max<size_t>(bytes, 1));
        return b.get();
    }
    template <class T>
    const T* in(const T* p, size_t n) {
        if (!host || !p) return p;
        void* d = stage(n * sizeof(T));
        if (n) HIP_CHECK(hipMemcpyAsync(d, p, n * sizeof(T), hipMemcpyHostToDevice, m.stream));
        return (const T*)d;
    }
    template <class T>
    T* inout(T* p, size_t n) {
        if (!host || !p) return p;
        T* d = (T*)in<T>(p, n);
        outs.push_back({p, d, n * sizeof(T)});
        return d;
    }
    template <class T>
    T* out(T* p, size_t n) {
        if (!host || !p) return p;
        T* d = (T*)stage(n * sizeof(T));
        outs.push_back({p, d, n * sizeof(T)});
        return d;
    }
    void finish() {
        for (const Out& o : outs)
            if (o.bytes) HIP_CHECK(hipMemcpyAsync(o.user, o.dev, o.bytes, hipMemcpyDeviceToHost, m.stream));
        if (host) HIP_CHECK(hipStreamSynchronize(m.stream));
    }
};

template <class F>
static hoefem_status guarded(F&& f) {
    try {
        f();
        return HOEFEM_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HOEFEM_ERROR_DEVICE;
    } catch (...) {
        g_last_error = "unknown error";
        return HOEFEM_ERROR_DEVICE;
    }
}

static void check_material(int32_t kind, double threshold) {
    if (kind != dev::MAT_SNH && kind != dev::MAT_INVERTIBLE_SNH && kind != dev::MAT_PROJECTED_SNH)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unknown material kind");
    if (kind == dev::MAT_INVERTIBLE_SNH && !(threshold >= 0.0 && std::isfinite(threshold)))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "inversion threshold must be finite and >= 0");
}

static Model& M_(hoefem_model* m) {
    if (!m) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null model");
    return m->m;
}

extern "C" {

const char* hoefem_last_error(void) { return g_last_error.c_str(); }

int32_t hoefem_nodes_per_element(int32_t t) { return nodes_per_element(t); }

void hoefem_lame_from_young_poisson(double young, double poisson, double* mu, double* lambda) {
    const double m = 0.5 * young / (1.0 + poisson);
    *mu = m;
    *lambda = 0.5 * m * poisson / (1.0 - 2.0 * poisson);
}

hoefem_status hoefem_material_eval(double mu, double lambda, int64_t nq, const double* F, double* P, double* dPdF) {
    return guarded([&] {
        if (nq < 0 || (nq > 0 && !F)) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        double mr, lr;
        dev::snh_reparam(mu, lambda, &mr, &lr);
        material_eval(dev::MAT_SNH, 0.0, mr, lr, nq, F, P, dPdF);
    });
}

hoefem_status hoefem_material_energy(double mu, double lambda, int64_t nq, const double* F, double* psi) {
    return guarded([&] {
        if (nq < 0 || (nq > 0 && (!F || !psi))) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        material_energy(mu, lambda, nq, F, psi);
    });
}

hoefem_status hoefem_material_eval_kind(int32_t kind, double inversion_threshold, double mu, double lambda, int64_t nq,
                                       const double* F, double* P, double* dPdF) {
    return guarded([&] {
        if (nq < 0 || (nq > 0 && !F)) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        check_material(kind, inversion_threshold);
        double mr, lr;
        dev::snh_reparam(mu, lambda, &mr, &lr);
        material_eval(kind, inversion_threshold, mr, lr, nq, F, P, dPdF);
    });
}

hoefem_status hoefem_rotation_svd(int64_t nq, const double* F, double* U, double* S, double* Vt) {
    return guarded([&] {
        if (nq < 0 || (nq > 0 && (!F || !U || !S || !Vt))) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        rotation_svd_eval(nq, F, U, S, Vt);
    });
}

hoefem_status hoefem_mesh_reorder(int32_t nodes_per_element, int64_t num_vertices, int64_t num_elements,
                                  const int32_t* connectivity, int64_t* vertex_permutation,
                                  int64_t* element_permutation) {
    return guarded([&] {
        if (nodes_per_element <= 0 || num_vertices < 0 || num_elements < 0 || (num_elements > 0 && !connectivity) ||
            (num_vertices > 0 && !vertex_permutation) || (num_elements > 0 && !element_permutation))
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        mesh_reorder(num_vertices, num_elements, nodes_per_element, connectivity, vertex_permutation,
                     element_permutation);
    });
}

hoefem_status hoefem_model_create(const hoefem_model_desc* desc, hoefem_model** out) {
    return guarded([&] {
        if (!out) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null output");
        if (!desc) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null descriptor");
        std::unique_ptr<hoefem_model> h(new hoefem_model());
        model_create(h->m, desc);
        *out = h.release();
    });
}

void hoefem_model_destroy(hoefem_model* model) { delete model; }

hoefem_status hoefem_model_set_stream(hoefem_model* model, void* stream) {
    return guarded([&] { M_(model).stream = (hipStream_t)stream; });
}

hoefem_status hoefem_model_set_material(hoefem_model* model, int32_t kind, double inversion_threshold) {
    return guarded([&] {
        Model& m = M_(model);
        check_material(kind, inversion_threshold);
        m.material = kind;
        m.inversion_threshold = kind == dev::MAT_INVERTIBLE_SNH ? inversion_threshold : 0.0;
    });
}

int64_t hoefem_model_ndof(const hoefem_model* m) { return m ? 3 * m->m.nv : 0; }
int64_t hoefem_model_nnz(const hoefem_model* m) { return m ? m->m.pattern.nnz() : 0; }

hoefem_status hoefem_model_pattern(const hoefem_model* h, int32_t* rowptr, int32_t* col) {
    return guarded([&] {
        if (!h || !rowptr || !col) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        const Model* m = &h->m;
        std::copy(m->pattern.rowptr.begin(), m->pattern.rowptr.end(), rowptr);
        std::copy(m->pattern.col.begin(), m->pattern.col.end(), col);
    });
}

int32_t hoefem_model_num_colors(const hoefem_model* h, int32_t pass) {
    if (!h || pass < 0 || pass > 1) return -1;
    return (int32_t)h->m.colors[pass].size();
}

hoefem_status hoefem_model_colors(const hoefem_model* h, int32_t pass, int32_t* offsets, int32_t* labels) {
    return guarded([&] {
        if (!h || pass < 0 || pass > 1 || !offsets) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        const Model* m = &h->m;
        int32_t k = 0;
        offsets[0] = 0;
        for (size_t c = 0; c < m->colors[pass].size(); ++c) {
            for (int32_t l : m->colors[pass][c]) labels[k++] = l;
            offsets[c + 1] = k;
        }
    });
}

hoefem_status hoefem_model_set_pointer_mode(hoefem_model* model, int32_t mode) {
    return guarded([&] {
        Model& m = M_(model);
        if (mode != HOEFEM_POINTERS_DEVICE && mode != HOEFEM_POINTERS_HOST)
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unknown pointer mode");
        m.pointer_mode = mode;
    });
}

hoefem_status hoefem_model_forces_add(hoefem_model* model, double mu, double lambda, const double* u, double* f) {
    return guarded([&] {
        Model& m = M_(model);
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* du = io.in(u, n);
        double* df = io.inout(f, n);
        double mr, lr;
        dev::snh_reparam(mu, lambda, &mr, &lr);
        launch_element_forces(m, mr, lr, du);
        launch_force_gather(m, df);
        raise_kernel_errors(m);
        io.finish();
    });
}

hoefem_status hoefem_model_stiffness_add_prec(hoefem_model* model, double mu, double lambda, const double* u,
                                              void* values, int32_t precision) {
    return guarded([&] {
        Model& m = M_(model);
        if (precision != HOEFEM_PRECISION_F64 && precision != HOEFEM_PRECISION_F32)
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unknown precision");
        CallIo io(m);
        const size_t nnz = (size_t)m.pattern.nnz();
        const double* du = io.in(u, 3 * (size_t)m.nv);
        double mr, lr;
        dev::snh_reparam(mu, lambda, &mr, &lr);
        if (precision == HOEFEM_PRECISION_F64) {
            double* dv = io.inout((double*)values, nnz);
            launch_element_stiffness(m, mr, lr, du, false);
            launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), dv, 0, 0.0, nullptr, false, m.stream);
        } else {
            // FP32 element tangents (FP32 arithmetic), added to the FP32 CSR values in color order
            float* dv = io.inout((float*)values, nnz);
            launch_element_stiffness(m, mr, lr, du, true);
            launch_block_gather(m.pattern, m.scatter, m.d_ke32.get(), dv, 0, 0.0, nullptr, false, m.stream);
        }
        raise_kernel_errors(m);
        io.finish();
    });
}

hoefem_status hoefem_model_stiffness_add(hoefem_model* model, double mu, double lambda, const double* u,
                                         double* values) {
    return hoefem_model_stiffness_add_prec(model, mu, lambda, u, values, HOEFEM_PRECISION_F64);
}

hoefem_status hoefem_model_mass(hoefem_model* model, double density, double* values) {
    return guarded([&] {
        Model& m = M_(model);
        CallIo io(m);
        double* dv = io.out(values, (size_t)m.pattern.nnz());
        launch_element_mass(m, density);
        launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), dv, 0, 0.0, nullptr, true, m.stream);
        raise_kernel_errors(m);
        io.finish();
    });
}

hoefem_status hoefem_model_dirichlet_csr(hoefem_model* model, int64_t num_nodes, const int32_t* nodes, double* values) {
    return guarded([&] {
        Model& m = M_(model);
        if (num_nodes < 0 || (num_nodes > 0 && !nodes)) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad node list");
        CallIo io(m);
        values = io.inout(values, (size_t)m.pattern.nnz());
        std::vector<int32_t> dir(nodes, nodes + num_nodes);
        DevBuf<uint8_t> mask;
        DevBuf<int32_t> bc;
        DevBuf<double> scale;
        scale.alloc(1);
        int64_t nbc = 0;
        build_bc(m, dir, mask, bc, nbc);
        launch_dirichlet(m.pattern, values, mask.get(), bc.get(), nbc, scale.get(), m.stream);
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_model_spmv(hoefem_model* model, const double* values, double beta, double* y, double alpha,
                                const double* x) {
    return guarded([&] {
        Model& m = M_(model);
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* dA = io.in(values, (size_t)m.pattern.nnz());
        const double* dx = io.in(x, n);
        double* dy = io.inout(y, n);
        launch_spmv(m.pattern, dA, beta, dy, alpha, dx, m.stream);
        io.finish();
    });
}

hoefem_status hoefem_schwarz_create(hoefem_model* model, int64_t num_subdomains, const int64_t* offsets,
                                    const int32_t* nodes, int64_t num_untreated, const int32_t* untreated,
                                    hoefem_schwarz** out) {
    return guarded([&] {
        Model& m = M_(model);
        if (!out || num_subdomains < 0 || num_untreated < 0 || (num_subdomains > 0 && (!offsets || !nodes)))
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad arguments");
        std::unique_ptr<hoefem_schwarz> s(new hoefem_schwarz());
        static const int64_t zero_off = 0;
        schwarz_setup(s->s, m, num_subdomains, num_subdomains > 0 ? offsets : &zero_off, nodes, num_untreated,
                      untreated);
        HIP_CHECK(hipStreamSynchronize(m.stream));
        *out = s.release();
    });
}

void hoefem_schwarz_destroy(hoefem_schwarz* s) { delete s; }
int64_t hoefem_schwarz_nnz(const hoefem_schwarz* s) { return s ? s->s.P.nnz() : 0; }

hoefem_status hoefem_schwarz_pattern(const hoefem_schwarz* s, int32_t* rowptr, int32_t* col) {
    return guarded([&] {
        if (!s || !rowptr || !col) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        std::copy(s->s.P.rowptr.begin(), s->s.P.rowptr.end(), rowptr);
        std::copy(s->s.P.col.begin(), s->s.P.col.end(), col);
    });
}

hoefem_status hoefem_schwarz_update(hoefem_schwarz* s, const double* J) {
    return guarded([&] {
        if (!s) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null preconditioner");
        Model& m = *s->s.model;
        CallIo io(m);
        schwarz_update(s->s, io.in(J, (size_t)m.pattern.nnz()));
        raise_kernel_errors(m);
    });
}

hoefem_status hoefem_schwarz_values(hoefem_schwarz* s, double* values) {
    return guarded([&] {
        if (!s) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null preconditioner");
        Model& m = *s->s.model;
        CallIo io(m);
        double* dv = io.out(values, (size_t)s->s.P.nnz());
        HIP_CHECK(hipMemcpyAsync(dv, s->s.d_vals.get(), sizeof(double) * s->s.P.nnz(), hipMemcpyDeviceToDevice,
                                 m.stream));
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_schwarz_apply(hoefem_schwarz* s, const double* x, double* y) {
    return guarded([&] {
        if (!s) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null preconditioner");
        Model& m = *s->s.model;
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* dx = io.in(x, n);
        double* dy = io.out(y, n);
        schwarz_apply(s->s, dx, dy);
        io.finish();
    });
}

hoefem_status hoefem_pcg(hoefem_model* model, const double* A, hoefem_schwarz* schwarz, const double* b, double* x,
                         double tolerance, int64_t max_iterations, int64_t* iterations) {
    hoefem_status st = HOEFEM_OK;
    int64_t it = 0;
    const hoefem_status g = guarded([&] {
        Model& m = M_(model);
        static thread_local CgWork* work = nullptr;
        if (!work) work = new CgWork();
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* dA = io.in(A, (size_t)m.pattern.nnz());
        const double* db = io.in(b, n);
        double* dx = io.inout(x, n);
        st = pcg_solve(m, *work, dA, schwarz ? &schwarz->s : nullptr, db, dx, tolerance, max_iterations, &it);
        io.finish();
    });
    if (iterations) *iterations = it;
    if (g != HOEFEM_OK) return g;
    if (st != HOEFEM_OK) g_last_error = "CG stopped with status " + std::to_string((int)st);
    return st;
}

hoefem_status hoefem_pcg_noalloc(hoefem_model* model, const double* A, hoefem_schwarz* schwarz, const double* b,
                                 double* x, double tolerance, int64_t max_iterations, int64_t* iterations) {
    hoefem_status st = HOEFEM_OK;
    int64_t it = 0;
    const hoefem_status g = guarded([&] {
        Model& m = M_(model);
        if (m.pointer_mode != HOEFEM_POINTERS_DEVICE)
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "hoefem_pcg_noalloc takes device pointers only");
        if (!A || !b || !x) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        st = pcg_solve_noalloc(m, A, schwarz ? &schwarz->s : nullptr, b, x, tolerance, max_iterations, &it);
    });
    if (iterations) *iterations = it;
    if (g != HOEFEM_OK) return g;
    if (st != HOEFEM_OK) g_last_error = "CG stopped with status " + std::to_string((int)st);
    return st;
}

hoefem_status hoefem_model_autotune(hoefem_model* model, int32_t* lanes_per_node, double* candidate_us) {
    return guarded([&] {
        Model& m = M_(model);
        const int g = autotune_spmv(m.pattern, m.stream, candidate_us);
        if (lanes_per_node) *lanes_per_node = g;
    });
}

hoefem_status hoefem_schwarz_autotune(hoefem_schwarz* s, int32_t* lanes_per_node, double* candidate_us) {
    return guarded([&] {
        if (!s) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null preconditioner");
        const int g = autotune_spmv(s->s.P, s->s.model->stream, candidate_us);
        if (lanes_per_node) *lanes_per_node = g;
    });
}

// Times a fixed number of preconditioned CG iterations for every pair (lanes per node of the
// Jacobian kernel, lanes per node of the preconditioner kernel) and keeps the fastest. The
// trial system is J = M (symmetric positive definite) with the Schwarz operator of M and
// tolerance 0, so every trial runs exactly the same number of full iterations. J and P are
// overwritten; the next step assembles them again.
static void sim_autotune_cg(Sim& s, int* ga_out, int* gb_out, double* candidate_us) {
    Model& m = *s.model;
    const int64_t n = 3 * m.nv, nnz = m.pattern.nnz();
    if (s.mixed) {
        launch_f64_to_f32(m, s.d_M.get(), s.d_J32.get(), nnz);
        if (s.schwarz) schwarz_update(*s.schwarz, s.d_J32.get());
    } else {
        HIP_CHECK(hipMemcpyAsync(s.d_J.get(), s.d_M.get(), sizeof(double) * nnz, hipMemcpyDeviceToDevice, m.stream));
        if (s.schwarz) schwarz_update(*s.schwarz, s.d_J.get());
    }
    launch_fill(m, s.d_F.get(), n, 1.0);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const int cand[3] = {8, 16, 32};
    const int nb = s.schwarz ? 3 : 1;
    int64_t it = 0;
    auto run = [&](int64_t iters) {
        launch_fill(m, s.d_sol.get(), n, 0.0);
        if (s.mixed)
            (void)pcg_solve(m, s.cg, s.d_J32.get(), s.schwarz.get(), s.d_F.get(), s.d_sol.get(), 0.0, iters, &it);
        else
            (void)pcg_solve(m, s.cg, s.d_J.get(), s.schwarz.get(), s.d_F.get(), s.d_sol.get(), 0.0, iters, &it);
    };
    // AUTOTUNE_ROUNDS rounds over all pairs, each pair's best 96-iteration solve kept: a drift of the
    // GPU's speed during tuning then reaches every pair instead of favouring the ones measured last
    constexpr int AUTOTUNE_ROUNDS = 5;
    float best_ms[9];
    std::fill(best_ms, best_ms + 9, 1e30f);
    for (int rep = 0; rep < AUTOTUNE_ROUNDS; ++rep)
        for (int ia = 0; ia < 3; ++ia)
            for (int ib = 0; ib < nb; ++ib) {
                m.pattern.tuned_group = cand[ia];
                if (s.schwarz) s.schwarz->P.tuned_group = cand[ib];
                run(4);  // also captures this pair's batch graph
                HIP_CHECK(hipEventRecord(e0, m.stream));
                run(96);
                HIP_CHECK(hipEventRecord(e1, m.stream));
                HIP_CHECK(hipEventSynchronize(e1));
                float t = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                best_ms[3 * ia + ib] = std::min(best_ms[3 * ia + ib], t);
            }
    int best_a = 8, best_b = 8;
    double best_t = 1e300;
    for (int ia = 0; ia < 3; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            const double ms = best_ms[3 * ia + ib];
            if (candidate_us) candidate_us[3 * ia + ib] = 1e3 * ms / 96.0;
            if (ms < best_t) {
                best_t = ms;
                best_a = cand[ia];
                best_b = cand[ib];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    m.pattern.tuned_group = best_a;
    if (s.schwarz) s.schwarz->P.tuned_group = best_b;
    *ga_out = best_a;
    *gb_out = s.schwarz ? best_b : 0;
}

hoefem_status hoefem_sim_autotune(hoefem_sim* sim, int32_t* lanes_matrix, int32_t* lanes_preconditioner,
                                  double* candidate_us) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        int ga = 0, gb = 0;
        if (candidate_us) std::fill(candidate_us, candidate_us + 9, 0.0);
        sim_autotune_cg(sim->s, &ga, &gb, candidate_us);
        if (lanes_matrix) *lanes_matrix = ga;
        if (lanes_preconditioner) *lanes_preconditioner = gb;
    });
}

hoefem_status hoefem_profile_enable(hoefem_model* model, int32_t enable) {
    return guarded([&] {
        Model& m = M_(model);
        HIP_CHECK(hipStreamSynchronize(m.stream));
        m.prof = enable != 0;
        m.prof_full_launches = m.prof_all_launches = 0;
        m.prof_full_ms = m.prof_all_ms = 0.0;
        for (KernelProf& k : m.kprof) k.reset();
    });
}

hoefem_status hoefem_profile_kernel(hoefem_model* model, int32_t kernel, int64_t* launches, double* total_ms) {
    return guarded([&] {
        Model& m = M_(model);
        if (kernel < 0 || kernel >= PROF_KINDS) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unknown profiled kernel");
        HIP_CHECK(hipStreamSynchronize(m.stream));
        KernelProf& k = m.kprof[kernel];
        k.resolve();
        if (launches) *launches = k.launches;
        if (total_ms) *total_ms = k.ms;
    });
}

hoefem_status hoefem_profile_read(hoefem_model* model, int64_t* full_launches, double* full_ms, int64_t* all_launches,
                                  double* all_ms) {
    return guarded([&] {
        Model& m = M_(model);
        if (full_launches) *full_launches = m.prof_full_launches;
        if (full_ms) *full_ms = m.prof_full_ms;
        if (all_launches) *all_launches = m.prof_all_launches;
        if (all_ms) *all_ms = m.prof_all_ms;
    });
}

hoefem_status hoefem_sim_create(hoefem_model* model, const hoefem_sim_desc* desc, hoefem_sim** out) {
    return guarded([&] {
        Model& m = M_(model);
        if (!desc || !out) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        if (desc->num_dirichlet > 0 && !desc->dirichlet_nodes) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null Dirichlet nodes");
        std::unique_ptr<hoefem_sim> h(new hoefem_sim());
        sim_create(h->s, m, desc);
        *out = h.release();
    });
}

void hoefem_sim_destroy(hoefem_sim* sim) {
    if (sim && sim->s.learned_handle) sim->s.learned_handle->sim = nullptr;
    delete sim;
}

double hoefem_sim_representative_force(const hoefem_sim* sim) { return sim ? sim->s.rep_force : 0.0; }

hoefem_status hoefem_sim_mass_values(hoefem_sim* sim, double* values) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        Model& m = *sim->s.model;
        CallIo io(m);
        double* dv = io.out(values, (size_t)m.pattern.nnz());
        HIP_CHECK(hipMemcpyAsync(dv, sim->s.d_M.get(), sizeof(double) * m.pattern.nnz(), hipMemcpyDeviceToDevice, m.stream));
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_sim_damping_values(hoefem_sim* sim, double* values) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        Model& m = *sim->s.model;
        const int64_t nnz = m.pattern.nnz();
        CallIo io(m);
        double* dv = io.out(values, (size_t)nnz);
        if (sim->s.damping)
            HIP_CHECK(hipMemcpyAsync(dv, sim->s.d_D.get(), sizeof(double) * nnz, hipMemcpyDeviceToDevice, m.stream));
        else
            HIP_CHECK(hipMemsetAsync(dv, 0, sizeof(double) * nnz, m.stream));
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_sim_jacobian_values(hoefem_sim* sim, double* values) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        Model& m = *sim->s.model;
        CallIo io(m);
        double* dv = io.out(values, (size_t)m.pattern.nnz());
        if (sim->s.fp32_jacobian())
            launch_f32_to_f64(m, sim->s.d_J32.get(), dv, m.pattern.nnz());
        else
            HIP_CHECK(hipMemcpyAsync(dv, sim->s.d_J.get(), sizeof(double) * m.pattern.nnz(), hipMemcpyDeviceToDevice,
                                     m.stream));
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_sim_residual(hoefem_sim* sim, double t, double dt, const double* u, const double* v,
                                  const double* w, double* F) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        (void)t;
        Model& m = *sim->s.model;
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* du = io.in(u, n);
        const double* dv = io.in(v, n);
        const double* dw = io.in(w, n);
        double* dF = io.out(F, n);
        launch_be_begin(m, du, dv, dt, sim->s.d_up.get(), sim->s.d_vn.get());
        const double n2 = sim_eval(sim->s, dt, dw);
        if (std::isnan(n2)) throw Error(HOEFEM_ERROR_SINGULAR, "singular deformation gradient");
        HIP_CHECK(hipMemcpyAsync(dF, sim->s.d_F.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, m.stream));
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_sim_set_dirichlet_values(hoefem_sim* sim, const double* u_values, const double* v_values) {
    return guarded([&] {
        if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
        Sim& s = sim->s;
        Model& m = *s.model;
        const int64_t n = 3 * m.nv;
        auto upload = [&](const double* vals, DevBuf<double>& buf, bool& has) {
            has = vals != nullptr;
            if (!has) return;
            std::vector<double> full((size_t)n, 0.0);
            for (size_t k = 0; k < s.dir.size(); ++k)
                for (int i = 0; i < 3; ++i) {
                    const double x = vals[3 * k + i];
                    if (!std::isfinite(x)) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "non-finite Dirichlet value");
                    full[3 * (size_t)s.dir[k] + i] = x;
                }
            buf.upload(full, m.stream);
        };
        upload(u_values, s.d_bcu, s.has_bcu);
        upload(v_values, s.d_bcv, s.has_bcv);
    });
}

struct hoefem_batch {
    Batch b;
};

hoefem_status hoefem_batch_create(hoefem_sim* const* sims, int32_t count, hoefem_batch** out) {
    return guarded([&] {
        if (!sims || !out || count < 1) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "a batch needs at least one sim");
        std::unique_ptr<hoefem_batch> h(new hoefem_batch());
        for (int32_t i = 0; i < count; ++i) {
            if (!sims[i]) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
            Sim* s = &sims[i]->s;
            if (std::find(h->b.sims.begin(), h->b.sims.end(), s) != h->b.sims.end())
                throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "a sim appears twice in the batch");
            if (s->model->pointer_mode != HOEFEM_POINTERS_DEVICE)
                throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "batched scenes take device pointers");
            if (s->mixed_mode == 2)
                throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "batched scenes run the FP64 or the FP32-operator step");
            if (i > 0 && s->mixed_mode != h->b.sims[0]->mixed_mode)
                throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "batched scenes share one precision mode");
            h->b.sims.push_back(s);
        }
        *out = h.release();
    });
}

void hoefem_batch_destroy(hoefem_batch* batch) { delete batch; }

hoefem_status hoefem_batch_step(hoefem_batch* batch, const double* dt, double* const* u, double* const* v,
                                hoefem_step_stats* stats, double* latency_ms, int32_t* status) {
    return guarded([&] {
        if (!batch || !dt || !u || !v) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        batch_step(batch->b, dt, u, v, stats, latency_ms, status);
    });
}

hoefem_status hoefem_sim_step(hoefem_sim* sim, double t, double dt, double* u, double* v, hoefem_step_stats* stats) {
    return guarded([&] {
        if (!sim || !u || !v) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        Model& m = *sim->s.model;
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        double* du = io.inout(u, n);
        double* dv = io.inout(v, n);
        sim_step(sim->s, t, dt, du, dv, stats);
        io.finish();
    });
}


// ------------------------------------------------------------------ dynamic system (hamilton2)
static Sim& S_(hoefem_sim* sim) {
    if (!sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null sim");
    return sim->s;
}

hoefem_status hoefem_sim_update_damping(hoefem_sim* sim, const double* u) {
    return guarded([&] {
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!u) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        if (!s.damping) return;
        CallIo io(m);
        launch_bc_copy(m, io.in(u, 3 * (size_t)m.nv), s.d_dirmask.get(), s.bcu(), s.d_uu.get());
        if (sim_damping(s, s.d_uu.get()) != HOEFEM_OK)
            throw Error(HOEFEM_ERROR_SINGULAR, "singular deformation gradient in the damping stiffness");
        s.solve_outdated = s.apply_outdated = true;
        io.finish();
    });
}

hoefem_status hoefem_sim_eval_f(hoefem_sim* sim, double t, const double* u, const double* v, double* f) {
    return guarded([&] {
        (void)t;
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!u || !v || !f) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* du = io.in(u, n);
        const double* dv = io.in(v, n);
        double* df = io.out(f, n);
        // the reference applies the boundary conditions to copies of u and v (newton_cg.rs:449-455)
        launch_bc_copy(m, du, s.d_dirmask.get(), s.bcu(), s.d_uu.get());
        launch_bc_copy(m, dv, s.d_dirmask.get(), s.bcv(), s.d_vv.get());
        sim_eval_f(s, s.d_uu.get(), s.d_vv.get(), df);
        raise_kernel_errors(m);
        io.finish();
    });
}

hoefem_status hoefem_sim_apply_mass_add(hoefem_sim* sim, double* y, const double* x) {
    return guarded([&] {
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!y || !x) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* dx = io.in(x, n);
        double* dy = io.inout(y, n);
        launch_mass_add(m, s.d_M.get(), dx, dy, s.cg, nullptr);
        io.finish();
    });
}

hoefem_status hoefem_sim_set_state(hoefem_sim* sim, double t, const double* u, const double* v) {
    return guarded([&] {
        (void)t;
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!u || !v) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        launch_bc_copy(m, io.in(u, n), s.d_dirmask.get(), s.bcu(), s.d_su.get());
        launch_bc_copy(m, io.in(v, n), s.d_dirmask.get(), s.bcv(), s.d_sv.get());
        s.state_set = true;
        s.has_alpha = s.has_beta = false;
        s.solve_outdated = s.apply_outdated = true;
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}

hoefem_status hoefem_sim_init_solve_jacobian_combination(hoefem_sim* sim, const double* alpha, const double* beta) {
    return guarded([&] {
        Sim& s = S_(sim);
        s.has_alpha = alpha != nullptr;
        s.has_beta = beta != nullptr;
        s.alpha = alpha ? *alpha : 0.0;
        s.beta = beta ? *beta : 0.0;
        s.solve_outdated = true;
    });
}

hoefem_status hoefem_sim_solve_jacobian_combination(hoefem_sim* sim, const double* rhs, double* sol,
                                                     int64_t* cg_iterations) {
    int64_t it = 0;
    const hoefem_status g = guarded([&] {
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!rhs || !sol) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        if (!s.state_set) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "set_state must precede solve_jacobian_combination");
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const double* drhs = io.in(rhs, n);
        double* dsol = io.out(sol, n);
        if (s.solve_outdated) {
            if (sim_assemble_jacobian(s, s.has_alpha, s.alpha, s.has_beta, s.beta, s.d_su.get()) != HOEFEM_OK)
                throw Error(HOEFEM_ERROR_SINGULAR, "singular deformation gradient or element Jacobian");
            s.solve_outdated = false;
        }
        const hoefem_status st = sim_cg(s, drhs, dsol, &it);
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
        if (st != HOEFEM_OK) throw Error(st, "CG stopped with status " + std::to_string((int)st));
    });
    if (cg_iterations) *cg_iterations = it;
    return g;
}

hoefem_status hoefem_sim_init_apply_jacobian_combination(hoefem_sim* sim, const double* alpha, const double* beta,
                                                         const double* gamma) {
    return guarded([&] {
        Sim& s = S_(sim);
        const double* c[3] = {alpha, beta, gamma};
        for (int k = 0; k < 3; ++k) {
            s.ap_has[k] = c[k] != nullptr;
            s.ap_coef[k] = c[k] ? *c[k] : 0.0;
        }
        s.apply_outdated = true;
        s.apply_ready = true;
    });
}

hoefem_status hoefem_sim_apply_jacobian_combination(hoefem_sim* sim, double* y, const double* x) {
    return guarded([&] {
        Sim& s = S_(sim);
        Model& m = *s.model;
        if (!y || !x) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        if (!s.state_set || !s.apply_ready)
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "set_state and init_apply_jacobian_combination must come first");
        CallIo io(m);
        const size_t n = 3 * (size_t)m.nv;
        const int64_t nnz = m.pattern.nnz();
        const double* dx = io.in(x, n);
        double* dy = io.inout(y, n);
        if (s.apply_outdated) {
            // H = -alpha K - beta D + gamma M (df/du = -K, df/dv = -D), with the Dirichlet treatment of
            // compute_jacobian_combination_into so that it is the operator solve_jacobian_combination inverts
            s.d_H.ensure((size_t)nnz);
            const double* D = (s.ap_has[1] && s.damping) ? s.d_D.get() : nullptr;
            const double cd = s.ap_has[1] ? -s.ap_coef[1] : 0.0;
            const double cm = s.ap_has[2] ? s.ap_coef[2] : 0.0;
            if (s.ap_has[0]) {
                launch_element_stiffness(m, s.mu_r, s.lambda_r, s.d_su.get(), false);
                launch_block_gather(m.pattern, m.scatter, m.d_ke.get(), s.d_H.get(), 1, -s.ap_coef[0],
                                    s.ap_has[2] ? s.d_M.get() : nullptr, true, m.stream, D, cd, cm);
            } else {
                launch_comb(m, nnz, s.d_H.get(), cd, D, cm, s.d_M.get());
            }
            launch_dirichlet(m.pattern, s.d_H.get(), s.d_dirmask.get(), s.d_bc_nodes.get(), s.n_bc, s.d_scale.get(),
                             m.stream);
            raise_kernel_errors(m);
            s.apply_outdated = false;
        }
        launch_spmv(m.pattern, s.d_H.get(), 1.0, dy, 1.0, dx, m.stream);
        io.finish();
        HIP_CHECK(hipStreamSynchronize(m.stream));
    });
}


// ------------------------------------------------------------------ learned coarse correction
static Learned& LC_(hoefem_learned* l) {
    if (!l) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null learned correction");
    return l->L;
}

int64_t hoefem_learned_param_count(void) { return LC_PARAMS; }

hoefem_status hoefem_learned_create(hoefem_sim* sim, const hoefem_learned_desc* desc, hoefem_learned** out) {
    return guarded([&] {
        if (!sim || !desc || !out) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null argument");
        if (!sim->s.schwarz) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "the learned correction needs Schwarz subdomains");
        Model& m = *sim->s.model;
        std::unique_ptr<hoefem_learned> h(new hoefem_learned());
        std::vector<uint8_t> mask(m.nv, 0);
        for (int32_t v : sim->s.dir) mask[v] = 1;
        learned_setup(h->L, m, desc->num_aggregates, mask, desc->seed);
        if (h->L.na > CG_MAX_BLOCKS) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "too many aggregates");
        if (sim->s.learned_handle) sim->s.learned_handle->sim = nullptr;  // replaces an earlier one
        h->sim = &sim->s;
        sim->s.learned = &h->L;
        sim->s.learned_handle = h.get();
        HIP_CHECK(hipStreamSynchronize(m.stream));
        *out = h.release();
    });
}

hoefem_status hoefem_learned_destroy(hoefem_learned* l) {
    return guarded([&] {
        if (!l) return;
        if (l->sim && l->sim->learned == &l->L) {
            l->sim->learned = nullptr;
            l->sim->learned_handle = nullptr;
        }
        delete l;
    });
}

hoefem_status hoefem_learned_harvest(hoefem_learned* l, int32_t enable, int32_t capacity, int32_t every) {
    return guarded([&] { learned_harvest_enable(LC_(l), enable != 0, capacity, every); });
}

hoefem_status hoefem_learned_info(hoefem_learned* l, hoefem_learned_stats* out) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (!out) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null output");
        out->num_aggregates = L.na;
        out->max_aggregate_dofs = L.max_dofs;
        out->harvested_pairs = L.pairs;
        out->stored_pairs = L.count;
        out->trained = L.trained ? 1 : 0;
        out->enabled = L.enabled ? 1 : 0;
        out->learned_solves = L.solves;
        out->fallbacks = L.fallbacks;
        out->learned_iterations = L.iterations;
        out->last_status = L.last_status;
    });
}

hoefem_status hoefem_learned_train(hoefem_learned* l, int32_t steps, double learning_rate, int32_t batch,
                                   double* initial_loss, double* final_loss) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (!l->sim) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "the simulation of this correction was destroyed");
        Sim& s = *l->sim;
        if (steps < 0 || batch < 1) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad training settings");
        const double fl = s.mixed ? learned_train(L, s.d_J32.get(), s.schwarz.get(), steps, learning_rate, batch, initial_loss)
                                  : learned_train(L, s.d_J.get(), s.schwarz.get(), steps, learning_rate, batch, initial_loss);
        if (final_loss) *final_loss = fl;
    });
}

hoefem_status hoefem_learned_enable(hoefem_learned* l, int32_t enable) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (enable && !L.trained) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "train the correction before enabling it");
        L.enabled = enable != 0;
    });
}

hoefem_status hoefem_learned_get_params(hoefem_learned* l, float* params) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (!params) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null output");
        HIP_CHECK(hipStreamSynchronize(L.model->stream));
        HIP_CHECK(hipMemcpy(params, L.d_params.get(), sizeof(float) * LC_PARAMS, hipMemcpyDeviceToHost));
    });
}

hoefem_status hoefem_learned_set_params(hoefem_learned* l, const float* params) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (!params) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "null input");
        learned_set_params(L, params, false);
        HIP_CHECK(hipStreamSynchronize(L.model->stream));
    });
}

hoefem_status hoefem_learned_mlp_forward(hoefem_learned* l, int64_t n, const float* x, float* y) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (n < 0 || (n > 0 && (!x || !y))) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad batch");
        if (n == 0) return;
        learned_mlp_forward(L, n, x, y);
        HIP_CHECK(hipStreamSynchronize(L.model->stream));
    });
}

hoefem_status hoefem_learned_mlp_grad(hoefem_learned* l, int64_t n, const float* x, const float* dy, float* grad) {
    return guarded([&] {
        Learned& L = LC_(l);
        if (n < 1 || !x || !dy || !grad) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad batch");
        learned_mlp_grad(L, n, x, dy, grad, 1.0f);
        HIP_CHECK(hipStreamSynchronize(L.model->stream));
    });
}

}  // extern "C"
