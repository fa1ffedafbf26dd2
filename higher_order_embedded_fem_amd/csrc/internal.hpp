// Internal declarations of the hoefem HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hoefem.h"
#include "mlp_device.hpp"

namespace hoefem {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    hoefem_status code;
    Error(hoefem_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                                    \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            throw ::hoefem::Error(HOEFEM_ERROR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

void check_launch(const char* what);

// Raises func's dynamic LDS limit on the current device to at least `bytes`. The attribute is
// shared by every thread of the process, so it is only ever raised (under a lock): a launch
// sized by one thread cannot find the limit lowered by another.
void raise_lds_limit(const void* func, size_t bytes);

// ------------------------------------------------------------------ buffers
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    }
    void ensure(size_t count) {
        if (count > n || p == nullptr) alloc(count ? count : 1);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // setup-time upload; blocking, so the host array may be released right after
    void upload(const T* h, size_t count, hipStream_t s) {
        ensure(count);
        if (count) {
            HIP_CHECK(hipStreamSynchronize(s));
            HIP_CHECK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
        }
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    T* get() const { return p; }
};

template <class T>
struct PinnedBuf {
    T* p = nullptr;
    explicit PinnedBuf(size_t count = 1, unsigned flags = hipHostMallocDefault) {
        HIP_CHECK(hipHostMalloc((void**)&p, count * sizeof(T), flags));
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
};

// ------------------------------------------------------------------ element types
enum : int { TET4 = 0, TET10 = 1, TET20 = 2, HEX8 = 3, HEX20 = 4, HEX27 = 5 };
int nodes_per_element(int t);
int geometry_nodes(int t);

// error flags raised by kernels (bitwise OR into a device int)
enum : int { KERR_SINGULAR = 1, KERR_PATTERN = 2, KERR_NONFINITE = 4, KERR_ZERO_DIAG = 8, KERR_SCHWARZ = 16 };

// ------------------------------------------------------------------ node-blocked CSR
// A CSR pattern whose rows come in xyz triples sharing one node-level column set (element
// node pairs x 3x3 blocks, fenris/src/assembly.rs:177-249). "Diagonal-only" nodes carry a
// single diagonal entry per row (untreated Schwarz nodes, schwarz_precond.rs:150-158).
// One node row block as the SpMV kernels consume it: node id, first neighbour (into ncol),
// neighbour count, offset of the block's first value (rowptr[3 node]). Nodes with neighbours
// come first, diagonal-only nodes after them.
struct __align__(16) NodeEntry {
    int32_t node, first, count, rowoff;
};

struct NodePattern {
    int64_t num_nodes = 0;
    std::vector<int32_t> nptr, ncol;  // node-level neighbour lists, sorted
    std::vector<int32_t> rowptr, col; // DOF-level CSR
    std::vector<int32_t> diag_pos;    // position of each row's diagonal entry
    std::vector<uint8_t> diag_only;   // per node
    int max_nbr = 0;
    int tuned_group = 0;  // lanes per node row block chosen by autotune_spmv (0: heuristic)
    DevBuf<int32_t> d_nptr, d_ncol, d_rowptr, d_col, d_diag_pos;
    DevBuf<uint8_t> d_diag_only;
    std::vector<NodeEntry> entries;   // block nodes, then diagonal-only nodes
    int64_t n_block = 0, n_diag = 0;
    DevBuf<NodeEntry> d_entries;
    int64_t nnz() const { return (int64_t)col.size(); }
    int64_t ndof() const { return 3 * num_nodes; }
    int32_t find(int32_t a, int32_t b) const;  // local index of b in nbr(a), or -1
    void upload(hipStream_t s);
};

// Build the pattern of the union of node sets (each set couples all its node pairs) plus
// diagonal-only nodes.
void mesh_reorder(int64_t nv, int64_t ne, int n, const int32_t* conn, int64_t* vperm, int64_t* eperm);
void build_node_pattern(int64_t num_nodes, int64_t nsets, const int64_t* off, const int32_t* nodes,
                        const std::vector<int32_t>& diag_only_nodes, NodePattern& out);

// First-fit greedy coloring of node sets (paradis/src/coloring.rs:48-72), labels in order.
std::vector<std::vector<int32_t>> greedy_coloring(int64_t nsets, const int64_t* off, const int32_t* nodes);

// Color-ordered incidence of dense blocks ("element matrices") into a node pattern; shared
// by the stiffness / mass assembly and the Schwarz preconditioner assembly.
struct BlockScatter {
    int64_t nblocks = 0;
    int max_m = 0;
    std::vector<int32_t> m;         // nodes per block
    std::vector<int64_t> val_off;   // offset of block b's (3m)^2 dense values in the staging
    std::vector<int64_t> slot_off;  // offset of block b's m*m slot table
    std::vector<int32_t> slot;      // slot[slot_off[b] + a*m + c] = local index of node c in nbr(node a)
    std::vector<int32_t> inc_ptr;   // per pattern node
    int max_inc = 0;                // most incident blocks of one node
    std::vector<int32_t> inc_blk, inc_loc;  // incident (block, local node) in color order
    DevBuf<int32_t> d_m, d_slot, d_inc_ptr, d_inc_blk, d_inc_loc;
    DevBuf<int64_t> d_val_off, d_slot_off;
    void build(const NodePattern& pat, int64_t nblocks, const int64_t* off, const int32_t* nodes,
               const std::vector<std::vector<int32_t>>* colors, int ncolor_lists);
    void upload(hipStream_t s);
};

// ------------------------------------------------------------------ kernel timing
// Kernel-timestamp event pairs of one kernel kind (profiling mode only): launches go through
// launch_timed, which brackets the kernel with its own start / end timestamps.
enum : int { PROF_TANGENT = 0, PROF_JACOBIAN_GATHER = 1, PROF_ELEMENT_FORCES = 2, PROF_FORCE_GATHER = 3, PROF_KINDS = 4 };
struct KernelProf {
    std::vector<hipEvent_t> ev;  // start / stop pairs
    size_t pending = 0;          // pairs recorded since the last resolve
    int64_t launches = 0;
    double ms = 0.0;
    void next(hipEvent_t& a, hipEvent_t& b);
    void resolve();  // the stream that recorded the events must be synchronized
    void reset() { pending = 0, launches = 0, ms = 0.0; }
    KernelProf() = default;
    KernelProf(const KernelProf&) = delete;
    KernelProf& operator=(const KernelProf&) = delete;
    ~KernelProf();
};

// ------------------------------------------------------------------ model
struct DeviceQuadrature {
    DevBuf<int32_t> off;  // [E + 1]
    DevBuf<double> w;     // [Q]
    DevBuf<double> xi;    // [Q * 3]
    int max_points = 0;
};

struct Model {
    int type = 0, n = 0, gn = 0;
    int64_t nv = 0, ne = 0, ninterior = 0;
    std::vector<double> vertices;
    std::vector<int32_t> conn;                    // element-major [E * n]
    std::vector<std::vector<int32_t>> colors[2];  // pass -> color -> element ids
    NodePattern pattern;
    BlockScatter scatter;                         // element blocks -> pattern, color order
    double mass_reg = 0.0;
    int material = 0;                 // dev::MAT_* (elastic model wrapping Stable Neo-Hookean)
    double inversion_threshold = 0.0; // InvertibleMaterial threshold
    hipStream_t stream = nullptr;
    int pointer_mode = HOEFEM_POINTERS_DEVICE;  // where the caller's vectors of the hot-path calls live
    DevBuf<char> io_stage[8];                   // device staging of host-pointer calls
    DevBuf<double> d_vertices;
    DevBuf<int32_t> d_conn;  // SoA: [k * E + e]
    DeviceQuadrature q_mass, q_stiff, q_ell, q_interior_mass;
    DevBuf<double> d_fe;  // element force staging [E * 3n]
    DevBuf<double> d_ke;  // element matrix staging [E * (3n)^2]
    DevBuf<float> d_ke32;  // FP32 element matrix staging (mixed precision)
    DevBuf<int> d_err;
    DevBuf<int32_t> d_elem_m;   // per element: nodes (n), for the batched element-matrix kernels
    DevBuf<int64_t> d_elem_off; // per element: offset of its (3n)^2 block in d_ke
    DevBuf<int> d_elem_need;    // per element: its matrix needs the eigen path (projection)
    PinnedBuf<int> h_err{1};
    // kernel timing of the CG system-matrix kernel (HIP events on the model stream)
    bool prof = false;
    int prof_tag = -1;           // PROF_* kind of the next launches of the tagged launchers, -1: none
    KernelProf kprof[PROF_KINDS];
    std::vector<hipEvent_t> prof_ev;
    int64_t prof_full_launches = 0, prof_all_launches = 0;
    double prof_full_ms = 0.0, prof_all_ms = 0.0;
    ~Model() {
        for (auto e : prof_ev) (void)hipEventDestroy(e);
    }
};

// ------------------------------------------------------------------ Schwarz
struct Schwarz {
    Model* model = nullptr;
    int64_t nsub = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> nodes, untreated;
    std::vector<std::vector<int32_t>> colors;
    NodePattern P;
    BlockScatter scatter;          // blocks -> P pattern
    std::vector<int32_t> jslot;    // per block entry (a, c): local index of node c in J's nbr(node a)
    DevBuf<int32_t> d_jslot, d_untreated, d_nodes;
    DevBuf<int64_t> d_off;
    bool fp32 = false;             // P stored in FP32 (mixed precision)
    DevBuf<double> d_vals;         // P values (FP64)
    DevBuf<float> d_vals32;        // P values (FP32)
    DevBuf<double> d_blocks;       // gathered / inverted blocks (always FP64)
    DevBuf<int> d_fallback;        // per block: 1 if the Gauss-Jordan path rejected it (eigen path)
    // blocks above SCHWARZ_SMALL_DIM: blocked Gauss-Jordan in global memory (32 x 32 MFMA tiles)
    int small_max_dim = 0, large_max_dim = 0;
    std::vector<int32_t> large;    // their block indices
    DevBuf<int32_t> d_large;
    DevBuf<int64_t> d_colp_off;    // per large block: offset of its column-panel copy (dim x 32)
    DevBuf<double> d_piv, d_colp, d_lnorm;  // pivot tile inverses, column panels, |B|_F^2 and |B^-1|_F^2
};
constexpr int SCHWARZ_SMALL_DIM = 96;  // blocks up to this size: register Gauss-Jordan / LDS eigen paths

// ------------------------------------------------------------------ CG workspace
struct CgScalars {
    double zTr, pAp, alpha, beta, rr, bnorm, tol;
    double zrec[2];  // zTr of iteration k in zrec[k & 1] (read by the next iteration's beta)
    int64_t iters, max_iter;  // iters: completed iterations (the A-kernel's k)
    int64_t kcur;             // k of the running iteration (written by the A-kernel for the B-kernel)
    int status;  // 0 running, 1 converged, hoefem_status codes for errors
    int pad;
};

// Per-block partial sums of one launch, summed again by every block of the next launch
// (no atomics or device-scope fences in the solver loop). Sized for the largest grid.
constexpr int CG_MAX_BLOCKS = 1024;

struct CgWork {
    int64_t n = 0;
    DevBuf<double> zp[2], ra[2];  // per node [z | p] and [r | Ap], six doubles, double-buffered
    DevBuf<double> diag_inv;
    DevBuf<double> part_pre, part_a, part_b;  // b.b (preamble 1), pAp (A-kernel), zTr and rr (B-kernel)
    DevBuf<double> partials;                  // residual kernel partials
    DevBuf<CgScalars> scalars;
    // every solve runs on its own stream (joined to the caller's by events); a batch of
    // iterations is captured once as a graph and replayed while its arguments stay the same
    hipStream_t cs = nullptr;
    hipEvent_t join[2] = {nullptr, nullptr};
    hipGraphExec_t batch_exec = nullptr;
    std::vector<const void*> batch_key;
    bool use_graphs = true;
    PinnedBuf<CgScalars> h_scalars{3};  // final readback (slot 2)
    PinnedBuf<int> h_status{1, hipHostMallocCoherent};  // the solve's status, written by the kernel that ends it
    hipEvent_t ev[2] = {nullptr, nullptr};
    void ensure(int64_t n, int nblocks);
    ~CgWork();
};

// ------------------------------------------------------------------ kernels (assembly.hip)
void launch_element_forces(Model& m, double mu_r, double lambda_r, const double* u);  // -> m.d_fe
void launch_element_stiffness(Model& m, double mu_r, double lambda_r, const double* u, bool fp32);  // -> d_ke / d_ke32
void launch_element_mass(Model& m, double density);  // -> m.d_ke (includes regularization)
// hipLaunchKernelGGL on stream s, or, with prof set, hipExtLaunchKernelGGL with a start / stop
// event pair of *prof around the kernel's own execution
template <class K, class... Args>
inline void launch_timed(KernelProf* prof, K kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s,
                         Args... args) {
    if (prof) {
        hipEvent_t a, b;
        prof->next(a, b);
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    }
}
inline KernelProf* prof_slot(Model& m) { return (m.prof && m.prof_tag >= 0) ? &m.kprof[m.prof_tag] : nullptr; }

// values (+)= gathered blocks; mode 0: values += sum; mode 1: values = (0 + c*sum) + 1*M
template <class TB, class TV>
// mode 0: values = sum of the blocks; mode 1: values = c * sum + cd * D2 + cm * M (D2, M optional)
void launch_block_gather(const NodePattern& pat, const BlockScatter& sc, const TB* blocks, TV* values, int mode,
                         double c, const double* M, bool zero_first, hipStream_t s, const double* D2 = nullptr,
                         double cd = 0.0, double cm = 1.0, KernelProf* prof = nullptr);
void launch_force_gather(Model& m, double* f);
template <class VT>
void launch_spmv(const NodePattern& pat, const VT* vals, double beta, double* y, double alpha, const double* x,
                 hipStream_t s);
template <class VT>
void launch_dirichlet(const NodePattern& pat, VT* vals, const uint8_t* d_dirmask, const int32_t* d_bc_nodes,
                      int64_t n_bc_nodes, double* d_scale, hipStream_t s);
int spmv_group_size(const NodePattern& pat);
// Times the node-blocked SpMV with 8, 16 and 32 lanes per node on this pattern and keeps the
// fastest in pat.tuned_group; returns it. Per-candidate averages (us) go to times_us[3] if given.
int autotune_spmv(NodePattern& pat, hipStream_t s, double* times_us);
void launch_element_psd_projection(Model& m);  // in place on m.d_ke
void launch_lincomb(hipStream_t s, int64_t n, double* out, double a, const double* A);  // out = 0 + a A
void rotation_svd_eval(int64_t nq, const double* F, double* U, double* S, double* Vt);
// psi[q] = Stable Neo-Hookean energy density of F[9 q ..] (Lame parameters as given, not reparametrized)
void material_energy(double mu, double lambda, int64_t nq, const double* F, double* psi);
void material_eval(int mat, double thr, double mu_r, double lambda_r, int64_t nq, const double* F, double* P,
                   double* dPdF);
int read_and_clear_err(Model& m);  // synchronizes the model stream

// ------------------------------------------------------------------ Schwarz (schwarz.hip)
void schwarz_setup(Schwarz& s, Model& m, int64_t nsub, const int64_t* off, const int32_t* nodes, int64_t nunt,
                   const int32_t* unt);  // s.fp32 selects the value type of P
void schwarz_update(Schwarz& s, const double* J);  // enqueues on the model stream; errors flagged in m.d_err
void schwarz_update(Schwarz& s, const float* J);
void schwarz_apply(Schwarz& s, const double* x, double* y);

// ------------------------------------------------------------------ CG (solver.hip)
// Solves A x = b. A has the model pattern; preconditioner = Schwarz or Jacobi(diag(A)).
hoefem_status pcg_solve(Model& m, CgWork& w, const double* A, Schwarz* P, const double* b, double* x, double tol,
                        int64_t max_iter, int64_t* iterations);
// pcg_solve (FP64) with every work buffer in a fixed per-device scratch array: no allocation;
// models of at most 262144 DOFs; calls overlapping in time on one GPU race
hoefem_status pcg_solve_noalloc(Model& m, const double* A, Schwarz* P, const double* b, double* x, double tol,
                                int64_t max_iter, int64_t* iterations);
// mixed precision: FP32 system matrix and FP32 Schwarz operator (P->fp32 must be set), FP64 vectors / dots
hoefem_status pcg_solve(Model& m, CgWork& w, const float* A, Schwarz* P, const double* b, double* x, double tol,
                        int64_t max_iter, int64_t* iterations);

// One system of a batched solve: the arguments of pcg_solve (A of value type double, or float
// with an FP32 Schwarz operator) and its results.
struct CgBatchSystem {
    Model* m = nullptr;
    CgWork* w = nullptr;
    const void* A = nullptr;
    Schwarz* P = nullptr;  // nullptr: Jacobi, whose diagonal must already be in w->diag_inv (launch_jacobi_diag)
    const double* b = nullptr;
    double* x = nullptr;
    double tol = 0.0;
    int64_t max_iter = -1;
    int64_t iterations = 0;            // out
    hoefem_status status = HOEFEM_OK;  // out: as pcg_solve returns it
};

// State of the batched solver that outlives one pool: its stream, the segment descriptors and
// work-group maps (pinned staging and device copies) and the captured iteration batches, keyed by
// their launch shapes.
struct CgBatch {
    hipStream_t cs = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};  // the two batches in flight
    hipEvent_t upload = nullptr;            // after the last descriptor / map upload
    std::vector<hipEvent_t> spare;          // reusable events
    std::unique_ptr<PinnedBuf<char>> h_seg;
    std::unique_ptr<PinnedBuf<int>> h_map;
    size_t seg_cap = 0, map_cap = 0;
    DevBuf<char> d_seg;
    DevBuf<int> d_map;
    struct Graph {
        std::vector<int64_t> key;
        hipGraphExec_t exec = nullptr;  // captured on the second sight of the key
        int seen = 0;
        uint64_t used = 0;
    };
    std::vector<Graph> graphs;
    uint64_t clock = 0;
    bool use_graphs = true;
    int64_t relayouts = 0;  // uploads of a new launch layout (statistics)
    void ensure(size_t seg_bytes, size_t map_ints);
    hipEvent_t take_event();
    ~CgBatch();
};

// Independent PCG solves that share launches (solver.hip): join() adds a system whose inputs are
// complete on the device (the caller has synchronized with them); progress() enqueues work and
// reports the systems whose solves ended, without blocking; system(id) then holds the iteration count
// and the status pcg_solve would return. Every system's x, iterations and status equal pcg_solve's bit
// for bit; each stops on the device when its solve ends.
struct CgPool {
    virtual ~CgPool() = default;
    virtual int join(const CgBatchSystem& s) = 0;
    virtual void progress(std::vector<int>& ended) = 0;
    virtual const CgBatchSystem& system(int id) const = 0;
    virtual bool idle() const = 0;
};
std::unique_ptr<CgPool> make_cg_pool(CgBatch& B, bool fp32, int capacity);
// the Jacobi diagonal of A into w.diag_inv on the model stream (zero diagonal entries flagged in m.d_err)
void launch_jacobi_diag(Model& m, CgWork& w, const void* A, bool fp32);

// ------------------------------------------------------------------ learned coarse correction (learned.hip)
constexpr int LC_MAX_AGG_NODES = 64;   // nodes per aggregate (LDS staging of the apply kernels)
constexpr int LC_SNAP_PER_SOLVE = 64;  // iterate snapshots kept per harvested solve

struct LearnedView {  // what the CG apply kernels read and write (no padding: its words key graphs)
    int32_t na;
    float kappa;              // output scale
    const int32_t* agg_ptr;   // [na + 1] offsets into agg_dofs (3 per node)
    const int32_t* agg_dofs;  // DOFs grouped by aggregate
    const float* basis;       // aggregate a: 64 modes x n_a DOFs at 64 * agg_ptr[a]
    const uint16_t* wbf;      // network weights (BF16) and
    const float* params;      // FP32 parameters (biases)
    float* x;                 // [na][64] scratch: normalized residual projections, BF16-rounded (network input)
    float* xf;                // [na][64] scratch: the same in FP32 (the sign guard's direction)
    float* gn;                // [na] scratch: their norms
    float* y;                 // [na][64] scratch: network outputs
};
static_assert(sizeof(LearnedView) == 8 * 10, "LearnedView has padding");

struct Learned {
    Model* model = nullptr;
    int32_t na = 0, max_dofs = 0;
    std::vector<int32_t> agg_ptr, agg_dofs;
    DevBuf<int32_t> d_agg_ptr, d_agg_dofs;
    DevBuf<uint8_t> d_dofmask;
    DevBuf<float> d_basis;
    bool has_basis = false, trained = false, enabled = false, harvesting = false;
    float kappa = 0.0f;  // mean |P r| / |r| over the first training batch
    DevBuf<float> d_cg_x, d_cg_xf, d_cg_gn, d_cg_y;  // per-iteration scratch of the CG apply (LearnedView)
    // network
    DevBuf<float> d_params, d_m, d_v, d_grad, d_gpart;
    DevBuf<uint16_t> d_wbf;
    int adam_t = 0;
    int64_t batch_cap = 0;
    DevBuf<float> d_z1, d_h1, d_z2, d_h2, d_y, d_dy, d_x, d_gn;
    void ensure_batch(int64_t N);
    // harvest rings: error snapshots x* - x_k and residuals r_k
    int cap = 0, every = 32, snap_max = 0;
    int64_t head = 0, count = 0, pairs = 0;
    DevBuf<double> d_err, d_res, d_snap_x, d_snap_r;
    // statistics of the learned solves
    int64_t solves = 0, fallbacks = 0, iterations = 0;
    int last_status = 0;  // status of the last learned solve (0: converged)
};

void learned_setup(Learned& L, Model& m, int target_aggregates, const std::vector<uint8_t>& dir_node_mask,
                   uint64_t seed);
void learned_set_params(Learned& L, const float* host_or_dev, bool on_device);
void learned_harvest_enable(Learned& L, bool on, int capacity, int every);
void learned_harvest_solve(Learned& L, const double* xstar, int nsnap);
void learned_build_basis(Learned& L);
double learned_train(Learned& L, const double* J, Schwarz* P, int steps, double lr, int B, double* initial_loss);
double learned_train(Learned& L, const float* J, Schwarz* P, int steps, double lr, int B, double* initial_loss);
void learned_mlp_forward(Learned& L, int64_t N, const float* X, float* Y);
void learned_mlp_grad(Learned& L, int64_t N, const float* X, const float* dY, float* grad, float scale);
LearnedView learned_view(Learned& L);
// Y = MLP(X) for N samples on stream s (the network of the view), without the training tape
void launch_mlp_forward(const LearnedView& lv, int64_t N, const float* X, float* Y, hipStream_t s);
// the sign guard of the correction (learned.hip): y of the samples with xf.y < 0 projected orthogonal to xf
void launch_lc_guard(int64_t N, const float* xf, float* y, uint8_t* flag, hipStream_t s);

// Snapshot hook of the plain CG: every `every` iterations x_k and r_k are copied into
// xs / rs (at most max snapshots); *taken receives the count.
struct CgHarvest {
    double* xs = nullptr;
    double* rs = nullptr;
    int max = 0, every = 32;
    int taken = 0;
};
hoefem_status pcg_solve_harvest(Model& m, CgWork& w, const double* A, Schwarz* P, const double* b, double* x,
                                double tol, int64_t max_iter, int64_t* iterations, CgHarvest* h);
hoefem_status pcg_solve_harvest(Model& m, CgWork& w, const float* A, Schwarz* P, const double* b, double* x,
                                double tol, int64_t max_iter, int64_t* iterations, CgHarvest* h);
// flexible PCG (Polak-Ribiere beta) with z = P r + learned coarse correction
hoefem_status pcg_solve_learned(Model& m, CgWork& w, const double* A, Schwarz* P, const LearnedView& lv,
                                const double* b, double* x, double tol, int64_t max_iter, int64_t* iterations);
hoefem_status pcg_solve_learned(Model& m, CgWork& w, const float* A, Schwarz* P, const LearnedView& lv,
                                const double* b, double* x, double tol, int64_t max_iter, int64_t* iterations);

}  // namespace hoefem
