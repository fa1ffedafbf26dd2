// Learned coarse correction of the Schwarz preconditioner.
//
// The nodes are split into aggregates: the breadth-first order of the node graph cut into the
// requested number of contiguous, equal pieces. Each aggregate a carries 64 orthonormal modes Phi_a: the restrictions to its
// DOFs of 64 error snapshots x* - x_k harvested from the plain Schwarz-CG solver's own solves
// (orthonormalized per aggregate by Gram-Schmidt, Dirichlet DOFs zeroed). One network shared by
// all aggregates maps the aggregate's residual projection to a correction in its modes:
//     g_a = Phi_a^T r,    c_a = kappa |g_a| MLP(g_a / |g_a|),    z = P r + sum_a Phi_a c_a,
// a 64-dim input and a 64-dim output per aggregate. A sign guard replaces an output y with
// (g / |g|).y < 0 by its projection orthogonal to g, so every aggregate adds a non-negative amount to
// r^T z >= r^T P r > 0 and the preconditioned residual keeps a positive inner product. The correction
// is positively homogeneous in r (CG is blind to the scale of z) but not linear, so the solver that
// uses it is the flexible (Polak-Ribiere) PCG; kappa = mean |P r| / |r| over the first training batch keeps the network's
// output O(1). Training minimizes, over harvested residuals r (unit norm),
//     L(z) = -(r^T z)^2 / (z^T J z)  = min over s of |J^-1 r - s z|_J^2 - |J^-1 r|_J^2,
// the energy left after the best step along z: the quality of z as a search direction,
// independent of its scale. The gradient reaches the network through
// dL/dz = (2 a / q) ((a / q) J z - r), a = r^T z, q = z^T J z, then dc_a = Phi_a^T dL/dz,
// dy_a = kappa |g_a| dc_a and, for a guarded sample, its projection orthogonal to g_a. The last layer starts at zero, so the untrained correction vanishes
// (the plain Schwarz preconditioner) while the gradients of its weights do not.
//
// Kernels: per-aggregate Gram-Schmidt basis, projection and prolongation, batched network
// forward / backward (FP32 accumulation over BF16-rounded weights and activations),
// deterministic gradient reduction, Adam. The per-iteration projection and apply kernels live
// with the CG kernels; they run this network through launch_mlp_forward.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.hpp"
#include "mlp_device.hpp"

namespace hoefem {

// ------------------------------------------------------------------ host setup
void learned_setup(Learned& L, Model& m, int target_aggregates, const std::vector<uint8_t>& dir_node_mask,
                   uint64_t seed) {
    L.model = &m;
    const int64_t nv = m.nv;
    if (target_aggregates < 1) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "need at least one aggregate");
    const int64_t na64 = std::min<int64_t>(target_aggregates, std::max<int64_t>(nv, 1));
    const int64_t target = (nv + na64 - 1) / na64;
    if (target > LC_MAX_AGG_NODES)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "aggregates would exceed " + std::to_string(LC_MAX_AGG_NODES) +
                                                       " nodes; ask for more aggregates");
    // breadth-first order of the node graph (restarting at the lowest unvisited node), cut into
    // na64 contiguous pieces of equal size: exactly the requested number of aggregates, each a run
    // of consecutive breadth-first layers
    std::vector<int32_t> order;
    order.reserve(nv);
    std::vector<uint8_t> seen(nv, 0);
    const NodePattern& pat = m.pattern;
    for (int64_t root = 0; root < nv; ++root) {
        if (seen[root]) continue;
        seen[root] = 1;
        size_t head = order.size();
        order.push_back((int32_t)root);
        for (; head < order.size(); ++head) {
            const int32_t a = order[head];
            for (int32_t k = pat.nptr[a]; k < pat.nptr[a + 1]; ++k) {
                const int32_t b = pat.ncol[k];
                if (!seen[b]) {
                    seen[b] = 1;
                    order.push_back(b);
                }
            }
        }
    }
    std::vector<int32_t> agg(nv, -1);
    for (int64_t pos = 0; pos < nv; ++pos) agg[order[pos]] = (int32_t)(pos * na64 / std::max<int64_t>(nv, 1));
    const int32_t na = (int32_t)na64;
    L.na = na;
    L.agg_ptr.assign(na + 1, 0);
    for (int64_t v = 0; v < nv; ++v) L.agg_ptr[agg[v] + 1] += 3;
    for (int32_t a = 0; a < na; ++a) L.agg_ptr[a + 1] += L.agg_ptr[a];
    L.agg_dofs.assign(3 * nv, 0);
    std::vector<int32_t> fill(L.agg_ptr.begin(), L.agg_ptr.end() - 1);
    for (int64_t v = 0; v < nv; ++v)
        for (int i = 0; i < 3; ++i) L.agg_dofs[fill[agg[v]]++] = (int32_t)(3 * v + i);
    L.max_dofs = 0;
    for (int32_t a = 0; a < na; ++a) L.max_dofs = std::max(L.max_dofs, L.agg_ptr[a + 1] - L.agg_ptr[a]);
    std::vector<uint8_t> dofmask(3 * nv, 0);
    for (int64_t v = 0; v < nv; ++v)
        if (v < (int64_t)dir_node_mask.size() && dir_node_mask[v]) dofmask[3 * v] = dofmask[3 * v + 1] = dofmask[3 * v + 2] = 1;
    L.d_agg_ptr.upload(L.agg_ptr, m.stream);
    L.d_agg_dofs.upload(L.agg_dofs, m.stream);
    L.d_dofmask.upload(dofmask, m.stream);
    L.d_basis.alloc((size_t)LC_IN * 3 * nv);
    HIP_CHECK(hipMemsetAsync(L.d_basis.get(), 0, sizeof(float) * L.d_basis.n, m.stream));
    // parameters: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for W1, W2 from a counter-based hash;
    // W3 and every bias zero, so the untrained correction vanishes
    std::vector<float> p(LC_PARAMS, 0.0f);
    auto unif = [&](uint64_t i) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);  // splitmix64
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < LC_HID * LC_IN; ++i) p[LC_W1 + i] = (float)((2.0 * unif(LC_W1 + i) - 1.0) / std::sqrt((double)LC_IN));
    for (int i = 0; i < LC_HID * LC_HID; ++i)
        p[LC_W2 + i] = (float)((2.0 * unif(LC_W2 + i) - 1.0) / std::sqrt((double)LC_HID));
    learned_set_params(L, p.data(), false);
    L.d_m.alloc(LC_PARAMS);
    L.d_v.alloc(LC_PARAMS);
    HIP_CHECK(hipMemsetAsync(L.d_m.get(), 0, sizeof(float) * LC_PARAMS, m.stream));
    HIP_CHECK(hipMemsetAsync(L.d_v.get(), 0, sizeof(float) * LC_PARAMS, m.stream));
    L.adam_t = 0;
    L.d_grad.alloc(LC_PARAMS);
}

__global__ void k_params_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < LC_PARAMS) w[i] = bf16_bits(p[i]);
}

void learned_set_params(Learned& L, const float* host_or_dev, bool on_device) {
    Model& m = *L.model;
    L.d_params.ensure(LC_PARAMS);
    L.d_wbf.ensure(LC_PARAMS);
    if (on_device)
        HIP_CHECK(hipMemcpyAsync(L.d_params.get(), host_or_dev, sizeof(float) * LC_PARAMS, hipMemcpyDeviceToDevice,
                                 m.stream));
    else {
        HIP_CHECK(hipStreamSynchronize(m.stream));
        HIP_CHECK(hipMemcpy(L.d_params.get(), host_or_dev, sizeof(float) * LC_PARAMS, hipMemcpyHostToDevice));
    }
    k_params_to_bf16<<<dim3((LC_PARAMS + 255) / 256), dim3(256), 0, m.stream>>>(L.d_params.get(), L.d_wbf.get());
    check_launch("k_params_to_bf16");
}

// ------------------------------------------------------------------ harvesting
// e_k = x* - x_k for the snapshots of one solve, appended to the error ring; r_k to the
// residual ring (both overwrite the oldest entries once full).
__global__ void k_lc_snap_errors(int64_t n, int nsnap, const double* __restrict__ xs, const double* __restrict__ rs,
                                 const double* __restrict__ xstar, double* __restrict__ err_ring,
                                 double* __restrict__ res_ring, int cap, int head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = xstar[i];
    for (int k = 0; k < nsnap; ++k) {
        const int slot = (head + k) % cap;
        err_ring[(int64_t)slot * n + i] = xi - xs[(int64_t)k * n + i];
        res_ring[(int64_t)slot * n + i] = rs[(int64_t)k * n + i];
    }
}

void learned_harvest_solve(Learned& L, const double* xstar, int nsnap) {
    Model& m = *L.model;
    if (nsnap <= 0) return;
    const int64_t n = 3 * m.nv;
    k_lc_snap_errors<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(
        n, nsnap, L.d_snap_x.get(), L.d_snap_r.get(), xstar, L.d_err.get(), L.d_res.get(), L.cap, L.head);
    check_launch("k_lc_snap_errors");
    L.head = (L.head + nsnap) % L.cap;
    L.count = std::min<int64_t>(L.cap, L.count + nsnap);
    L.pairs += nsnap;
}

void learned_harvest_enable(Learned& L, bool on, int capacity, int every) {
    Model& m = *L.model;
    const int64_t n = 3 * m.nv;
    if (on) {
        if (capacity < LC_IN) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "harvest capacity must be >= 64 snapshots");
        if (L.cap != capacity) {
            L.cap = capacity;
            L.d_err.alloc((size_t)capacity * n);
            L.d_res.alloc((size_t)capacity * n);
            L.head = 0;
            L.count = 0;
        }
        L.every = std::max(1, every);
        L.snap_max = LC_SNAP_PER_SOLVE;
        L.d_snap_x.ensure((size_t)L.snap_max * n);
        L.d_snap_r.ensure((size_t)L.snap_max * n);
    }
    L.harvesting = on;
}

// ------------------------------------------------------------------ basis
// Classical Gram-Schmidt with re-orthogonalization over the aggregate's restrictions of
// nsnap error snapshots (FP64 in LDS). A vector whose norm falls below 1e-10 of its original
// norm is dropped (its mode stays zero).
__global__ __launch_bounds__(256) void k_lc_basis(const int32_t* __restrict__ agg_ptr, const int32_t* __restrict__ agg_dofs,
                                                  const uint8_t* __restrict__ dofmask, const double* __restrict__ err,
                                                  const int32_t* __restrict__ snap_slot, int nsnap, int64_t n,
                                                  float* __restrict__ basis, int ld) {
    extern __shared__ double S[];  // [64][ld]
    __shared__ double s_c[LC_IN];
    __shared__ double s_red[256];
    const int a = blockIdx.x;
    const int o = agg_ptr[a], nd = agg_ptr[a + 1] - o;
    const int t = threadIdx.x;
    for (int idx = t; idx < LC_IN * nd; idx += 256) {
        const int j = idx / nd, d = idx % nd;
        const int dof = agg_dofs[o + d];
        S[j * ld + d] = (j < nsnap && !dofmask[dof]) ? err[(int64_t)snap_slot[j] * n + dof] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < LC_IN; ++j) {
        // original norm
        double loc = 0.0;
        for (int d = t; d < nd; d += 256) loc += S[j * ld + d] * S[j * ld + d];
        s_red[t] = loc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) s_red[t] += s_red[t + s];
            __syncthreads();
        }
        const double n0 = sqrt(s_red[0]);
        __syncthreads();
        for (int pass = 0; pass < 2; ++pass) {
            // c_i = q_i . v_j for i < j: 4 threads per i
            for (int i = t / 4; i < j; i += 64) {
                double c = 0.0;
                for (int d = t % 4; d < nd; d += 4) c += S[i * ld + d] * S[j * ld + d];
                c += __shfl_xor(c, 1, 4);
                c += __shfl_xor(c, 2, 4);
                if (t % 4 == 0) s_c[i] = c;
            }
            __syncthreads();
            for (int d = t; d < nd; d += 256) {
                double v = S[j * ld + d];
                for (int i = 0; i < j; ++i) v -= s_c[i] * S[i * ld + d];
                S[j * ld + d] = v;
            }
            __syncthreads();
        }
        loc = 0.0;
        for (int d = t; d < nd; d += 256) loc += S[j * ld + d] * S[j * ld + d];
        s_red[t] = loc;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) s_red[t] += s_red[t + s];
            __syncthreads();
        }
        const double nr = sqrt(s_red[0]);
        __syncthreads();
        const double scale = (nr > 1e-10 * n0 && nr > 0.0) ? 1.0 / nr : 0.0;
        for (int d = t; d < nd; d += 256) S[j * ld + d] *= scale;
        __syncthreads();
    }
    for (int idx = t; idx < LC_IN * nd; idx += 256) {
        const int j = idx / nd, d = idx % nd;
        basis[(int64_t)LC_IN * o + (int64_t)j * nd + d] = (float)S[j * ld + d];
    }
}

void learned_build_basis(Learned& L) {
    Model& m = *L.model;
    if (L.count < LC_IN)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT,
                    "the learned correction needs at least 64 harvested snapshots (have " + std::to_string(L.count) + ")");
    // 64 snapshots spread evenly over the ring's contents, oldest first
    std::vector<int32_t> slots(LC_IN);
    const int64_t oldest = (L.head - L.count + L.cap) % L.cap;
    for (int j = 0; j < LC_IN; ++j) {
        const int64_t k = (int64_t)j * (L.count - 1) / (LC_IN - 1);
        slots[j] = (int32_t)((oldest + k) % L.cap);
    }
    DevBuf<int32_t> d_slots;
    d_slots.upload(slots, m.stream);
    const int ld = L.max_dofs + 1;
    const size_t lds = (size_t)LC_IN * ld * sizeof(double);
    if (lds > 150 * 1024) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "aggregate too large for the basis kernel");
    raise_lds_limit((const void*)k_lc_basis, 150 * 1024);
    k_lc_basis<<<dim3((unsigned)L.na), dim3(256), lds, m.stream>>>(L.d_agg_ptr.get(), L.d_agg_dofs.get(),
                                                                   L.d_dofmask.get(), L.d_err.get(), d_slots.get(),
                                                                   LC_IN, 3 * m.nv, L.d_basis.get(), ld);
    check_launch("k_lc_basis");
    HIP_CHECK(hipStreamSynchronize(m.stream));
    L.has_basis = true;
}

// ------------------------------------------------------------------ projection / prolongation
// G[b][a][j] = g_j / |g| (normalize 1: rounded to BF16, 2: FP32; 0: g itself) with g = Phi_a^T vec_b
// restricted to a; gnorm[b][a] = |g| when given
__global__ __launch_bounds__(256) void k_lc_project(int na, const int32_t* __restrict__ agg_ptr,
                                                    const int32_t* __restrict__ agg_dofs, const float* __restrict__ basis,
                                                    const double* __restrict__ vecs, int64_t n, float* __restrict__ G,
                                                    float* __restrict__ gnorm, int normalize) {
    __shared__ double s_r[LC_MAX_AGG_NODES * 3];
    __shared__ double s_g[LC_IN];
    const int a = blockIdx.x, b = blockIdx.y;
    const int o = agg_ptr[a], nd = agg_ptr[a + 1] - o;
    const int t = threadIdx.x;
    const double* vec = vecs + (int64_t)b * n;
    for (int d = t; d < nd; d += 256) s_r[d] = vec[agg_dofs[o + d]];
    __syncthreads();
    const float* B = basis + (int64_t)LC_IN * o;
    {
        const int j = t / 4;
        double c = 0.0;
        for (int d = t % 4; d < nd; d += 4) c += (double)B[j * nd + d] * s_r[d];
        c += __shfl_xor(c, 1, 4);
        c += __shfl_xor(c, 2, 4);
        if (t % 4 == 0) s_g[j] = c;
    }
    __syncthreads();
    if (t < 64) {
        double q = s_g[t] * s_g[t];
        for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
        const double nrm = sqrt(q);
        float* out = G + ((int64_t)b * na + a) * LC_IN;
        if (normalize) {
            const float xn = nrm > 0.0 ? (float)(s_g[t] / nrm) : 0.0f;
            out[t] = normalize == 1 ? bf16_round(xn) : xn;
            if (t == 0 && gnorm) gnorm[(int64_t)b * na + a] = (float)nrm;
        } else {
            out[t] = (float)s_g[t];
        }
    }
}

// out_b[dof] = sum_j Phi_a[j][dof] * kappa * scale_b_a * C[b][a][j]  (every DOF belongs to one aggregate)
__global__ __launch_bounds__(256) void k_lc_prolong(int na, const int32_t* __restrict__ agg_ptr,
                                                    const int32_t* __restrict__ agg_dofs, const float* __restrict__ basis,
                                                    const float* __restrict__ C, const float* __restrict__ scale,
                                                    float kappa, int64_t n, double* __restrict__ outs) {
    __shared__ float s_c[LC_IN];
    const int a = blockIdx.x, b = blockIdx.y;
    const int o = agg_ptr[a], nd = agg_ptr[a + 1] - o;
    const int t = threadIdx.x;
    const float sc = (scale ? scale[(int64_t)b * na + a] : 1.0f) * kappa;
    if (t < LC_IN) s_c[t] = C[((int64_t)b * na + a) * LC_IN + t] * sc;
    __syncthreads();
    const float* B = basis + (int64_t)LC_IN * o;
    for (int d = t; d < nd; d += 256) {
        double acc = 0.0;
        for (int j = 0; j < LC_IN; ++j) acc += (double)B[j * nd + d] * (double)s_c[j];
        outs[(int64_t)b * n + agg_dofs[o + d]] = acc;
    }
}

// ------------------------------------------------------------------ network, batched
constexpr int MLP_TS = 32;  // samples per work-group

// Z1 = W1 x + b1, H1 = bf16(GELU(Z1)), Z2 = W2 H1 + b2, H2 = bf16(GELU(Z2)), Y = W3 H2 + b3
// (the tape Z1, H1, Z2, H2 for the backward pass is written when Z1 is given)
__global__ __launch_bounds__(256) void k_mlp_fwd(int64_t N, const float* __restrict__ X, const uint16_t* __restrict__ wbf,
                                                 const float* __restrict__ params, float* __restrict__ Z1,
                                                 float* __restrict__ H1, float* __restrict__ Z2, float* __restrict__ H2,
                                                 float* __restrict__ Y) {
    __shared__ float sx[MLP_TS][LC_IN + 1];
    __shared__ float sh1[MLP_TS][LC_HID + 1];
    __shared__ float sh2[MLP_TS][LC_HID + 1];
    const int t = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * MLP_TS;
    for (int idx = t; idx < MLP_TS * LC_IN; idx += 256) {
        const int s = idx / LC_IN, i = idx % LC_IN;
        sx[s][i] = (s0 + s < N) ? bf16_round(X[(s0 + s) * LC_IN + i]) : 0.0f;
    }
    __syncthreads();
    {
        const int o = t % LC_HID;
        float w[LC_IN];
#pragma unroll
        for (int i = 0; i < LC_IN; ++i) w[i] = bf16_float(wbf[LC_W1 + o * LC_IN + i]);
        const float bias = params[LC_B1 + o];
        for (int s = t / LC_HID; s < MLP_TS; s += 256 / LC_HID) {
            float acc = 0.0f;
#pragma unroll
            for (int i = 0; i < LC_IN; ++i) acc += w[i] * sx[s][i];
            acc += bias;
            const float h = bf16_round(gelu(acc));
            sh1[s][o] = h;
            if (Z1 && s0 + s < N) {
                Z1[(s0 + s) * LC_HID + o] = acc;
                H1[(s0 + s) * LC_HID + o] = h;
            }
        }
    }
    __syncthreads();
    {
        const int o = t % LC_HID;
        const float bias = params[LC_B2 + o];
        for (int s = t / LC_HID; s < MLP_TS; s += 256 / LC_HID) {
            float acc = 0.0f;
            for (int i = 0; i < LC_HID; ++i) acc += bf16_float(wbf[LC_W2 + o * LC_HID + i]) * sh1[s][i];
            acc += bias;
            const float h = bf16_round(gelu(acc));
            sh2[s][o] = h;
            if (Z2 && s0 + s < N) {
                Z2[(s0 + s) * LC_HID + o] = acc;
                H2[(s0 + s) * LC_HID + o] = h;
            }
        }
    }
    __syncthreads();
    {
        const int o = t % LC_OUT;
        const float bias = params[LC_B3 + o];
        for (int s = t / LC_OUT; s < MLP_TS; s += 256 / LC_OUT) {
            float acc = 0.0f;
            for (int i = 0; i < LC_HID; ++i) acc += bf16_float(wbf[LC_W3 + o * LC_HID + i]) * sh2[s][i];
            acc += bias;
            if (s0 + s < N) Y[(s0 + s) * LC_OUT + o] = acc;
        }
    }
}

// Parameter gradients of sum_s dY[s] . Y[s] over the work-group's samples -> gpart[blockIdx]
__global__ __launch_bounds__(256) void k_mlp_bwd(int64_t N, const float* __restrict__ X, const uint16_t* __restrict__ wbf,
                                                 const float* __restrict__ Z1, const float* __restrict__ H1,
                                                 const float* __restrict__ Z2, const float* __restrict__ H2,
                                                 const float* __restrict__ dY, float* __restrict__ gpart) {
    __shared__ float sx[MLP_TS][LC_IN + 1];
    __shared__ float sa[MLP_TS][LC_HID + 1];  // H1, later dZ1
    __shared__ float sb[MLP_TS][LC_HID + 1];  // H2, later dZ2
    __shared__ float sd[MLP_TS][LC_OUT + 1];  // dY
    const int t = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * MLP_TS;
    float* g = gpart + (int64_t)blockIdx.x * LC_PARAMS;
    for (int idx = t; idx < MLP_TS * LC_HID; idx += 256) {
        const int s = idx / LC_HID, i = idx % LC_HID;
        const bool ok = s0 + s < N;
        sa[s][i] = ok ? H1[(s0 + s) * LC_HID + i] : 0.0f;
        sb[s][i] = ok ? H2[(s0 + s) * LC_HID + i] : 0.0f;
        if (i < LC_IN) sx[s][i] = ok ? bf16_round(X[(s0 + s) * LC_IN + i]) : 0.0f;
        if (i < LC_OUT) sd[s][i] = ok ? dY[(s0 + s) * LC_OUT + i] : 0.0f;
    }
    __syncthreads();
    // dW3[o][i] = sum_s dY[s][o] H2[s][i]; db3[o] = sum_s dY[s][o]
    for (int idx = t; idx < LC_OUT * LC_HID; idx += 256) {
        const int o = idx / LC_HID, i = idx % LC_HID;
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sd[s][o] * sb[s][i];
        g[LC_W3 + idx] = acc;
    }
    if (t < LC_OUT) {
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sd[s][t];
        g[LC_B3 + t] = acc;
    }
    __syncthreads();
    // dZ2[s][i] = (sum_o W3[o][i] dY[s][o]) GELU'(Z2[s][i]) -> sb (H2 is no longer read)
    {
        const int i = t % LC_HID;
        for (int s = t / LC_HID; s < MLP_TS; s += 256 / LC_HID) {
            float acc = 0.0f;
            for (int o = 0; o < LC_OUT; ++o) acc += bf16_float(wbf[LC_W3 + o * LC_HID + i]) * sd[s][o];
            const float z = (s0 + s < N) ? Z2[(s0 + s) * LC_HID + i] : 0.0f;
            sb[s][i] = (s0 + s < N) ? acc * gelu_grad(z) : 0.0f;
        }
    }
    __syncthreads();
    // dW2[o][i] = sum_s dZ2[s][o] H1[s][i]; db2
    for (int idx = t; idx < LC_HID * LC_HID; idx += 256) {
        const int o = idx / LC_HID, i = idx % LC_HID;
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sb[s][o] * sa[s][i];
        g[LC_W2 + idx] = acc;
    }
    if (t < LC_HID) {
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sb[s][t];
        g[LC_B2 + t] = acc;
    }
    __syncthreads();
    // dZ1[s][i] = (sum_o W2[o][i] dZ2[s][o]) GELU'(Z1[s][i]) -> sa (H1 is no longer read)
    {
        const int i = t % LC_HID;
        for (int s = t / LC_HID; s < MLP_TS; s += 256 / LC_HID) {
            float acc = 0.0f;
            for (int o = 0; o < LC_HID; ++o) acc += bf16_float(wbf[LC_W2 + o * LC_HID + i]) * sb[s][o];
            const float z = (s0 + s < N) ? Z1[(s0 + s) * LC_HID + i] : 0.0f;
            sa[s][i] = (s0 + s < N) ? acc * gelu_grad(z) : 0.0f;
        }
    }
    __syncthreads();
    // dW1[o][i] = sum_s dZ1[s][o] X[s][i]; db1
    for (int idx = t; idx < LC_HID * LC_IN; idx += 256) {
        const int o = idx / LC_IN, i = idx % LC_IN;
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sa[s][o] * sx[s][i];
        g[LC_W1 + idx] = acc;
    }
    if (t < LC_HID) {
        float acc = 0.0f;
        for (int s = 0; s < MLP_TS; ++s) acc += sa[s][t];
        g[LC_B1 + t] = acc;
    }
}

// grad[p] = scale * sum over work-groups (fixed order) of gpart[wg][p]
__global__ void k_mlp_reduce(int nwg, const float* __restrict__ gpart, float scale, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= LC_PARAMS) return;
    double acc = 0.0;
    for (int w = 0; w < nwg; ++w) acc += (double)gpart[(int64_t)w * LC_PARAMS + p];
    grad[p] = (float)(acc * scale);
}

// Adam (Kingma & Ba, 2015) on the FP32 master parameters, then the BF16 weight copy
__global__ void k_adam(const float* __restrict__ grad, float* __restrict__ p, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ wbf, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= LC_PARAMS) return;
    const float gi = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float w = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = w;
    wbf[i] = bf16_bits(w);
}

void learned_mlp_forward(Learned& L, int64_t N, const float* X, float* Y) {
    Model& m = *L.model;
    L.ensure_batch(N);
    const unsigned nwg = (unsigned)((N + MLP_TS - 1) / MLP_TS);
    k_mlp_fwd<<<dim3(nwg), dim3(256), 0, m.stream>>>(N, X, L.d_wbf.get(), L.d_params.get(), L.d_z1.get(),
                                                     L.d_h1.get(), L.d_z2.get(), L.d_h2.get(), Y);
    check_launch("k_mlp_fwd");
}

// grad = d(sum_s dY[s].Y[s]) / d params (forward recomputed on X)
void learned_mlp_grad(Learned& L, int64_t N, const float* X, const float* dY, float* grad, float scale) {
    Model& m = *L.model;
    L.ensure_batch(N);
    const unsigned nwg = (unsigned)((N + MLP_TS - 1) / MLP_TS);
    k_mlp_fwd<<<dim3(nwg), dim3(256), 0, m.stream>>>(N, X, L.d_wbf.get(), L.d_params.get(), L.d_z1.get(),
                                                     L.d_h1.get(), L.d_z2.get(), L.d_h2.get(), L.d_y.get());
    k_mlp_bwd<<<dim3(nwg), dim3(256), 0, m.stream>>>(N, X, L.d_wbf.get(), L.d_z1.get(), L.d_h1.get(), L.d_z2.get(),
                                                     L.d_h2.get(), dY, L.d_gpart.get());
    k_mlp_reduce<<<dim3((LC_PARAMS + 255) / 256), dim3(256), 0, m.stream>>>((int)nwg, L.d_gpart.get(), scale, grad);
    check_launch("k_mlp_bwd");
}

void Learned::ensure_batch(int64_t N) {
    if (batch_cap >= N) return;
    batch_cap = N;
    d_z1.alloc((size_t)N * LC_HID);
    d_h1.alloc((size_t)N * LC_HID);
    d_z2.alloc((size_t)N * LC_HID);
    d_h2.alloc((size_t)N * LC_HID);
    d_y.alloc((size_t)N * LC_OUT);
    d_dy.alloc((size_t)N * LC_OUT);
    d_x.alloc((size_t)N * LC_IN);
    d_gn.alloc((size_t)N);
    d_gpart.alloc((size_t)((N + MLP_TS - 1) / MLP_TS) * LC_PARAMS);
}

// ------------------------------------------------------------------ training
__global__ void k_lc_select(int64_t n, int B, const double* __restrict__ ring, const int32_t* __restrict__ slots,
                            const double* __restrict__ inv_norm, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int b = 0; b < B; ++b) out[(int64_t)b * n + i] = ring[(int64_t)slots[b] * n + i] * inv_norm[b];
}

// z = zs + y ; grad vector dz = 2 (Jz - r)  (computed after Jz is known); loss terms per DOF
__global__ void k_lc_add(int64_t m, const double* __restrict__ a, double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) b[i] += a[i];
}

// dz = (2 a / q) ((a / q) Jz - r) per sample (a = r.z, q = z.Jz in aq[2 b], aq[2 b + 1])
__global__ void k_lc_graddz(int64_t n, int B, const double* __restrict__ aq, const double* __restrict__ Jz,
                            const double* __restrict__ r, double* __restrict__ dz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * B) return;
    const int b = (int)(i / n);
    const double a = aq[2 * b], q = aq[2 * b + 1];
    const double ratio = q > 0.0 ? a / q : 0.0;
    dz[i] = 2.0 * ratio * (ratio * Jz[i] - r[i]);
}

// per-sample squared norm (one work-group per sample vector)
__global__ __launch_bounds__(256) void k_lc_norm2(int64_t n, const double* __restrict__ z, double* __restrict__ out) {
    __shared__ double s_red[256];
    const int b = blockIdx.x;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double zi = z[(int64_t)b * n + i];
        acc += zi * zi;
    }
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_red[threadIdx.x] += s_red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = s_red[0];
}

// per-sample a = r.z and q = z.Jz -> aq[2 b], aq[2 b + 1] (one work-group per sample vector)
__global__ __launch_bounds__(256) void k_lc_dots(int64_t n, const double* __restrict__ z, const double* __restrict__ Jz,
                                                 const double* __restrict__ r, double* __restrict__ aq) {
    __shared__ double s_a[256], s_q[256];
    const int b = blockIdx.x;
    double a = 0.0, q = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double zi = z[(int64_t)b * n + i];
        a += r[(int64_t)b * n + i] * zi;
        q += zi * Jz[(int64_t)b * n + i];
    }
    s_a[threadIdx.x] = a;
    s_q[threadIdx.x] = q;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_a[threadIdx.x] += s_a[threadIdx.x + s];
            s_q[threadIdx.x] += s_q[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        aq[2 * b] = s_a[0];
        aq[2 * b + 1] = s_q[0];
    }
}

// dY[i] = kappa gnorm[i / 64] dC[i]: the gradient through c = kappa |g| MLP(g / |g|)
__global__ void k_lc_out_grad(int64_t total, const float* __restrict__ gnorm, const float* __restrict__ dC, float kappa,
                              float* __restrict__ dY) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    dY[i] = kappa * gnorm[i / LC_OUT] * dC[i];
}

// The sign guard of the correction c = kappa |g| y: with xf = g / |g|, r^T Phi_a c = kappa |g|^2 xf.y,
// so a sample with xf.y < 0 has y replaced by its projection y - (xf.y / xf.xf) xf, orthogonal to xf.
// Every aggregate's correction then adds a non-negative amount to r^T z = r^T P r + sum_a r^T Phi_a c_a,
// which keeps the preconditioned residual's inner product positive (no IndefinitePreconditioner
// breakdown). flag[s] records the samples that were projected (for the backward pass). One 64-lane
// wavefront per sample.
__global__ __launch_bounds__(256) void k_lc_guard(int64_t N, const float* __restrict__ xf, float* __restrict__ y,
                                                  uint8_t* __restrict__ flag) {
    const int64_t smp = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int j = threadIdx.x % 64;
    if (smp >= N) return;
    const float xv = xf[smp * LC_IN + j], yv = y[smp * LC_OUT + j];
    float sy = xv * yv, sq = xv * xv;
    for (int off = 32; off > 0; off >>= 1) {
        sy += __shfl_xor(sy, off, 64);
        sq += __shfl_xor(sq, off, 64);
    }
    const bool proj = sy < 0.0f && sq > 0.0f;
    if (proj) y[smp * LC_OUT + j] = yv - (sy / sq) * xv;
    if (flag && j == 0) flag[smp] = proj ? 1 : 0;
}

// dY <- (I - xf xf^T / xf.xf) dY for the samples the guard projected
__global__ __launch_bounds__(256) void k_lc_guard_grad(int64_t N, const float* __restrict__ xf,
                                                       const uint8_t* __restrict__ flag, float* __restrict__ dy) {
    const int64_t smp = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int j = threadIdx.x % 64;
    if (smp >= N || !flag[smp]) return;
    const float xv = xf[smp * LC_IN + j], dv = dy[smp * LC_OUT + j];
    float sd = xv * dv, sq = xv * xv;
    for (int off = 32; off > 0; off >>= 1) {
        sd += __shfl_xor(sd, off, 64);
        sq += __shfl_xor(sq, off, 64);
    }
    dy[smp * LC_OUT + j] = dv - (sd / sq) * xv;
}

void launch_lc_guard(int64_t N, const float* xf, float* y, uint8_t* flag, hipStream_t s) {
    if (N <= 0) return;
    k_lc_guard<<<dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s>>>(N, xf, y, flag);
    check_launch("k_lc_guard");
}

void launch_mlp_forward(const LearnedView& lv, int64_t N, const float* X, float* Y, hipStream_t s) {
    if (N <= 0) return;
    const unsigned nwg = (unsigned)((N + MLP_TS - 1) / MLP_TS);
    k_mlp_fwd<<<dim3(nwg), dim3(256), 0, s>>>(N, X, lv.wbf, lv.params, nullptr, nullptr, nullptr, nullptr, Y);
    check_launch("k_mlp_fwd");
}

template <class VT>
static double learned_train_t(Learned& L, const VT* Jvals, Schwarz* P, int steps, double lr, int B,
                              double* initial_loss) {
    Model& m = *L.model;
    const hipStream_t s = m.stream;
    const int64_t n = 3 * m.nv;
    if (!L.has_basis) learned_build_basis(L);
    if (L.count < 1) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "no harvested residuals");
    B = std::max(1, std::min<int>(B, (int)L.count));
    const int64_t NS = (int64_t)B * L.na;   // network samples of a batch: one per (residual, aggregate)
    const int64_t NG = NS * LC_IN;          // their projections
    L.ensure_batch(NS);
    DevBuf<double> r, zs, y, Jz, dz, aq;
    r.alloc((size_t)B * n);
    zs.alloc((size_t)B * n);
    y.alloc((size_t)B * n);
    Jz.alloc((size_t)B * n);
    dz.alloc((size_t)B * n);
    aq.alloc(2 * (size_t)B);
    DevBuf<float> G, Gf, gnorm, dC;
    DevBuf<uint8_t> gflag;
    G.alloc(NG);
    Gf.alloc(NG);
    gnorm.alloc(NS);
    dC.alloc(NG);
    gflag.alloc(NS);
    DevBuf<int32_t> d_slots;
    d_slots.alloc(B);
    DevBuf<double> d_inv;
    d_inv.alloc(B);
    // residual norms of the ring (host, once)
    std::vector<double> rnorm(L.cap, 0.0);
    std::vector<int64_t> usable;
    {
        std::vector<double> h((size_t)n);
        const int64_t oldest = (L.head - L.count + L.cap) % L.cap;
        HIP_CHECK(hipStreamSynchronize(s));
        for (int64_t k = 0; k < L.count; ++k) {
            const int64_t slot = (oldest + k) % L.cap;
            HIP_CHECK(hipMemcpy(h.data(), L.d_res.get() + slot * n, sizeof(double) * n, hipMemcpyDeviceToHost));
            double q = 0.0;
            for (double v : h) q += v * v;
            rnorm[slot] = std::sqrt(q);
            if (rnorm[slot] > 0.0 && std::isfinite(rnorm[slot])) usable.push_back(slot);
        }
    }
    if (usable.empty()) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "all harvested residuals are zero");
    const unsigned gn = (unsigned)((n + 255) / 256);
    const unsigned gBn = (unsigned)((B * n + 255) / 256);
    const unsigned gG = (unsigned)((NG + 255) / 256);
    const unsigned nwg = (unsigned)((NS + MLP_TS - 1) / MLP_TS);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<int32_t> slots(B);
    std::vector<double> inv(B), haq(2 * (size_t)B), hl(B);
    double last = 0.0;
    uint64_t cursor = 0;
    for (int step = 0; step <= steps; ++step) {
        // the last pass only evaluates, on the first batch's residuals: the reported losses before and
        // after training are on the same samples
        if (step == steps) cursor = 0;
        for (int b = 0; b < B; ++b) {
            const int64_t slot = usable[(cursor + (uint64_t)b * 7919u) % usable.size()];
            slots[b] = (int32_t)slot;
            inv[b] = 1.0 / rnorm[slot];
        }
        cursor += B;
        d_slots.upload(slots, s);
        d_inv.upload(inv, s);
        k_lc_select<<<dim3(gn), dim3(256), 0, s>>>(n, B, L.d_res.get(), d_slots.get(), d_inv.get(), r.get());
        for (int b = 0; b < B; ++b) {
            if (P->fp32)
                launch_spmv(P->P, P->d_vals32.get(), 0.0, zs.get() + (int64_t)b * n, 1.0, r.get() + (int64_t)b * n, s);
            else
                launch_spmv(P->P, P->d_vals.get(), 0.0, zs.get() + (int64_t)b * n, 1.0, r.get() + (int64_t)b * n, s);
        }
        if (L.kappa == 0.0f) {  // |r_b| = 1
            k_lc_norm2<<<dim3((unsigned)B), dim3(256), 0, s>>>(n, zs.get(), aq.get());
            HIP_CHECK(hipMemcpyAsync(hl.data(), aq.get(), sizeof(double) * B, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            double k = 0.0;
            for (double q : hl) k += std::sqrt(q);
            L.kappa = (float)(k / B);
            if (!(L.kappa > 0.0f) || !std::isfinite(L.kappa)) throw Error(HOEFEM_ERROR_NOT_FINITE, "bad output scale");
        }
        // g = Phi^T r per aggregate, c = kappa |g| guard(MLP(g / |g|)), z = P r + Phi c
        k_lc_project<<<dim3((unsigned)L.na, (unsigned)B), dim3(256), 0, s>>>(
            L.na, L.d_agg_ptr.get(), L.d_agg_dofs.get(), L.d_basis.get(), r.get(), n, G.get(), gnorm.get(), 1);
        k_lc_project<<<dim3((unsigned)L.na, (unsigned)B), dim3(256), 0, s>>>(
            L.na, L.d_agg_ptr.get(), L.d_agg_dofs.get(), L.d_basis.get(), r.get(), n, Gf.get(), nullptr, 2);
        k_mlp_fwd<<<dim3(nwg), dim3(256), 0, s>>>(NS, G.get(), L.d_wbf.get(), L.d_params.get(), L.d_z1.get(),
                                                  L.d_h1.get(), L.d_z2.get(), L.d_h2.get(), L.d_y.get());
        launch_lc_guard(NS, Gf.get(), L.d_y.get(), gflag.get(), s);
        k_lc_prolong<<<dim3((unsigned)L.na, (unsigned)B), dim3(256), 0, s>>>(
            L.na, L.d_agg_ptr.get(), L.d_agg_dofs.get(), L.d_basis.get(), L.d_y.get(), gnorm.get(), L.kappa, n,
            y.get());
        k_lc_add<<<dim3(gBn), dim3(256), 0, s>>>(B * n, zs.get(), y.get());  // y <- z = zs + y
        for (int b = 0; b < B; ++b) launch_spmv(m.pattern, Jvals, 0.0, Jz.get() + (int64_t)b * n, 1.0, y.get() + (int64_t)b * n, s);
        check_launch("learned training forward");
        k_lc_dots<<<dim3((unsigned)B), dim3(256), 0, s>>>(n, y.get(), Jz.get(), r.get(), aq.get());
        if (step == 0 || step == steps) {
            HIP_CHECK(hipMemcpyAsync(haq.data(), aq.get(), sizeof(double) * 2 * B, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            double tot = 0.0;
            for (int b = 0; b < B; ++b) tot += haq[2 * b + 1] > 0.0 ? -haq[2 * b] * haq[2 * b] / haq[2 * b + 1] : 0.0;
            tot /= B;
            if (step == 0 && initial_loss) *initial_loss = tot;
            last = tot;
        }
        if (step == steps) break;
        k_lc_graddz<<<dim3(gBn), dim3(256), 0, s>>>(n, B, aq.get(), Jz.get(), r.get(), dz.get());
        k_lc_project<<<dim3((unsigned)L.na, (unsigned)B), dim3(256), 0, s>>>(
            L.na, L.d_agg_ptr.get(), L.d_agg_dofs.get(), L.d_basis.get(), dz.get(), n, dC.get(), nullptr, 0);
        k_lc_out_grad<<<dim3(gG), dim3(256), 0, s>>>(NG, gnorm.get(), dC.get(), L.kappa, L.d_dy.get());
        k_lc_guard_grad<<<dim3((unsigned)((NS + 3) / 4)), dim3(256), 0, s>>>(NS, Gf.get(), gflag.get(), L.d_dy.get());
        k_mlp_bwd<<<dim3(nwg), dim3(256), 0, s>>>(NS, G.get(), L.d_wbf.get(), L.d_z1.get(), L.d_h1.get(),
                                                  L.d_z2.get(), L.d_h2.get(), L.d_dy.get(), L.d_gpart.get());
        k_mlp_reduce<<<dim3((LC_PARAMS + 255) / 256), dim3(256), 0, s>>>((int)nwg, L.d_gpart.get(), 1.0f / (float)B,
                                                                        L.d_grad.get());
        L.adam_t += 1;
        const float c1 = 1.0f - std::pow(b1, (float)L.adam_t), c2 = 1.0f - std::pow(b2, (float)L.adam_t);
        k_adam<<<dim3((LC_PARAMS + 255) / 256), dim3(256), 0, s>>>(L.d_grad.get(), L.d_params.get(), L.d_m.get(),
                                                                   L.d_v.get(), L.d_wbf.get(), (float)lr, b1, b2, eps,
                                                                   c1, c2);
        check_launch("learned training step");
    }
    HIP_CHECK(hipStreamSynchronize(s));
    L.trained = true;
    return last;
}

double learned_train(Learned& L, const double* J, Schwarz* P, int steps, double lr, int B, double* initial_loss) {
    if (!P || P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "training needs the FP64 Schwarz operator here");
    return learned_train_t<double>(L, J, P, steps, lr, B, initial_loss);
}

double learned_train(Learned& L, const float* J, Schwarz* P, int steps, double lr, int B, double* initial_loss) {
    if (!P || !P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "training needs the FP32 Schwarz operator here");
    return learned_train_t<float>(L, J, P, steps, lr, B, initial_loss);
}

LearnedView learned_view(Learned& L) {
    L.d_cg_x.ensure((size_t)L.na * LC_IN);
    L.d_cg_xf.ensure((size_t)L.na * LC_IN);
    L.d_cg_gn.ensure((size_t)L.na);
    L.d_cg_y.ensure((size_t)L.na * LC_OUT);
    return LearnedView{L.na,           L.kappa,         L.d_agg_ptr.get(), L.d_agg_dofs.get(), L.d_basis.get(),
                       L.d_wbf.get(),  L.d_params.get(), L.d_cg_x.get(),   L.d_cg_xf.get(),    L.d_cg_gn.get(),
                       L.d_cg_y.get()};
}

}  // namespace hoefem
