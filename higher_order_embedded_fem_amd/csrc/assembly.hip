// Assembly kernels for gfx950: element forces, element tangents and element mass into
// staging buffers, then color-ordered, atomic-free gathers into node-owned CSR row blocks.
//
// Summation order matches the reference's colored assembly (fenris/src/assembly.rs:253-313,
// :915-966): every CSR entry / force component receives its element contributions in color
// order, because the incidence lists are built color by color and one work-group owns each
// node's row block.
#include <type_traits>
#include <algorithm>

#include "fem_device.hpp"
#include "internal.hpp"

namespace hoefem {

using namespace dev;

// ------------------------------------------------------------------ element forces
// assemble_generalized_element_elliptic_term (fenris/src/assembly.rs:674-711) with
// g = -P(F)^T (fenris/src/solid/assembly.rs:33-38). One 64-lane work-group per element; quadrature
// points in chunks of FC_PTS:
//  - phase 1, FC_SUB lanes per point: the Jacobian (each lane), the node sum of F = I + u_grad split over
//    the point's lanes and reduced by shuffles, then the stress; the point's first lane stores
//    A = -P J^-T, w and |det J| (and the point's position for phase 2) in LDS;
//  - phase 2, lanes = element nodes: f_I += A G_ref(I) w |det J| over the chunk's points in order.
// The first chunk's point data is loaded with the element's nodes, not at the start of phase 1.
constexpr int FC_PTS = 8, FC_SUB = 64 / FC_PTS;

template <int ET, int MAT>
__global__ __launch_bounds__(64) void k_element_forces(int64_t ne, const int32_t* __restrict__ conn,
                                                       const double* __restrict__ verts,
                                                       const int32_t* __restrict__ qoff,
                                                       const double* __restrict__ qw,
                                                       const double* __restrict__ qxi, double thr,
                                                       double mu, double lambda, const double* __restrict__ u,
                                                       double* __restrict__ fe, int* err) {
    constexpr int N = ElemTraits<ET>::N, GN = ElemTraits<ET>::GN;
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x, pq = t / FC_SUB, sub = t % FC_SUB;
    __shared__ int s_node[N];
    __shared__ double s_u[3 * N];
    __shared__ double s_X[3 * GN];
    __shared__ double s_A[FC_PTS][12];
    __shared__ double s_xi[FC_PTS][3];
    const int q0 = qoff[e], q1 = qoff[e + 1];
    if (t < N) s_node[t] = conn[(int64_t)t * ne + e];
    __syncthreads();
    for (int k = t; k < 3 * N; k += 64) s_u[k] = u[3 * (int64_t)s_node[k / 3] + k % 3];
    for (int k = t; k < 3 * GN; k += 64) s_X[k] = verts[3 * (int64_t)s_node[k / 3] + k % 3];
    double pxi[3] = {0.0, 0.0, 0.0}, pw = 0.0;  // the lane's point of the first chunk
    if (q1 > q0) {
        const int q = min(q0 + pq, q1 - 1);
        pxi[0] = qxi[3 * q];
        pxi[1] = qxi[3 * q + 1];
        pxi[2] = qxi[3 * q + 2];
        pw = qw[q];
    }
    __syncthreads();
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int qb = q0; qb < q1; qb += FC_PTS) {
        const int q = qb + pq;
        const bool act = q < q1;
        double xi[3], wq = 0.0;
        if (qb == q0) {
            xi[0] = pxi[0];
            xi[1] = pxi[1];
            xi[2] = pxi[2];
            wq = pw;
        } else {
            const int qc = act ? q : q1 - 1;
            xi[0] = qxi[3 * qc];
            xi[1] = qxi[3 * qc + 1];
            xi[2] = qxi[3 * qc + 2];
            wq = qw[qc];
        }
        double J[9], Jinv[9], psi[4];
        ref_jacobian<ET>(s_X, xi, J);
        const double detJ = det3(J);
        bool ok = inv3(J, Jinv);
        tet4_psi(xi, psi);
        double F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (act && ok) {
            // mm(b, a) = sum_I u_I[b] * (G_ref^T J^-1)(I, a), nodes split over the point's lanes
            for (int I = sub; I < N; I += FC_SUB) {
                double G[3];
                grad_node<ET>(I, xi, psi, G);
                double tmp[3];
                for (int a = 0; a < 3; ++a) tmp[a] = G[0] * Jinv[0 + 3 * a] + G[1] * Jinv[1 + 3 * a] + G[2] * Jinv[2 + 3 * a];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) F[b + 3 * a] += s_u[3 * I + b] * tmp[a];
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k)
#pragma unroll
            for (int off = 1; off < FC_SUB; off <<= 1) F[k] += __shfl_xor(F[k], off, 64);
        if (act && sub == 0) {
            s_xi[pq][0] = xi[0];
            s_xi[pq][1] = xi[1];
            s_xi[pq][2] = xi[2];
            double P[9];
            if (ok) {
                F[0] += 1.0;
                F[4] += 1.0;
                F[8] += 1.0;
                ok = material_stress(MAT, thr, mu, lambda, F, P);
            }
            if (ok) {
                // A = (-P) * J^-T  : A(i, a) = sum_c -P(i, c) * Jinv(a, c)
                for (int a = 0; a < 3; ++a)
                    for (int i = 0; i < 3; ++i)
                        s_A[pq][i + 3 * a] = (-P[i]) * Jinv[a] + (-P[i + 3]) * Jinv[a + 3] + (-P[i + 6]) * Jinv[a + 6];
                s_A[pq][9] = wq;
                s_A[pq][10] = fabs(detJ);
            } else {
                atomicOr(err, KERR_SINGULAR);
                for (int k = 0; k < 11; ++k) s_A[pq][k] = 0.0;
            }
        }
        __syncthreads();
        if (t < N) {
            const int cnt = min(FC_PTS, q1 - qb);
            for (int k = 0; k < cnt; ++k) {
                const double xk[3] = {s_xi[k][0], s_xi[k][1], s_xi[k][2]};
                double pk[4], G[3];
                tet4_psi(xk, pk);
                grad_node<ET>(t, xk, pk, G);
                const double* A = s_A[k];
                const double w = A[9], dj = A[10];
                f0 += (A[0] * G[0] + A[3] * G[1] + A[6] * G[2]) * w * dj;
                f1 += (A[1] * G[0] + A[4] * G[1] + A[7] * G[2]) * w * dj;
                f2 += (A[2] * G[0] + A[5] * G[1] + A[8] * G[2]) * w * dj;
            }
        }
        __syncthreads();
    }
    if (t < N) {
        double* out = fe + e * 3 * N + 3 * t;
        out[0] = f0;
        out[1] = f1;
        out[2] = f2;
    }
}

// ------------------------------------------------------------------ element tangents
// assemble_generalized_element_stiffness (fenris/src/assembly.rs:747-797) with the SNH
// contraction (fenris/src/solid/materials.rs:349-376, :690-743). One 256-lane work-group per
// element; lanes = quadrature points build dP/dF and the scaled gradients in LDS, then every
// lane accumulates a fixed set of K_e entries (lower block triangle, mirrored on write).
// Output: row-major (3N)^2 block per element in the staging buffer.
constexpr int STIFF_THREADS = 256;
constexpr int STIFF_QC = 16;

template <int ET, class T, int MAT>
__global__ __launch_bounds__(STIFF_THREADS) void k_element_stiffness(
    int64_t ne, const int32_t* __restrict__ conn, const double* __restrict__ verts, const int32_t* __restrict__ qoff,
    const double* __restrict__ qw, const double* __restrict__ qxi, T thr, T mu, T lambda,
    const double* __restrict__ u, T* __restrict__ ke, int* err) {
    constexpr int N = ElemTraits<ET>::N, GN = ElemTraits<ET>::GN, M3 = 3 * N;
    constexpr int PAIRS = N * (N + 1) / 2;
    constexpr int ENTRIES = 9 * PAIRS;
    constexpr int PER = (ENTRIES + STIFF_THREADS - 1) / STIFF_THREADS;
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x;
    __shared__ int s_node[N];
    __shared__ T s_u[3 * N];
    __shared__ T s_X[3 * GN];
    __shared__ T s_C[STIFF_QC][81];
    __shared__ T s_a[STIFF_QC][M3];
    __shared__ unsigned char s_pI[PAIRS], s_pJ[PAIRS];
    if (t < N) s_node[t] = conn[(int64_t)t * ne + e];
    if (t == 0) {
        int p = 0;
        for (int J = 0; J < N; ++J)
            for (int I = J; I < N; ++I) {
                s_pI[p] = (unsigned char)I;
                s_pJ[p] = (unsigned char)J;
                ++p;
            }
    }
    __syncthreads();
    for (int k = t; k < 3 * N; k += STIFF_THREADS) s_u[k] = T(u[3 * (int64_t)s_node[k / 3] + k % 3]);
    for (int k = t; k < 3 * GN; k += STIFF_THREADS) s_X[k] = T(verts[3 * (int64_t)s_node[k / 3] + k % 3]);
    __syncthreads();
    T acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = T(0);
    const int q0 = qoff[e], q1 = qoff[e + 1];
    for (int qb = q0; qb < q1; qb += STIFF_QC) {
        const int cnt = min(STIFF_QC, q1 - qb);
        if (t < cnt) {
            const int q = qb + t;
            const T xi[3] = {T(qxi[3 * q]), T(qxi[3 * q + 1]), T(qxi[3 * q + 2])};
            T J[9], Jinv[9];
            ref_jacobian<ET>(s_X, xi, J);
            const T detJ = det3(J);
            bool ok = inv3(J, Jinv);
            T psi[4];
            tet4_psi(xi, psi);
            T ug[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};  // u_grad(a, b) = sum_I G(a, I) u(b, I)
            if (ok) {
                for (int I = 0; I < N; ++I) {
                    T Gr[3];
                    grad_node<ET>(I, xi, psi, Gr);
                    T G[3];
                    // G = J^-T G_ref : G(a) = sum_c Jinv(c, a) G_ref(c)
                    for (int a = 0; a < 3; ++a) G[a] = Jinv[0 + 3 * a] * Gr[0] + Jinv[1 + 3 * a] * Gr[1] + Jinv[2 + 3 * a] * Gr[2];
                    for (int b = 0; b < 3; ++b)
                        for (int a = 0; a < 3; ++a) ug[a + 3 * b] += G[a] * s_u[3 * I + b];
                    for (int a = 0; a < 3; ++a) s_a[t][3 * I + a] = G[a];
                }
                const T sq = sqrt(T(qw[q]) * fabs(detJ));
                for (int k = 0; k < M3; ++k) s_a[t][k] *= sq;
                T F[9];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) F[b + 3 * a] = ug[a + 3 * b];
                F[0] += T(1);
                F[4] += T(1);
                F[8] += T(1);
                ok = material_dpdf(MAT, thr, mu, lambda, F, s_C[t]);
            }
            if (!ok) {
                atomicOr(err, KERR_SINGULAR);
                for (int k = 0; k < 81; ++k) s_C[t][k] = T(0);
                for (int k = 0; k < M3; ++k) s_a[t][k] = T(0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int ent = t + STIFF_THREADS * j;
            if (ent < ENTRIES) {
                const int p = ent / 9, ij = ent % 9;
                const int i = ij % 3, jj = ij / 3;
                const int I = s_pI[p], Jn = s_pJ[p];
                for (int k2 = 0; k2 < cnt; ++k2) {
                    const T* C = s_C[k2];
                    const T* aI = s_a[k2] + 3 * I;
                    const T* aJ = s_a[k2] + 3 * Jn;
                    T res = T(0);
                    for (int l = 0; l < 3; ++l)
                        for (int k = 0; k < 3; ++k) res += C[(3 * k + i) + 9 * (3 * l + jj)] * aI[k] * aJ[l];
                    acc[j] += res;
                }
            }
        }
        __syncthreads();
    }
    T* out = ke + e * (int64_t)M3 * M3;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int ent = t + STIFF_THREADS * j;
        if (ent < ENTRIES) {
            const int p = ent / 9, ij = ent % 9;
            const int i = ij % 3, jj = ij / 3;
            const int I = s_pI[p], Jn = s_pJ[p];
            out[(3 * I + i) * M3 + 3 * Jn + jj] = acc[j];
            if (I != Jn) out[(3 * Jn + jj) * M3 + 3 * I + i] = acc[j];
        }
    }
}

// FP64 element tangents on the matrix cores. With G~_q = sqrt(w_q |det J_q|) G_q (3 x N spatial
// gradients) and C_q = dP/dF (column-major 9 x 9, entry (3k + i, 3l + j)), the block
//     K[(I, i), (J, j)] = sum_q sum_{k, l} G~_q(k, I) C_q(3k + i, 3l + j) G~_q(l, J)
// is, for each component pair (i, j), the product K^{ij} = A B with A(I, kk) = G~_q(k, I) and
// B(kk, J) = U^{ij}_q(k, J) = sum_l C_q(3k + i, 3l + j) G~_q(l, J) over kk = 3q + k. The six pairs
// i <= j suffice (C has the major symmetry, so K^{ji} = (K^{ij})^T). Nodes are padded to 16 or 32,
// so a pair is one or four 16 x 16 tiles, accumulated in registers across the quadrature chunks.
// Lane layout of v_mfma_f64_16x16x4 (lane l): A operand A(l % 16, l / 16), B operand
// B(l / 16, l % 16), result item v = D(4 v + l / 16, l % 16). Each entry of K_e is written exactly
// once (diagonal pairs from their lower triangle), so K_e is exactly symmetric.
typedef double mf_double4 __attribute__((ext_vector_type(4)));

// One element per 128-lane work-group (two wavefronts; many work-groups per CU, so the SIMDs
// interleave the phases of different elements). Quadrature points go in chunks of TW_QC:
//  - per-point phase with TW_SUB lanes per point: the Jacobian (each lane), the node gradients split
//    over the point's lanes (scaled by sqrt(w |det J|) into s_G), u_grad reduced over the lanes by
//    shuffles, then dP/dF as nine 3 x 3 blocks on nine of the point's lanes (snh_dpdf_block; the
//    material variants evaluate dP/dF on the point's first lane). No memory access waits in this
//    phase: the first chunk's point data is loaded with the element's nodes, and the reference node
//    coordinates are decoded from constants (hexn);
//  - product phase: K^{ij} += G~^T U^{ij} on v_mfma_f64_16x16x4, wavefront w taking the pairs
//    (i, j) = (w, w..2) for w = 0 and (1, 1), (1, 2), (2, 2) for w = 1, with the B operand
//    U^{ij}(kk, col) = sum_m C_q(3k + i, 3m + j) G~(3q + m, col) formed in registers at each k-step
//    (no U staging in LDS); a diagonal pair skips its tile above the diagonal.
constexpr int TW_QC = 8, TW_THREADS = 128, TW_SUB = TW_THREADS / TW_QC;

template <int ET, int MAT>
__global__ __launch_bounds__(TW_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_element_stiffness_tw(
    int64_t ne, const int32_t* __restrict__ conn, const double* __restrict__ verts, const int32_t* __restrict__ qoff,
    const double* __restrict__ qw, const double* __restrict__ qxi, double thr, double mu, double lambda,
    const double* __restrict__ u, double* __restrict__ ke, int* err) {
    constexpr int N = ElemTraits<ET>::N, GN = ElemTraits<ET>::GN, M3 = 3 * N;
    constexpr int NP = N > 16 ? 32 : 16, NT = NP / 16, JW = 3 * NT * NT;  // jobs per wavefront
    constexpr int KS = (3 * TW_QC) / 4;                                    // k-steps of a full chunk
    static_assert(N <= 32 && (3 * TW_QC) % 4 == 0, "tile layout");
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    __shared__ int s_node[N];
    __shared__ double s_u[3 * N];
    __shared__ double s_X[3 * GN];
    __shared__ double s_C[TW_QC][81];
    __shared__ double s_G[3 * TW_QC * NP];
    const int pq = t / TW_SUB, sub = t % TW_SUB;
    const int q0 = qoff[e], q1 = qoff[e + 1];
    if (t < N) s_node[t] = conn[(int64_t)t * ne + e];
    __syncthreads();
    for (int k = t; k < 3 * N; k += TW_THREADS) s_u[k] = u[3 * (int64_t)s_node[k / 3] + k % 3];
    for (int k = t; k < 3 * GN; k += TW_THREADS) s_X[k] = verts[3 * (int64_t)s_node[k / 3] + k % 3];
    // the lane's point of the first chunk (rule position and weight), loaded with the element's data
    // rather than at the start of the per-point phase
    double pxi[3] = {0.0, 0.0, 0.0}, pw = 0.0;
    if (q1 > q0) {
        const int q = min(q0 + pq, q1 - 1);
        pxi[0] = qxi[3 * q];
        pxi[1] = qxi[3 * q + 1];
        pxi[2] = qxi[3 * q + 2];
        pw = qw[q];
    }
    mf_double4 acc[JW];
    // One chunk of up to TW_QC points. The first chunk is its own instance: its products start from
    // zero, so the accumulators hold nothing during its per-point phase (fewer live registers there).
    auto chunk = [&](int qc, auto first_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;
        const int cnt = min(TW_QC, q1 - qc);
        for (int k = t; k < 3 * TW_QC * NP; k += TW_THREADS) s_G[k] = 0.0;
        __syncthreads();
        // ---- per-point phase: TW_SUB lanes per quadrature point
        const bool act = pq < cnt;
        bool ok = act;
        double ug[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // u_grad(a, b) = sum_I G(a, I) u(b, I)
        if (act) {
            const int q = qc + pq;
            double xi[3], wq;
            if constexpr (FIRST) {
                xi[0] = pxi[0];
                xi[1] = pxi[1];
                xi[2] = pxi[2];
                wq = pw;
            } else {
                xi[0] = qxi[3 * q];
                xi[1] = qxi[3 * q + 1];
                xi[2] = qxi[3 * q + 2];
                wq = qw[q];
            }
            double J[9], Jinv[9], psi[4];
            ref_jacobian<ET>(s_X, xi, J);
            const double detJ = det3(J);
            ok = inv3(J, Jinv);
            tet4_psi(xi, psi);
            if (ok) {
                const double sq = sqrt(wq * fabs(detJ));
                for (int I = sub; I < N; I += TW_SUB) {
                    double Gr[3], G[3];
                    grad_node<ET>(I, xi, psi, Gr);
                    for (int a = 0; a < 3; ++a) G[a] = Jinv[0 + 3 * a] * Gr[0] + Jinv[1 + 3 * a] * Gr[1] + Jinv[2 + 3 * a] * Gr[2];
                    for (int b = 0; b < 3; ++b)
                        for (int a = 0; a < 3; ++a) ug[a + 3 * b] += G[a] * s_u[3 * I + b];
                    for (int a = 0; a < 3; ++a) s_G[(3 * pq + a) * NP + I] = G[a] * sq;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k)
#pragma unroll
            for (int off = 1; off < TW_SUB; off <<= 1) ug[k] += __shfl_xor(ug[k], off, 64);
        double* C = s_C[pq];
        if (ok) {
            double F[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) F[b + 3 * a] = ug[a + 3 * b];
            F[0] += 1.0;
            F[4] += 1.0;
            F[8] += 1.0;
            if constexpr (MAT == MAT_SNH) {
                const double alpha = 1.0 + mu / lambda * 0.75;
                double I_C = 0.0;
                for (int k = 0; k < 9; ++k) I_C += F[k] * F[k];
                const double Jf = det3(F);
                if (Jf == 0.0) {
                    ok = false;
                } else {
                    double g[9];
                    cof3(F, g);
                    const double beta = I_C + 1.0;
                    const double diag = mu * (1.0 - 1.0 / beta), cf = mu * 2.0 / (beta * beta);
                    const double scale = lambda * (Jf - alpha);
                    // the 81 entries of dP/dF as nine 3 x 3 blocks, one per lane of the point
                    if (sub < 9) snh_dpdf_block(sub % 3, sub / 3, F, g, diag, cf, lambda, scale, C);
                }
            } else {
                int okl = 1;
                if (sub == 0) okl = material_dpdf(MAT, thr, mu, lambda, F, C) ? 1 : 0;
                okl = __shfl(okl, l - (l % TW_SUB), 64);  // a point's lanes lie in one wavefront
                ok = okl != 0;
            }
        }
        if (act && !ok) {  // singular element Jacobian or deformation gradient: the point adds nothing
            if (sub == 0) atomicOr(err, KERR_SINGULAR);
            for (int idx = sub; idx < 81; idx += TW_SUB) C[idx] = 0.0;
            for (int k = sub; k < 3 * NP; k += TW_SUB) s_G[3 * pq * NP + k] = 0.0;
        }
        __syncthreads();
        // ---- product phase: rows kk of the chunk (3 per point); rows >= 3 cnt are zero in s_G
        const int ksteps = (3 * cnt + 3) / 4;
        double a[NT][KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int ti = 0; ti < NT; ++ti) a[ti][ks] = s_G[(4 * ks + (l >> 4)) * NP + 16 * ti + (l & 15)];
#pragma unroll
        for (int pp = 0; pp < 3; ++pp) {
            const int p = 3 * w + pp;  // pairs (0,0) (0,1) (0,2) | (1,1) (1,2) (2,2)
            const int i = p < 3 ? 0 : (p < 5 ? 1 : 2);
            const int j = p < 3 ? p : (p < 5 ? p - 2 : 2);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                if (ks == 0 || ks < ksteps) {  // a chunk has at least one point: k-step 0 always runs
                    const int kk = 4 * ks + (l >> 4), qq = kk / 3, k = kk % 3;
                    const bool live = qq < cnt;
                    const double* Cq = s_C[live ? qq : 0];
                    const double c0 = Cq[(3 * k + i) + 9 * j], c1 = Cq[(3 * k + i) + 9 * (3 + j)],
                                 c2 = Cq[(3 * k + i) + 9 * (6 + j)];
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        const int col = 16 * tj + (l & 15);
                        double b = 0.0;
                        if (live)
                            b = c0 * s_G[(3 * qq) * NP + col] + c1 * s_G[(3 * qq + 1) * NP + col] +
                                c2 * s_G[(3 * qq + 2) * NP + col];
#pragma unroll
                        for (int ti = 0; ti < NT; ++ti) {
                            if (i == j && ti < tj) continue;  // a diagonal pair's tile above its diagonal is never stored
                            mf_double4& A = acc[(pp * NT + ti) * NT + tj];
                            A = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a[ti][ks], b, (FIRST && ks == 0) ? mf_double4{0.0, 0.0, 0.0, 0.0} : A, 0, 0, 0);
                        }
                    }
                }
            }
        }
        __syncthreads();
    };
    if (q1 > q0) {
        chunk(q0, std::true_type{});
        for (int qc = q0 + TW_QC; qc < q1; qc += TW_QC) chunk(qc, std::false_type{});
    } else {
#pragma unroll
        for (int j = 0; j < JW; ++j) acc[j] = mf_double4{0.0, 0.0, 0.0, 0.0};
    }
    double* out = ke + e * (int64_t)M3 * M3;
#pragma unroll
    for (int jb = 0; jb < JW; ++jb) {
        const int pp = jb / (NT * NT), tile = jb % (NT * NT);
        const int ti = tile / NT, tj = tile % NT;
        const int p = 3 * w + pp;
        const int i = p < 3 ? 0 : (p < 5 ? 1 : 2);
        const int j = p < 3 ? p : (p < 5 ? p - 2 : 2);
        if (i == j && ti < tj) continue;  // not computed: every entry of it lies above the diagonal
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int I = 16 * ti + (4 * v + (l >> 4)), Jn = 16 * tj + (l & 15);
            if (I < N && Jn < N && (i != j || I >= Jn)) {
                out[(3 * I + i) * M3 + 3 * Jn + j] = acc[jb][v];
                out[(3 * Jn + j) * M3 + 3 * I + i] = acc[jb][v];
            }
        }
    }
}

// ------------------------------------------------------------------ element mass
// assemble_generalized_element_mass (fenris/src/assembly.rs:624-671): the scalar matrix
// rho * sum_q |det J| w phi_i phi_j duplicated on the diagonal of each 3x3 block. With a
// regularization factor > 0 the interior rule is added with density * factor
// (fenris/src/embedding/mod.rs:438-451).
template <int ET>
__global__ __launch_bounds__(256) void k_element_mass(int64_t ne, const int32_t* __restrict__ conn,
                                                      const double* __restrict__ verts,
                                                      const int32_t* __restrict__ qoff, const double* __restrict__ qw,
                                                      const double* __restrict__ qxi, double density,
                                                      int nreg, const double* __restrict__ rw,
                                                      const double* __restrict__ rxi, double reg_density,
                                                      double* __restrict__ ke) {
    constexpr int N = ElemTraits<ET>::N, GN = ElemTraits<ET>::GN, M3 = 3 * N;
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x;
    __shared__ int s_node[N];
    __shared__ double s_X[3 * GN];
    if (t < N) s_node[t] = conn[(int64_t)t * ne + e];
    __syncthreads();
    for (int k = t; k < 3 * GN; k += 256) s_X[k] = verts[3 * (int64_t)s_node[k / 3] + k % 3];
    __syncthreads();
    double* out = ke + e * (int64_t)M3 * M3;
    for (int ent = t; ent < N * N; ent += 256) {
        const int i = ent / N, j = ent % N;
        double m = 0.0;
        for (int q = qoff[e]; q < qoff[e + 1]; ++q) {
            const double xi[3] = {qxi[3 * q], qxi[3 * q + 1], qxi[3 * q + 2]};
            double J[9], psi[4];
            ref_jacobian<ET>(s_X, xi, J);
            tet4_psi(xi, psi);
            m += density * fabs(det3(J)) * qw[q] * basis_node<ET>(i, xi, psi) * basis_node<ET>(j, xi, psi);
        }
        for (int q = 0; q < nreg; ++q) {
            const double xi[3] = {rxi[3 * q], rxi[3 * q + 1], rxi[3 * q + 2]};
            double J[9], psi[4];
            ref_jacobian<ET>(s_X, xi, J);
            tet4_psi(xi, psi);
            m += reg_density * fabs(det3(J)) * rw[q] * basis_node<ET>(i, xi, psi) * basis_node<ET>(j, xi, psi);
        }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[(3 * i + a) * M3 + 3 * j + b] = (a == b) ? m : 0.0;
    }
}

// ------------------------------------------------------------------ block gather
// One 64-lane work-group owns the row block (3 rows) of pattern node v and accumulates the
// incident element blocks in color order in LDS (add_element_row_to_csr_row semantics,
// fenris/src/assembly.rs:1042-1075, without the linear search: slots are precomputed).
// mode 0: values = (values or 0) + sum; mode 1: values = (0 + c * sum) + 1 * M (J combination,
// simulation_toolbox/src/fem/system_assembly.rs:196-235).
// A row block wider than `chunk` neighbour nodes (large Schwarz subdomains) is done in column
// chunks of `chunk` nodes, each a pass over the incident blocks; every entry still receives its
// terms in color order, so the sums do not depend on the chunking.
constexpr int GATHER_CHUNK = 896;  // 9 * 896 doubles = 63 KiB of LDS
// threads of k_block_gather: four waves share a node row, so a row of hundreds of blocks (the
// Schwarz operator of large subdomains) is not walked by one wave
constexpr int GB_THREADS = 256;
template <class TB, class TV>
__global__ __launch_bounds__(GB_THREADS) void k_block_gather(
    const int32_t* __restrict__ nptr, const int32_t* __restrict__ rowptr, const uint8_t* __restrict__ diag_only,
    const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_blk, const int32_t* __restrict__ inc_loc,
    const int32_t* __restrict__ msz, const int64_t* __restrict__ val_off, const int64_t* __restrict__ slot_off,
    const int32_t* __restrict__ slot, const TB* __restrict__ blocks, TV* __restrict__ values, int mode,
    double c, const double* __restrict__ M, int zero_first, const double* __restrict__ D2, double cd, double cm,
    int chunk) {
    extern __shared__ double s_rb[];
    const int64_t v = blockIdx.x;
    const int t = threadIdx.x;
    if (diag_only[v]) return;
    const int k = nptr[v + 1] - nptr[v];
    if (k == 0) return;
    const int rw = 3 * k;
    const int64_t r0 = rowptr[3 * v];
    // rows 3v, 3v+1, 3v+2 are contiguous: row i starts at r0 + i * rw
    for (int c0 = 0; c0 < k; c0 += chunk) {
        const int cw = min(chunk, k - c0);
        const int cr = 3 * cw;  // a row's width within the chunk
        const int R = 3 * cr;
        for (int idx = t; idx < R; idx += GB_THREADS) {
            const int i = idx / cr, cc = idx % cr;
            s_rb[idx] = (zero_first || mode == 1) ? 0.0 : (double)values[r0 + i * rw + 3 * c0 + cc];
        }
        __syncthreads();
        for (int it = inc_ptr[v]; it < inc_ptr[v + 1]; ++it) {
            const int b = inc_blk[it], a = inc_loc[it];
            const int mb = msz[b];
            const int bw = 3 * mb;
            const TB* base = blocks + val_off[b] + (int64_t)(3 * a) * bw;
            const int32_t* sl = slot + slot_off[b] + (int64_t)a * mb;
            for (int idx = t; idx < 3 * bw; idx += GB_THREADS) {
                const int i = idx / bw, cc = idx % bw;
                const int bb = cc / 3, jj = cc % 3;
                const int col = sl[bb] - c0;
                if (col >= 0 && col < cw) s_rb[i * cr + 3 * col + jj] += (double)base[idx];
            }
            __syncthreads();
        }
        for (int idx = t; idx < R; idx += GB_THREADS) {
            const int i = idx / cr, cc = idx % cr;
            const int64_t o = r0 + i * rw + 3 * c0 + cc;
            const double s = s_rb[idx];
            if (mode == 0)
                values[o] = (TV)s;
            else {
                // compute_jacobian_combination_into (system_assembly.rs:197-235): the terms in the order
                // stiffness, damping, mass
                double v2 = 0.0;
                v2 += c * s;
                if (D2) v2 += cd * D2[o];
                if (M) v2 += cm * M[o];
                values[o] = (TV)v2;
            }
        }
        __syncthreads();  // the next chunk reuses s_rb
    }
}

// The same gather with the incident blocks' metadata held one incidence per lane (moved by
// shuffles instead of loaded per block) and each block's row values and slots loaded one block
// ahead, so a node row's blocks cost one memory latency between them instead of three. For blocks of
// at most 28 nodes (3 rows x 3m <= 256 values, four per lane) and rows that fit the LDS at once;
// every entry receives its terms in the same order as in k_block_gather.
constexpr int GP_VALS = 4;
template <class TB, class TV>
__global__ __launch_bounds__(64) void k_block_gather_pf(
    const int32_t* __restrict__ nptr, const int32_t* __restrict__ rowptr, const uint8_t* __restrict__ diag_only,
    const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_blk, const int32_t* __restrict__ inc_loc,
    const int32_t* __restrict__ msz, const int64_t* __restrict__ val_off, const int64_t* __restrict__ slot_off,
    const int32_t* __restrict__ slot, const TB* __restrict__ blocks, TV* __restrict__ values, int mode,
    double c, const double* __restrict__ M, int zero_first, const double* __restrict__ D2, double cd, double cm) {
    extern __shared__ double s_rb[];
    const int64_t v = blockIdx.x;
    const int t = threadIdx.x;
    if (diag_only[v]) return;
    const int k = nptr[v + 1] - nptr[v];
    if (k == 0) return;
    const int rw = 3 * k, R = 3 * rw;
    const int64_t r0 = rowptr[3 * v];
    const int i0 = inc_ptr[v], ninc = inc_ptr[v + 1] - i0;
    // lane j holds incidence j (at most 64, checked by the launch)
    int a_l = 0, mb_l = 0;
    int64_t vo_l = 0, so_l = 0;
    if (t < ninc) {
        const int b = inc_blk[i0 + t];
        a_l = inc_loc[i0 + t];
        mb_l = msz[b];
        vo_l = val_off[b];
        so_l = slot_off[b];
    }
    for (int idx = t; idx < R; idx += 64) s_rb[idx] = (zero_first || mode == 1) ? 0.0 : (double)values[r0 + idx];
    // incidence j's row values and slots: every lane loads (a lane past the block's 3 x 3m values
    // re-reads its first one and does not use it), so the loads in flight are counted statically and
    // adding one incidence waits only for its own loads
    auto fetch = [&](int j, double* vv, int* ss, int& bwj) {
        const int a = __shfl(a_l, j, 64), mb = __shfl(mb_l, j, 64);
        const int64_t vo = __shfl(vo_l, j, 64), so = __shfl(so_l, j, 64);
        const int bw = 3 * mb;
        bwj = bw;
        const TB* base = blocks + vo + (int64_t)(3 * a) * bw;
        const int32_t* sl = slot + so + (int64_t)a * mb;
#pragma unroll
        for (int q = 0; q < GP_VALS; ++q) {
            const int idx = t + 64 * q;
            const int id = idx < 3 * bw ? idx : 0;
            vv[q] = (double)base[id];
            ss[q] = sl[(id % bw) / 3];
        }
    };
    // two incidence buffers used in turn (no register copy between them, which would wait for the
    // loads): the next incidence's loads are in flight while this one's values are added
    double av[GP_VALS] = {0.0, 0.0, 0.0, 0.0}, bv[GP_VALS] = {0.0, 0.0, 0.0, 0.0};
    int as[GP_VALS] = {0, 0, 0, 0}, bs[GP_VALS] = {0, 0, 0, 0};
    int abw = 0, bbw = 0;
    auto add = [&](const double* vv, const int* ss, int bw) {
#pragma unroll
        for (int q = 0; q < GP_VALS; ++q) {
            const int idx = t + 64 * q;
            if (idx < 3 * bw) {
                const int i = idx / bw, cc = idx % bw;
                s_rb[i * rw + 3 * ss[q] + cc % 3] += vv[q];
            }
        }
    };
    // The work-group is one wavefront, whose LDS accesses take effect in program order: the row's sums
    // need no s_barrier, only a compiler barrier that keeps the LDS accesses of consecutive incidences
    // in order (an s_barrier's fence would also wait for the loads of the incidence ahead). The fetches
    // are unconditional inside the block (the last incidence again past the end), so every path reaches
    // the loop with the same loads in flight and each addition waits only for its own incidence.
    if (ninc > 0) {
        fetch(0, av, as, abw);
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < ninc; j += 2) {
            fetch(min(j + 1, ninc - 1), bv, bs, bbw);
            add(av, as, abw);
            __builtin_amdgcn_wave_barrier();
            fetch(min(j + 2, ninc - 1), av, as, abw);
            if (j + 1 < ninc) add(bv, bs, bbw);
            __builtin_amdgcn_wave_barrier();
        }
    }
    for (int idx = t; idx < R; idx += 64) {
        const int64_t o = r0 + idx;
        const double sum = s_rb[idx];
        if (mode == 0)
            values[o] = (TV)sum;
        else {
            double v2 = 0.0;  // stiffness, damping, mass (system_assembly.rs:197-235)
            v2 += c * sum;
            if (D2) v2 += cd * D2[o];
            if (M) v2 += cm * M[o];
            values[o] = (TV)v2;
        }
    }
}

// f += sum of element force blocks in color order (one thread per node)
__global__ void k_force_gather(int64_t nn, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_blk,
                               const int32_t* __restrict__ inc_loc, int n, const double* __restrict__ fe,
                               double* __restrict__ f) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double a0 = f[3 * v], a1 = f[3 * v + 1], a2 = f[3 * v + 2];
    for (int it = inc_ptr[v]; it < inc_ptr[v + 1]; ++it) {
        const double* src = fe + ((int64_t)inc_blk[it] * n + inc_loc[it]) * 3;
        a0 += src[0];
        a1 += src[1];
        a2 += src[2];
    }
    f[3 * v] = a0;
    f[3 * v + 1] = a1;
    f[3 * v + 2] = a2;
}

// ------------------------------------------------------------------ SpMV
// y = beta*y + alpha*A x on a node-blocked CSR matrix: G lanes per node share the node's
// column list; each lane reads 3 contiguous values of each of the node's 3 rows.
template <int G, class VT>
__global__ __launch_bounds__(256) void k_spmv(int64_t nn, const int32_t* __restrict__ nptr,
                                              const int32_t* __restrict__ ncol, const int32_t* __restrict__ rowptr,
                                              const uint8_t* __restrict__ diag_only, const VT* __restrict__ A,
                                              double beta, double* __restrict__ y, double alpha,
                                              const double* __restrict__ x) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t v = gid / G;
    const int lane = (int)(gid % G);
    VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
    const bool valid = v < nn;
    if (valid) {
        if (diag_only[v]) {
            if (lane == 0) {
                a0 = A[rowptr[3 * v]] * VT(x[3 * v]);
                a1 = A[rowptr[3 * v + 1]] * VT(x[3 * v + 1]);
                a2 = A[rowptr[3 * v + 2]] * VT(x[3 * v + 2]);
            }
        } else {
            const int kb = nptr[v], k = nptr[v + 1] - kb;
            const VT* R0 = A + rowptr[3 * v];
            const VT* R1 = R0 + 3 * k;
            const VT* R2 = R1 + 3 * k;
            for (int j = lane; j < k; j += G) {
                const int64_t b = ncol[kb + j];
                const VT x0 = VT(x[3 * b]), x1 = VT(x[3 * b + 1]), x2 = VT(x[3 * b + 2]);
                a0 += R0[3 * j] * x0 + R0[3 * j + 1] * x1 + R0[3 * j + 2] * x2;
                a1 += R1[3 * j] * x0 + R1[3 * j + 1] * x1 + R1[3 * j + 2] * x2;
                a2 += R2[3 * j] * x0 + R2[3 * j + 1] * x1 + R2[3 * j + 2] * x2;
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, G);
        a1 += __shfl_xor(a1, off, G);
        a2 += __shfl_xor(a2, off, G);
    }
    if (valid && lane == 0) {
        if (beta == 0.0) {
            y[3 * v] = alpha * double(a0);
            y[3 * v + 1] = alpha * double(a1);
            y[3 * v + 2] = alpha * double(a2);
        } else {
            double y0 = beta * y[3 * v], y1 = beta * y[3 * v + 1], y2 = beta * y[3 * v + 2];
            y0 += alpha * double(a0);
            y1 += alpha * double(a1);
            y2 += alpha * double(a2);
            y[3 * v] = y0;
            y[3 * v + 1] = y1;
            y[3 * v + 2] = y2;
        }
    }
}

// ------------------------------------------------------------------ Dirichlet
// apply_homogeneous_dirichlet_bc_csr (fenris/src/assembly.rs:799-869): scale = |first non-zero
// diagonal entry|; Dirichlet rows and (by symmetry of the pattern) Dirichlet columns zeroed,
// Dirichlet diagonal := scale.
template <class VT>
__global__ __launch_bounds__(256) void k_first_nonzero_diag(int64_t ndof, const int32_t* __restrict__ diag_pos,
                                                            const VT* __restrict__ vals, double* scale) {
    __shared__ int64_t s_min[256];
    for (int64_t base = 0; base < ndof; base += 256) {
        const int64_t r = base + threadIdx.x;
        int64_t found = INT64_MAX;
        if (r < ndof && diag_pos[r] >= 0 && vals[diag_pos[r]] != VT(0)) found = r;
        s_min[threadIdx.x] = found;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) s_min[threadIdx.x] = min(s_min[threadIdx.x], s_min[threadIdx.x + s]);
            __syncthreads();
        }
        const int64_t m = s_min[0];
        __syncthreads();
        if (m != INT64_MAX) {
            if (threadIdx.x == 0) *scale = fabs((double)vals[diag_pos[m]]);
            return;
        }
    }
    if (threadIdx.x == 0) *scale = 1.0;
}

template <class VT>
__global__ void k_dirichlet_rows(const int32_t* __restrict__ bc_nodes, int64_t nbc, const int32_t* __restrict__ nptr,
                                 const int32_t* __restrict__ ncol, const int32_t* __restrict__ rowptr,
                                 const uint8_t* __restrict__ diag_only, const uint8_t* __restrict__ dirmask,
                                 VT* __restrict__ vals, const double* __restrict__ scale) {
    const int64_t idx = blockIdx.x;
    if (idx >= nbc) return;
    const int64_t v = bc_nodes[idx];
    const bool vdir = dirmask[v] != 0;
    const VT sc = (VT)*scale;
    if (diag_only[v]) {
        if (vdir && threadIdx.x < 3) vals[rowptr[3 * v + threadIdx.x]] = sc;
        return;
    }
    const int kb = nptr[v], k = nptr[v + 1] - kb;
    for (int ent = threadIdx.x; ent < 9 * k; ent += blockDim.x) {
        const int i = ent / (3 * k), cc = ent % (3 * k);
        const int j = cc / 3, jj = cc % 3;
        const int64_t b = ncol[kb + j];
        const int64_t pos = rowptr[3 * v + i] + cc;
        if (vdir)
            vals[pos] = (b == v && jj == i) ? sc : VT(0);
        else if (dirmask[b])
            vals[pos] = VT(0);
    }
}

// ------------------------------------------------------------------ material (batched)
__global__ void k_material(int mat, double thr, double mu, double lambda, int64_t nq, const double* __restrict__ F,
                           double* __restrict__ P, double* __restrict__ dP, int* err) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double f[9];
    for (int k = 0; k < 9; ++k) f[k] = F[9 * q + k];
    if (P) {
        double p[9];
        if (!material_stress(mat, thr, mu, lambda, f, p)) {
            atomicOr(err, KERR_SINGULAR);
            for (int k = 0; k < 9; ++k) p[k] = 0.0;
        }
        for (int k = 0; k < 9; ++k) P[9 * q + k] = p[k];
    }
    if (dP) {
        double d[81];
        if (!material_dpdf(mat, thr, mu, lambda, f, d)) {
            atomicOr(err, KERR_SINGULAR);
            for (int k = 0; k < 81; ++k) d[k] = 0.0;
        }
        for (int k = 0; k < 81; ++k) dP[81 * q + k] = d[k];
    }
}

// Stable Neo-Hookean energy density of nq deformation gradients (column-major F[9 q .. 9 q + 8])
__global__ void k_material_energy(double mu, double lambda, double mu_r, double lambda_r, int64_t nq,
                                  const double* __restrict__ F, double* __restrict__ psi) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double f[9];
    for (int k = 0; k < 9; ++k) f[k] = F[9 * q + k];
    psi[q] = snh_energy(mu, lambda, mu_r, lambda_r, f);
}

// rotation_svd (fenris/src/util.rs:68-94) on nq matrices: U, Vt column-major, S[3] signed.
__global__ void k_rotation_svd(int64_t nq, const double* __restrict__ F, double* __restrict__ U,
                               double* __restrict__ S, double* __restrict__ Vt) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double f[9], u[9], sv[3], v[9];
    for (int k = 0; k < 9; ++k) f[k] = F[9 * q + k];
    rotation_svd3(f, u, sv, v);
    for (int k = 0; k < 9; ++k) U[9 * q + k] = u[k];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Vt[9 * q + r + 3 * c] = v[c + 3 * r];
    for (int k = 0; k < 3; ++k) S[3 * q + k] = sv[k];
}

// ================================================================== launchers
#define HOEFEM_DISPATCH(type, ...)                                                 \
    switch (type) {                                                                \
        case TET4: { constexpr int ET = 0; __VA_ARGS__; } break;                   \
        case TET10: { constexpr int ET = 1; __VA_ARGS__; } break;                  \
        case TET20: { constexpr int ET = 2; __VA_ARGS__; } break;                  \
        case HEX8: { constexpr int ET = 3; __VA_ARGS__; } break;                   \
        case HEX20: { constexpr int ET = 4; __VA_ARGS__; } break;                  \
        case HEX27: { constexpr int ET = 5; __VA_ARGS__; } break;                  \
        default: throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad element type");  \
    }

#define HOEFEM_MATERIAL_DISPATCH(kind, ...)                                          \
    switch (kind) {                                                                  \
        case MAT_SNH: { constexpr int MK = MAT_SNH; __VA_ARGS__; } break;            \
        case MAT_INVERTIBLE_SNH: { constexpr int MK = MAT_INVERTIBLE_SNH; __VA_ARGS__; } break; \
        case MAT_PROJECTED_SNH: { constexpr int MK = MAT_PROJECTED_SNH; __VA_ARGS__; } break;   \
        default: throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "bad material kind");   \
    }

void launch_element_forces(Model& m, double mu_r, double lambda_r, const double* u) {
    m.d_fe.ensure((size_t)m.ne * 3 * m.n);
    if (m.ne == 0) return;
    KernelProf* prof = prof_slot(m);
    HOEFEM_MATERIAL_DISPATCH(
        m.material,
        HOEFEM_DISPATCH(m.type, launch_timed(prof, k_element_forces<ET, MK>, dim3((unsigned)m.ne), dim3(64), 0, m.stream,
                                             m.ne, (const int32_t*)m.d_conn.get(), (const double*)m.d_vertices.get(),
                                             (const int32_t*)m.q_ell.off.get(), (const double*)m.q_ell.w.get(),
                                             (const double*)m.q_ell.xi.get(), m.inversion_threshold, mu_r, lambda_r,
                                             u, m.d_fe.get(), m.d_err.get())));
    check_launch("k_element_forces");
}

void launch_element_stiffness(Model& m, double mu_r, double lambda_r, const double* u, bool fp32) {
    if (m.ne == 0) return;
    if (fp32) {
        m.d_ke32.ensure((size_t)m.ne * 9 * m.n * m.n);
        HOEFEM_MATERIAL_DISPATCH(
            m.material,
            HOEFEM_DISPATCH(m.type,
                            k_element_stiffness<ET, float, MK><<<dim3((unsigned)m.ne), dim3(STIFF_THREADS), 0, m.stream>>>(
                                m.ne, m.d_conn.get(), m.d_vertices.get(), m.q_stiff.off.get(), m.q_stiff.w.get(),
                                m.q_stiff.xi.get(), (float)m.inversion_threshold, (float)mu_r, (float)lambda_r, u,
                                m.d_ke32.get(), m.d_err.get())));
    } else {
        m.d_ke.ensure((size_t)m.ne * 9 * m.n * m.n);
        HOEFEM_MATERIAL_DISPATCH(
            m.material,
            HOEFEM_DISPATCH(m.type,
                            launch_timed(prof_slot(m), k_element_stiffness_tw<ET, MK>, dim3((unsigned)m.ne),
                                         dim3(TW_THREADS), 0, m.stream, m.ne, (const int32_t*)m.d_conn.get(),
                                         (const double*)m.d_vertices.get(), (const int32_t*)m.q_stiff.off.get(),
                                         (const double*)m.q_stiff.w.get(), (const double*)m.q_stiff.xi.get(),
                                         m.inversion_threshold, mu_r, lambda_r, u, m.d_ke.get(), m.d_err.get())));
    }
    check_launch("k_element_stiffness");
}

void launch_element_mass(Model& m, double density) {
    m.d_ke.ensure((size_t)m.ne * 9 * m.n * m.n);
    if (m.ne == 0) return;
    const int nreg = m.mass_reg > 0.0 ? (int)m.q_interior_mass.max_points : 0;
    HOEFEM_DISPATCH(m.type, k_element_mass<ET><<<dim3((unsigned)m.ne), dim3(256), 0, m.stream>>>(
                                m.ne, m.d_conn.get(), m.d_vertices.get(), m.q_mass.off.get(), m.q_mass.w.get(),
                                m.q_mass.xi.get(), density, nreg, m.q_interior_mass.w.get(),
                                m.q_interior_mass.xi.get(), m.mass_reg * density, m.d_ke.get()));
    check_launch("k_element_mass");
}

template <class TB, class TV>
void launch_block_gather(const NodePattern& pat, const BlockScatter& sc, const TB* blocks, TV* values, int mode,
                         double c, const double* M, bool zero_first, hipStream_t s, const double* D2, double cd,
                         double cm, KernelProf* prof) {
    if (pat.num_nodes == 0) return;
    if (pat.max_nbr <= GATHER_CHUNK && 9 * sc.max_m <= 64 * GP_VALS && sc.max_inc <= 64) {
        const size_t lds = (size_t)9 * std::max(pat.max_nbr, 1) * sizeof(double);
        launch_timed(prof, k_block_gather_pf<TB, TV>, dim3((unsigned)pat.num_nodes), dim3(64), (uint32_t)lds, s,
                     (const int32_t*)pat.d_nptr.get(), (const int32_t*)pat.d_rowptr.get(),
                     (const uint8_t*)pat.d_diag_only.get(), (const int32_t*)sc.d_inc_ptr.get(),
                     (const int32_t*)sc.d_inc_blk.get(), (const int32_t*)sc.d_inc_loc.get(),
                     (const int32_t*)sc.d_m.get(), (const int64_t*)sc.d_val_off.get(),
                     (const int64_t*)sc.d_slot_off.get(), (const int32_t*)sc.d_slot.get(), blocks, values, mode, c,
                     M, zero_first ? 1 : 0, D2, cd, cm);
        check_launch("k_block_gather_pf");
        return;
    }
    const int chunk = std::min(std::max(pat.max_nbr, 1), GATHER_CHUNK);
    const size_t lds = (size_t)9 * chunk * sizeof(double);
    launch_timed(prof, k_block_gather<TB, TV>, dim3((unsigned)pat.num_nodes), dim3(GB_THREADS), (uint32_t)lds, s,
                 (const int32_t*)pat.d_nptr.get(), (const int32_t*)pat.d_rowptr.get(),
                 (const uint8_t*)pat.d_diag_only.get(), (const int32_t*)sc.d_inc_ptr.get(),
                 (const int32_t*)sc.d_inc_blk.get(), (const int32_t*)sc.d_inc_loc.get(), (const int32_t*)sc.d_m.get(),
                 (const int64_t*)sc.d_val_off.get(), (const int64_t*)sc.d_slot_off.get(),
                 (const int32_t*)sc.d_slot.get(), blocks, values, mode, c, M, zero_first ? 1 : 0, D2, cd, cm, chunk);
    check_launch("k_block_gather");
}
template void launch_block_gather<double, double>(const NodePattern&, const BlockScatter&, const double*,
                                                  double*, int, double, const double*, bool, hipStream_t,
                                                  const double*, double, double, KernelProf*);
template void launch_block_gather<float, float>(const NodePattern&, const BlockScatter&, const float*,
                                                  float*, int, double, const double*, bool, hipStream_t,
                                                  const double*, double, double, KernelProf*);
template void launch_block_gather<double, float>(const NodePattern&, const BlockScatter&, const double*,
                                                  float*, int, double, const double*, bool, hipStream_t,
                                                  const double*, double, double, KernelProf*);
template void launch_block_gather<float, double>(const NodePattern&, const BlockScatter&, const float*,
                                                  double*, int, double, const double*, bool, hipStream_t,
                                                  const double*, double, double, KernelProf*);

void launch_force_gather(Model& m, double* f) {
    if (m.nv == 0) return;
    hipLaunchKernelGGL(k_force_gather, dim3((unsigned)((m.nv + 255) / 256)), dim3(256), 0, m.stream, m.nv,
                       m.scatter.d_inc_ptr.get(), m.scatter.d_inc_blk.get(), m.scatter.d_inc_loc.get(), m.n,
                       m.d_fe.get(), f);
    check_launch("k_force_gather");
}

int spmv_group_size(const NodePattern& pat) {
    if (pat.tuned_group) return pat.tuned_group;
    const double avg = pat.num_nodes ? (double)pat.ncol.size() / (double)pat.num_nodes : 1.0;
    if (avg >= 40.0) return 32;
    if (avg >= 12.0) return 16;
    return 8;
}

template <class VT>
void launch_spmv(const NodePattern& pat, const VT* vals, double beta, double* y, double alpha, const double* x,
                 hipStream_t s) {
    if (pat.num_nodes == 0) return;
    const int G = spmv_group_size(pat);
    const int64_t threads = pat.num_nodes * G;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
#define SPMV_CASE(GG)                                                                                             \
    k_spmv<GG, VT><<<dim3(blocks), dim3(256), 0, s>>>(pat.num_nodes, pat.d_nptr.get(), pat.d_ncol.get(),           \
                                                      pat.d_rowptr.get(), pat.d_diag_only.get(), vals, beta, y, alpha, x)
    if (G == 32)
        SPMV_CASE(32);
    else if (G == 16)
        SPMV_CASE(16);
    else
        SPMV_CASE(8);
#undef SPMV_CASE
    check_launch("k_spmv");
}
template void launch_spmv<double>(const NodePattern&, const double*, double, double*, double, const double*, hipStream_t);
template void launch_spmv<float>(const NodePattern&, const float*, double, double*, double, const double*, hipStream_t);

template <class VT>
void launch_dirichlet(const NodePattern& pat, VT* vals, const uint8_t* d_dirmask, const int32_t* d_bc_nodes,
                      int64_t n_bc_nodes, double* d_scale, hipStream_t s) {
    k_first_nonzero_diag<VT><<<dim3(1), dim3(256), 0, s>>>(pat.ndof(), pat.d_diag_pos.get(), vals, d_scale);
    check_launch("k_first_nonzero_diag");
    if (n_bc_nodes == 0) return;
    k_dirichlet_rows<VT><<<dim3((unsigned)n_bc_nodes), dim3(64), 0, s>>>(d_bc_nodes, n_bc_nodes, pat.d_nptr.get(),
                                                                        pat.d_ncol.get(), pat.d_rowptr.get(),
                                                                        pat.d_diag_only.get(), d_dirmask, vals, d_scale);
    check_launch("k_dirichlet_rows");
}
template void launch_dirichlet<double>(const NodePattern&, double*, const uint8_t*, const int32_t*, int64_t, double*,
                                       hipStream_t);
template void launch_dirichlet<float>(const NodePattern&, float*, const uint8_t*, const int32_t*, int64_t, double*,
                                      hipStream_t);

__global__ void k_fill_pattern_values(int64_t n, double* __restrict__ v, double a, double b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = a + b * (double)(i % 7);
}

int autotune_spmv(NodePattern& pat, hipStream_t s, double* times_us) {
    if (pat.num_nodes == 0) return 8;
    const int64_t nnz = pat.nnz(), ndof = pat.ndof();
    DevBuf<double> vals, x, y;
    vals.alloc(std::max<int64_t>(nnz, 1));
    x.alloc(ndof);
    y.alloc(ndof);
    k_fill_pattern_values<<<dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s>>>(nnz, vals.get(), 1.0, 0.125);
    k_fill_pattern_values<<<dim3((unsigned)((ndof + 255) / 256)), dim3(256), 0, s>>>(ndof, x.get(), -1.0, 0.5);
    check_launch("k_fill_pattern_values");
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const int cand[3] = {8, 16, 32};
    int best = 8;
    double best_t = 1e300;
    const int saved = pat.tuned_group;
    for (int c = 0; c < 3; ++c) {
        pat.tuned_group = cand[c];
        for (int w = 0; w < 2; ++w) launch_spmv(pat, vals.get(), 0.0, y.get(), 1.0, x.get(), s);
        HIP_CHECK(hipEventRecord(e0, s));
        const int reps = 10;
        for (int r = 0; r < reps; ++r) launch_spmv(pat, vals.get(), 0.0, y.get(), 1.0, x.get(), s);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / reps;
        if (times_us) times_us[c] = us;
        if (us < best_t) {
            best_t = us;
            best = cand[c];
        }
    }
    (void)saved;
    pat.tuned_group = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

void material_energy(double mu, double lambda, int64_t nq, const double* F, double* psi) {
    if (nq <= 0) return;
    double mu_r, lambda_r;
    snh_reparam(mu, lambda, &mu_r, &lambda_r);
    hipLaunchKernelGGL(k_material_energy, dim3((unsigned)((nq + 127) / 128)), dim3(128), 0, nullptr, mu, lambda, mu_r,
                       lambda_r, nq, F, psi);
    check_launch("k_material_energy");
    HIP_CHECK(hipDeviceSynchronize());
}

void material_eval(int mat, double thr, double mu_r, double lambda_r, int64_t nq, const double* F, double* P,
                   double* dPdF) {
    DevBuf<int> err;
    err.alloc(1);
    HIP_CHECK(hipMemset(err.get(), 0, sizeof(int)));
    if (nq > 0) {
        hipLaunchKernelGGL(k_material, dim3((unsigned)((nq + 127) / 128)), dim3(128), 0, nullptr, mat, thr, mu_r, lambda_r, nq, F,
                           P, dPdF, err.get());
        check_launch("k_material");
    }
    int h = 0;
    HIP_CHECK(hipMemcpy(&h, err.get(), sizeof(int), hipMemcpyDeviceToHost));
    if (h & KERR_SINGULAR) throw Error(HOEFEM_ERROR_SINGULAR, "singular deformation gradient");
}

void rotation_svd_eval(int64_t nq, const double* F, double* U, double* S, double* Vt) {
    if (nq == 0) return;
    hipLaunchKernelGGL(k_rotation_svd, dim3((unsigned)((nq + 127) / 128)), dim3(128), 0, nullptr, nq, F, U, S, Vt);
    check_launch("k_rotation_svd");
    HIP_CHECK(hipDeviceSynchronize());
}

__global__ void k_lincomb(int64_t n, double* __restrict__ out, double a, const double* __restrict__ A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    v += a * A[i];
    out[i] = v;
}

void launch_lincomb(hipStream_t s, int64_t n, double* out, double a, const double* A) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_lincomb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, out, a, A);
    check_launch("k_lincomb");
}

int read_and_clear_err(Model& m) {
    HIP_CHECK(hipMemcpyAsync(m.h_err.p, m.d_err.get(), sizeof(int), hipMemcpyDeviceToHost, m.stream));
    HIP_CHECK(hipStreamSynchronize(m.stream));
    const int e = *m.h_err.p;
    if (e) HIP_CHECK(hipMemsetAsync(m.d_err.get(), 0, sizeof(int), m.stream));
    return e;
}

}  // namespace hoefem
