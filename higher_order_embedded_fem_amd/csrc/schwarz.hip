// Additive Schwarz preconditioner (simulation_toolbox/src/fem/schwarz_precond.rs:134-342):
// P = sum_s R_s^T B_s^+ R_s + sum_untreated 1 / J_ii, with B_s the restriction of J to the
// subdomain's DOFs and B_s^+ = V diag(1 / max(|lambda|, 1e-12 max|lambda|)) V^T.
//
// Per update: (1) gather every B_s from J (precomputed slots), (2) batched FP64 Gauss-Jordan
// inversion in LDS, one work-group per block, accepted where it certifies that B_s is positive
// definite with kappa_2 < 1e11 (then the eigenvalue clamp is inactive and B_s^+ = B_s^-1),
// (3) for the remaining blocks, symmetric eigen-decomposition by parallel (round-robin) Jacobi
// in LDS and recomposition of B_s^+, (4) color-ordered gather of the blocks into P's row
// blocks, (5) reciprocal diagonal for untreated nodes.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace hoefem {

// the partial sums per norm of a block above 96 x 96 (k_bgj_norm)
constexpr int BGJ_NORM_SPLIT = 32;

void schwarz_setup(Schwarz& s, Model& m, int64_t nsub, const int64_t* off, const int32_t* nodes, int64_t nunt,
                   const int32_t* unt) {
    s.model = &m;
    s.nsub = nsub;
    s.off.assign(off, off + nsub + 1);
    if (s.off[0] != 0) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "subdomain offsets must start at 0");
    for (int64_t k = 0; k < nsub; ++k)
        if (s.off[k + 1] < s.off[k]) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "subdomain offsets must be sorted");
    s.nodes.assign(nodes, nodes + s.off[nsub]);
    s.untreated.assign(unt, unt + nunt);
    for (int32_t v : s.nodes)
        if (v < 0 || v >= m.nv) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "subdomain node out of range");
    s.colors = greedy_coloring(nsub, s.off.data(), s.nodes.data());
    build_node_pattern(m.nv, nsub, s.off.data(), s.nodes.data(), s.untreated, s.P);
    s.scatter.build(s.P, nsub, s.off.data(), s.nodes.data(), &s.colors, 1);
    // slots of every block entry in J's (model) pattern. SchwarzElementAssembler gathers with
    // csr_row.get(col).expect(..) (schwarz_precond.rs:304-320), which panics on a node pair that J
    // does not couple; here such entries are the zeros they are in J (-1 slot), so subdomains may be
    // any node sets, e.g. patches of many elements.
    s.jslot.resize(s.scatter.slot.size());
    s.small_max_dim = s.large_max_dim = 0;
    s.large.clear();
    std::vector<int64_t> colp_off;
    int64_t colp_total = 0;
    for (int64_t b = 0; b < nsub; ++b) {
        const int32_t* g = s.nodes.data() + s.off[b];
        const int mb = (int)(s.off[b + 1] - s.off[b]);
        for (int a = 0; a < mb; ++a)
            for (int c = 0; c < mb; ++c) s.jslot[s.scatter.slot_off[b] + (int64_t)a * mb + c] = m.pattern.find(g[a], g[c]);
        if (3 * mb > SCHWARZ_SMALL_DIM) {
            s.large.push_back((int32_t)b);
            s.large_max_dim = std::max(s.large_max_dim, 3 * mb);
            colp_off.push_back(colp_total);
            colp_total += (int64_t)3 * mb * 32;
        } else {
            s.small_max_dim = std::max(s.small_max_dim, 3 * mb);
        }
    }
    if (s.large_max_dim > 4096) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "Schwarz subdomain above 4096 DOFs");
    for (int32_t v : s.untreated)
        if (v < 0 || v >= m.nv || !s.P.diag_only[v])
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "untreated node must not belong to a subdomain");
    s.P.upload(m.stream);
    s.scatter.upload(m.stream);
    s.d_jslot.upload(s.jslot, m.stream);
    s.d_untreated.upload(s.untreated, m.stream);
    s.d_off.upload(s.off, m.stream);
    s.d_nodes.upload(s.nodes, m.stream);
    if (s.fp32) {
        s.d_vals32.ensure(std::max<int64_t>(s.P.nnz(), 1));
        HIP_CHECK(hipMemsetAsync(s.d_vals32.get(), 0, sizeof(float) * s.d_vals32.n, m.stream));
    } else {
        s.d_vals.ensure(std::max<int64_t>(s.P.nnz(), 1));
        HIP_CHECK(hipMemsetAsync(s.d_vals.get(), 0, sizeof(double) * s.d_vals.n, m.stream));
    }
    s.d_blocks.ensure(std::max<int64_t>(s.scatter.val_off[nsub], 1));
    s.d_fallback.ensure(std::max<int64_t>(nsub, 1));
    if (!s.large.empty()) {
        s.d_large.upload(s.large, m.stream);
        s.d_colp_off.upload(colp_off, m.stream);
        s.d_piv.ensure(s.large.size() * 32 * 32);
        s.d_colp.ensure((size_t)colp_total);
        s.d_lnorm.ensure(2 * s.large.size() * BGJ_NORM_SPLIT);
    }
}

// B_s(r, c) = J[3 g_a + i, 3 g_c + j], row-major (3m)^2, always FP64. Work-group (s, y) gathers
// entries [y SG_CHUNK, (y + 1) SG_CHUNK) of block s, so a block of hundreds of DOFs (the 64-element
// patches) is spread over many work-groups instead of walking its entries on one.
constexpr int SG_CHUNK = 2048;
template <class VJ>
__global__ void k_schwarz_gather(const int32_t* __restrict__ msz, const int64_t* __restrict__ val_off,
                                 const int64_t* __restrict__ slot_off, const int64_t* __restrict__ noff,
                                 const int32_t* __restrict__ nodes, const int32_t* __restrict__ jslot,
                                 const int32_t* __restrict__ jrowptr, const VJ* __restrict__ J,
                                 double* __restrict__ blocks) {
    const int64_t b = blockIdx.x;
    const int mb = msz[b], dim = 3 * mb;
    const int c0 = (int)blockIdx.y * SG_CHUNK;
    if (c0 >= dim * dim) return;
    const int c1 = min(dim * dim, c0 + SG_CHUNK);
    const int32_t* g = nodes + noff[b];
    double* out = blocks + val_off[b];
    for (int idx = c0 + (int)threadIdx.x; idx < c1; idx += blockDim.x) {
        const int r = idx / dim, c = idx % dim;
        const int a = r / 3, i = r % 3, cb = c / 3, j = c % 3;
        const int k = jslot[slot_off[b] + (int64_t)a * mb + cb];
        out[idx] = k < 0 ? 0.0 : (double)J[(int64_t)jrowptr[3 * (int64_t)g[a] + i] + 3 * k + j];
    }
}

// Parallel cyclic Jacobi (Brent-Luk round-robin ordering) on a symmetric dim x dim block in
// LDS, then B^+ = V diag(f(lambda)) V^T with f = 1/max(|lambda|, 1e-12 max|lambda|)
// (invert_element_block, schwarz_precond.rs:329-342). Rotations follow Golub & Van Loan,
// "Matrix Computations", sym.schur2.
constexpr int EIG_THREADS = 256;
// kappa_2 bound below which the eigenvalue clamp max(|lambda|, 1e-12 max|lambda|) is provably
// inactive, with a 2x margin for the rounding of the computed inverse (relative error ~kappa eps)
constexpr double SPD_COND_LIMIT = 5e11;

__device__ inline double block_max(double v, double* s_red) {
    const int t = threadIdx.x;
    s_red[t] = v;
    __syncthreads();
    for (int s = EIG_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) s_red[t] = fmax(s_red[t], s_red[t + s]);
        __syncthreads();
    }
    const double r = s_red[0];
    __syncthreads();
    return r;
}

__device__ inline double block_sum(double v, double* s_red) {
    const int t = threadIdx.x;
    s_red[t] = v;
    __syncthreads();
    for (int s = EIG_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) s_red[t] += s_red[t + s];
        __syncthreads();
    }
    const double r = s_red[0];
    __syncthreads();
    return r;
}

// Gauss-Jordan inversion of a symmetric block without pivoting, held in registers: the 16 x 16
// threads own the entries (i, j) with i = 16 r + tr, j = 16 c + tc (r, c < 6, so dim <= 96), and
// each pivot step broadcasts row k and column k through LDS (double-buffered: one barrier per
// pivot). Each pivot is a diagonal entry of a Schur complement, so all pivots are positive iff B
// is positive definite. Results per block (fallback[b]):
//   0  pivots positive and min(|B|_F |B^-1|_F, |B|_inf |B^-1|_inf) < SPD_COND_LIMIT (both upper
//      bounds of kappa_2(B) for symmetric B): the clamp of the pseudo-inverse is inactive, so
//      B^+ = B^-1, and the symmetrized inverse replaces the block;
//   2  pivots positive, the cheap bound too loose: the inverse replaces the block and
//      k_spd_power_certify decides with a sharper bound;
//   1  a pivot not positive: B is kept for the eigen path.
// RR x RR register tiles per thread: RR = 6 covers blocks up to 96 x 96 (used above 64 x 64; the
// smaller blocks go to k_spd_inverse_mf on the matrix cores, with the same results per block).

// Pivot row k = 16 R + kt and column k of the register-held block into LDS (owners only).
template <int R, int RR>
__device__ inline void gj_publish(const double (&a)[RR][RR], double* srow, double* scol, int tr, int tc, int kt) {
    if constexpr (R >= RR) return;
    else {
    if (tr == kt) {
#pragma unroll
        for (int c = 0; c < RR; ++c) srow[16 * c + tc] = a[R][c];
    }
    if (tc == kt) {
#pragma unroll
        for (int r = 0; r < RR; ++r) scol[16 * r + tr] = a[r][R];
    }
    }
}

// After the rank-1 update: row k -> a_kj / a_kk, column k -> -a_ik / a_kk, the pivot -> 1 / a_kk.
template <int R, int RR>
__device__ inline void gj_fix(double (&a)[RR][RR], const double (&rk)[RR], const double (&ck)[RR], double ip, int tr,
                              int tc, int kt) {
    if constexpr (R >= RR) return;
    else {
    if (tr == kt) {
#pragma unroll
        for (int c = 0; c < RR; ++c) a[R][c] = rk[c];
    }
    if (tc == kt) {
#pragma unroll
        for (int r = 0; r < RR; ++r) a[r][R] = -ck[r] * ip;
    }
    if (tr == kt && tc == kt) a[R][R] = ip;
    }
}

template <int RR>
__global__ __launch_bounds__(EIG_THREADS) void k_spd_inverse(const int32_t* __restrict__ msz,
                                                             const int64_t* __restrict__ val_off,
                                                             double* __restrict__ blocks, int* __restrict__ fallback) {
    __shared__ double s_row[2][16 * RR], s_col[2][16 * RR];
    __shared__ double s_x[RR][EIG_THREADS];  // row-sum partials, then the transpose exchange
    __shared__ double s_red[EIG_THREADS];
    const int64_t b = blockIdx.x;
    const int dim = 3 * msz[b];
    if (dim > SCHWARZ_SMALL_DIM) {  // handled by the blocked path
        if (threadIdx.x == 0) fallback[b] = 4;
        return;
    }
    const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
    double* blk = blocks + val_off[b];
    double a[RR][RR];
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
        for (int c = 0; c < RR; ++c) {
            const int i = 16 * r + tr, j = 16 * c + tc;
            a[r][c] = (i < dim && j < dim) ? blk[(int64_t)i * dim + j] : 0.0;
        }
    // |.|_F^2 and |.|_inf of the register-held matrix
    auto norms = [&](double& f2, double& finf) {
        double loc = 0.0, rs[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            rs[r] = 0.0;
#pragma unroll
            for (int c = 0; c < RR; ++c) {
                loc += a[r][c] * a[r][c];
                rs[r] += fabs(a[r][c]);
            }
        }
        f2 = block_sum(loc, s_red);  // synchronizes
        double* part = &s_x[0][0];   // [16 (tc)][96 (row)]
#pragma unroll
        for (int r = 0; r < RR; ++r) part[tc * (16 * RR) + 16 * r + tr] = rs[r];
        __syncthreads();
        double row = 0.0;
        if (t < 16 * RR)
            for (int q = 0; q < 16; ++q) row += part[q * (16 * RR) + t];
        finf = block_max(row, s_red);  // synchronizes
    };
    double nb2, nbinf;
    norms(nb2, nbinf);
    for (int k = 0; k < dim; ++k) {
        const int buf = k & 1, kr = k >> 4, kt = k & 15;
        // kr is the same in every thread: the switch branches on it instead of selecting per entry
        switch (kr) {
            case 0: gj_publish<0, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
            case 1: gj_publish<1, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
            case 2: gj_publish<2, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
            case 3: gj_publish<3, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
            case 4: gj_publish<4, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
            default: gj_publish<5, RR>(a, s_row[buf], s_col[buf], tr, tc, kt); break;
        }
        __syncthreads();
        const double piv = s_row[buf][k];
        if (!(piv > 0.0) || !isfinite(piv)) {  // same value in every thread: uniform exit
            if (t == 0) fallback[b] = 1;
            return;
        }
        const double ip = 1.0 / piv;
        double rk[RR], ck[RR];
#pragma unroll
        for (int c = 0; c < RR; ++c) rk[c] = s_row[buf][16 * c + tc] * ip;
#pragma unroll
        for (int r = 0; r < RR; ++r) ck[r] = s_col[buf][16 * r + tr];
        // a_ij - a_ik a_kj / a_kk everywhere, then row k -> a_kj / a_kk, column k -> -a_ik / a_kk,
        // the pivot -> 1 / a_kk
#pragma unroll
        for (int r = 0; r < RR; ++r)
#pragma unroll
            for (int c = 0; c < RR; ++c) a[r][c] -= ck[r] * rk[c];
        switch (kr) {
            case 0: gj_fix<0, RR>(a, rk, ck, ip, tr, tc, kt); break;
            case 1: gj_fix<1, RR>(a, rk, ck, ip, tr, tc, kt); break;
            case 2: gj_fix<2, RR>(a, rk, ck, ip, tr, tc, kt); break;
            case 3: gj_fix<3, RR>(a, rk, ck, ip, tr, tc, kt); break;
            case 4: gj_fix<4, RR>(a, rk, ck, ip, tr, tc, kt); break;
            default: gj_fix<5, RR>(a, rk, ck, ip, tr, tc, kt); break;
        }
    }
    double ni2, niinf;
    norms(ni2, niinf);
    const double cert = fmin(sqrt(nb2) * sqrt(ni2), nbinf * niinf);
    // symmetrize: the transpose of entry (r, c) sits in thread (tc, tr) at (c, r); round r
    // exchanges the entries (r, 0..5)
    const int partner = tc * 16 + tr;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
#pragma unroll
        for (int x = 0; x < RR; ++x) s_x[x][t] = a[x][r];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < RR; ++c) {
            const int i = 16 * r + tr, j = 16 * c + tc;
            if (i < dim && j < dim) blk[(int64_t)i * dim + j] = 0.5 * (a[r][c] + s_x[c][partner]);
        }
        __syncthreads();
    }
    if (t == 0) fallback[b] = cert < SPD_COND_LIMIT ? 0 : 2;
}

typedef double sgj_double4 __attribute__((ext_vector_type(4)));

// The same inversion on the matrix cores: in-place blocked Gauss-Jordan over NT x NT tiles of 16 x 16
// (dim <= 16 NT; the padding is the identity), one work-group of four wavefronts per block, the block
// in LDS (row stride LD = 16 NT + 1). For each diagonal tile k:
//   D = A_kk^-1        wavefront 0, pivot by pivot (the pivots are those of the unblocked elimination,
//                      so a non-positive one marks the block exactly as k_spd_inverse does)
//   A_kj = D A_kj      (j != k)
//   A_ij -= A_ik A_kj  (i, j != k; the new A_kj, the old A_ik)
//   A_ik = -A_ik D     (i != k);  A_kk = D
// The tile products run on v_mfma_f64_16x16x4 (lane l: A operand X(l % 16, 4 s + l / 16), B operand
// Y(4 s + l / 16, l % 16), result item v = (4 v + l / 16, l % 16)), the tiles of each step dealt to
// the wavefronts in turn. The certificate and the symmetrized write-back are k_spd_inverse's.
template <int NT>
__global__ __launch_bounds__(EIG_THREADS) void k_spd_inverse_mf(const int32_t* __restrict__ msz,
                                                                const int64_t* __restrict__ val_off,
                                                                double* __restrict__ blocks, int* __restrict__ fallback) {
    constexpr int PD = 16 * NT, LD = PD + 1;
    extern __shared__ double s_a[];  // [PD][LD]
    __shared__ double s_red[EIG_THREADS];
    __shared__ int s_bad;
    const int64_t b = blockIdx.x;
    const int dim = 3 * msz[b];
    if (dim > SCHWARZ_SMALL_DIM || dim > PD) {  // handled by the blocked path in global memory
        if (threadIdx.x == 0) fallback[b] = 4;
        return;
    }
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    double* blk = blocks + val_off[b];
    for (int idx = t; idx < PD * PD; idx += EIG_THREADS) {
        const int i = idx / PD, j = idx % PD;
        s_a[i * LD + j] = (i < dim && j < dim) ? blk[(int64_t)i * dim + j] : (i == j ? 1.0 : 0.0);
    }
    if (t == 0) s_bad = 0;
    __syncthreads();
    // |.|_F^2 and |.|_inf of the block's dim x dim part
    auto norms = [&](double& f2, double& finf) {
        double loc = 0.0, row = 0.0;
        for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
            const double x = s_a[(idx / dim) * LD + idx % dim];
            loc += x * x;
        }
        if (t < dim)
            for (int j = 0; j < dim; ++j) row += fabs(s_a[t * LD + j]);
        f2 = block_sum(loc, s_red);
        finf = block_max(row, s_red);
    };
    double nb2, nbinf;
    norms(nb2, nbinf);
    // C = (neg ? -1 : 1) X Y (+ Cin), 16 x 16 tiles at (xr, xc), (yr, yc) of s_a
    auto tile_mul = [&](int xr, int xc, int yr, int yc, bool neg, sgj_double4 acc) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const double x = s_a[(xr + (l & 15)) * LD + xc + 4 * s4 + (l >> 4)];
            const double y = s_a[(yr + 4 * s4 + (l >> 4)) * LD + yc + (l & 15)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(neg ? -x : x, y, acc, 0, 0, 0);
        }
        return acc;
    };
    auto tile_store = [&](int r0, int c0, const sgj_double4& acc) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_a[(r0 + 4 * v + (l >> 4)) * LD + c0 + (l & 15)] = acc[v];
    };
    auto tile_load = [&](int r0, int c0) {
        sgj_double4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = s_a[(r0 + 4 * v + (l >> 4)) * LD + c0 + (l & 15)];
        return acc;
    };
    const sgj_double4 zero = {0.0, 0.0, 0.0, 0.0};
    for (int kt = 0; kt < NT && 16 * kt < dim; ++kt) {
        const int K = 16 * kt;
        if (w == 0) {  // D = A_kk^-1 in place, one wavefront: its LDS accesses take effect in order
            for (int p = 0; p < 16; ++p) {
                const double piv = s_a[(K + p) * LD + K + p];
                if (!(piv > 0.0) || !isfinite(piv)) {
                    if (l == 0) s_bad = 1;
                    break;
                }
                const double ip = 1.0 / piv;
                double nv[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int e = l + 64 * m, i = e >> 4, j = e & 15;
                    const double aij = s_a[(K + i) * LD + K + j];
                    const double aip = s_a[(K + i) * LD + K + p], apj = s_a[(K + p) * LD + K + j];
                    nv[m] = (i == p && j == p) ? ip : (i == p ? apj * ip : (j == p ? -aip * ip : aij - aip * apj * ip));
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int e = l + 64 * m;
                    s_a[(K + (e >> 4)) * LD + K + (e & 15)] = nv[m];
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        if (s_bad) {  // the same in every thread: uniform exit
            if (t == 0) fallback[b] = 1;
            return;
        }
        // row panel A_kj = D A_kj: the tile's old values are read (by its wavefront) before it is written
        for (int j = 0, n = 0; j < NT; ++j) {
            if (j == kt) continue;
            if ((n++ & 3) == w) tile_store(K, 16 * j, tile_mul(K, K, K, 16 * j, false, zero));
        }
        __syncthreads();
        for (int i = 0, n = 0; i < NT; ++i) {
            if (i == kt) continue;
            for (int j = 0; j < NT; ++j) {
                if (j == kt) continue;
                if ((n++ & 3) == w) tile_store(16 * i, 16 * j, tile_mul(16 * i, K, K, 16 * j, true, tile_load(16 * i, 16 * j)));
            }
        }
        __syncthreads();
        for (int i = 0, n = 0; i < NT; ++i) {
            if (i == kt) continue;
            if ((n++ & 3) == w) tile_store(16 * i, K, tile_mul(16 * i, K, K, K, true, zero));
        }
        __syncthreads();
    }
    double ni2, niinf;
    norms(ni2, niinf);
    const double cert = fmin(sqrt(nb2) * sqrt(ni2), nbinf * niinf);
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const int i = idx / dim, j = idx % dim;
        blk[(int64_t)i * dim + j] = 0.5 * (s_a[i * LD + j] + s_a[j * LD + i]);
    }
    if (t == 0) fallback[b] = cert < SPD_COND_LIMIT ? 0 : 2;
}

// Positive semi-definiteness certificate of each symmetric block: Cholesky of B + delta I in LDS
// with delta = 1e-11 |B|_F. When every pivot is positive, lambda_min(B) > -delta, so clamping the
// eigenvalues at zero would move B by less than sqrt(dim) delta (the order of the eigensolver's
// own rounding): need[b] = 0 and the block is kept. Otherwise need[b] = 1 and the block goes to
// the eigen path.
__global__ __launch_bounds__(EIG_THREADS) void k_psd_certify(const int32_t* __restrict__ msz,
                                                             const int64_t* __restrict__ val_off,
                                                             const double* __restrict__ blocks, int* __restrict__ need,
                                                             int ldpad) {
    extern __shared__ double lds[];
    __shared__ double s_red[EIG_THREADS];
    const int64_t b = blockIdx.x;
    const int dim = 3 * msz[b];
    const int ld = ldpad;
    double* A = lds;
    const double* blk = blocks + val_off[b];
    const int t = threadIdx.x;
    double loc = 0.0;
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const double v = blk[idx];
        A[(idx / dim) * ld + idx % dim] = v;
        loc += v * v;
    }
    const double delta = 1e-11 * sqrt(block_sum(loc, s_red));  // block_sum synchronizes
    for (int i = t; i < dim; i += EIG_THREADS) A[i * ld + i] += delta;
    __syncthreads();
    // right-looking Cholesky on the lower triangle
    for (int k = 0; k < dim; ++k) {
        const double piv = A[k * ld + k];
        if (!(piv > 0.0)) {  // uniform: every thread reads the same LDS value
            if (t == 0) need[b] = 1;
            return;
        }
        const double d = sqrt(piv);
        __syncthreads();
        for (int i = k + 1 + t; i < dim; i += EIG_THREADS) A[i * ld + k] /= d;
        __syncthreads();
        const int m = dim - k - 1;
        for (int idx = t; idx < m * m; idx += EIG_THREADS) {
            const int i = k + 1 + idx / m, j = k + 1 + idx % m;
            if (j <= i) A[i * ld + j] -= A[i * ld + k] * A[j * ld + k];
        }
        __syncthreads();
    }
    if (t == 0) need[b] = 0;
}

// Positive semi-definite projection P = (B + B sign(B)) / 2 of each flagged symmetric block by the
// Newton-Schulz iteration for the matrix sign, X_0 = B / |B|_F, X_{k+1} = X_k (3 I - X_k^2) / 2
// (Higham, "Functions of Matrices", SIAM 2008, sec. 5.3). Every X_k is a polynomial in B, so the
// eigenvectors are B's and each eigenvalue x follows the scalar map: x in (0, 1] goes to 1, x < 0
// to -1, x = 0 stays. PSD_NS_ITERS steps settle every eigenvalue with |lambda| >= 1.5^-59 |B|_F
// (~4e-11 |B|_F); smaller ones contribute at most |lambda| to the error. The block is dim x dim
// in LDS padded to a multiple of 4; products in 4 x 4 register tiles, one or two per thread.
constexpr int PSD_NS_ITERS = 64;

__device__ inline void tile_gemm(const double* __restrict__ X, const double* __restrict__ Y, int dp, int ld, int ti,
                                 int tj, double (&acc)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
    for (int k = 0; k < dp; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) x[a] = X[(4 * ti + a) * ld + k];
#pragma unroll
        for (int c = 0; c < 4; ++c) y[c] = Y[k * ld + 4 * tj + c];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[a][c] += x[a] * y[c];
    }
}

__global__ __launch_bounds__(EIG_THREADS) void k_psd_newton_schulz(const int32_t* __restrict__ msz,
                                                                   const int64_t* __restrict__ val_off,
                                                                   double* __restrict__ blocks,
                                                                   const int* __restrict__ need, int ldpad) {
    if (need[blockIdx.x] == 0) return;
    extern __shared__ double lds[];
    __shared__ double s_red[EIG_THREADS];
    const int64_t b = blockIdx.x;
    const int dim = 3 * msz[b];
    const int dp = (dim + 3) & ~3;
    const int ld = ldpad;  // >= dp
    const int nt = dp / 4, ntiles = nt * nt;
    double* X = lds;
    double* T = lds + (size_t)dp * ld;
    double* blk = blocks + val_off[b];
    const int t = threadIdx.x;
    double loc = 0.0;
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const double v = blk[idx];
        loc += v * v;
    }
    const double nrm = sqrt(block_sum(loc, s_red));
    if (nrm == 0.0) return;  // uniform
    const double inv = 1.0 / nrm;
    for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
        const int r = idx / dp, c = idx % dp;
        X[r * ld + c] = (r < dim && c < dim) ? 0.5 * (blk[r * dim + c] + blk[c * dim + r]) * inv : 0.0;
    }
    __syncthreads();
    double acc[2][4][4];
    for (int it = 0; it < PSD_NS_ITERS; ++it) {
        // T = X X
        for (int q = 0; q < 2; ++q) {
            const int tile = t + q * EIG_THREADS;
            if (tile < ntiles) tile_gemm(X, X, dp, ld, tile / nt, tile % nt, acc[q]);
        }
        for (int q = 0; q < 2; ++q) {
            const int tile = t + q * EIG_THREADS;
            if (tile < ntiles) {
                const int ti = tile / nt, tj = tile % nt;
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) T[(4 * ti + a) * ld + 4 * tj + c] = acc[q][a][c];
            }
        }
        __syncthreads();
        // X <- 1.5 X - 0.5 X T
        for (int q = 0; q < 2; ++q) {
            const int tile = t + q * EIG_THREADS;
            if (tile < ntiles) tile_gemm(X, T, dp, ld, tile / nt, tile % nt, acc[q]);
        }
        __syncthreads();
        for (int q = 0; q < 2; ++q) {
            const int tile = t + q * EIG_THREADS;
            if (tile < ntiles) {
                const int ti = tile / nt, tj = tile % nt;
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        double* x = &X[(4 * ti + a) * ld + 4 * tj + c];
                        *x = 1.5 * *x - 0.5 * acc[q][a][c];
                    }
            }
        }
        __syncthreads();
    }
    // P = (B + B S) / 2 with B = nrm X_0: X_0 back into T, then the product
    for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
        const int r = idx / dp, c = idx % dp;
        T[r * ld + c] = (r < dim && c < dim) ? 0.5 * (blk[r * dim + c] + blk[c * dim + r]) : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < 2; ++q) {
        const int tile = t + q * EIG_THREADS;
        if (tile < ntiles) tile_gemm(T, X, dp, ld, tile / nt, tile % nt, acc[q]);
    }
    __syncthreads();
    for (int q = 0; q < 2; ++q) {
        const int tile = t + q * EIG_THREADS;
        if (tile < ntiles) {
            const int ti = tile / nt, tj = tile % nt;
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 4; ++c) {
                    const int r = 4 * ti + a, cc = 4 * tj + c;
                    if (r < dim && cc < dim) X[r * ld + cc] = 0.5 * (T[r * ld + cc] + acc[q][a][c]);
                }
        }
    }
    __syncthreads();
    // symmetrize on the way out (B S is symmetric in exact arithmetic)
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const int r = idx / dim, c = idx % dim;
        blk[idx] = 0.5 * (X[r * ld + c] + X[c * ld + r]);
    }
}

// Sharper certificate for the blocks k_spd_inverse marked 2. For symmetric X with eigenvalues
// lambda, |X^8|_F = (sum lambda^16)^(1/2) >= max|lambda|^8, so
//     |X|_2 <= s |(X / s)^8|_F^(1/8),  s = |X|_F,
// an upper bound within a factor dim^(1/16) (<= 1.33) of |X|_2 instead of the factor up to dim of
// the norms above. Three squarings in LDS for B (gathered again from J) and for the inverse now in
// the block; kappa_2(B) <= |B|_2 |B^-1|_2 < SPD_COND_LIMIT accepts the inverse, otherwise B goes
// back into the block for the eigen path.
template <class VJ>
__global__ __launch_bounds__(EIG_THREADS) void k_spd_power_certify(
    const int32_t* __restrict__ msz, const int64_t* __restrict__ val_off, const int64_t* __restrict__ slot_off,
    const int64_t* __restrict__ noff, const int32_t* __restrict__ nodes, const int32_t* __restrict__ jslot,
    const int32_t* __restrict__ jrowptr, const VJ* __restrict__ J, double* __restrict__ blocks,
    int* __restrict__ fallback, int ldpad) {
    if (fallback[blockIdx.x] != 2) return;
    extern __shared__ double lds[];
    __shared__ double s_red[EIG_THREADS];
    const int64_t b = blockIdx.x;
    const int mb = msz[b], dim = 3 * mb, dp = (dim + 3) & ~3, ld = ldpad;
    if (dim > SCHWARZ_SMALL_DIM) return;
    double* X = lds;
    double* Y = lds + (size_t)dp * ld;
    double* blk = blocks + val_off[b];
    const int32_t* g = nodes + noff[b];
    const int t = threadIdx.x;
    auto entry = [&](int r, int c) -> double {  // B(r, c) = J[3 g_a + i, 3 g_cb + j]
        const int a = r / 3, i = r % 3, cb = c / 3, j = c % 3;
        const int k = jslot[slot_off[b] + (int64_t)a * mb + cb];
        return k < 0 ? 0.0 : (double)J[(int64_t)jrowptr[3 * (int64_t)g[a] + i] + 3 * k + j];
    };
    auto bound = [&]() -> double {  // spectral-norm bound of the symmetric matrix in X
        double loc = 0.0;
        for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
            const double v = X[(idx / dp) * ld + idx % dp];
            loc += v * v;
        }
        const double s = sqrt(block_sum(loc, s_red));  // synchronizes
        if (!(s > 0.0)) return 0.0;                    // uniform
        const double is = 1.0 / s;
        for (int idx = t; idx < dp * dp; idx += EIG_THREADS) X[(idx / dp) * ld + idx % dp] *= is;
        __syncthreads();
        const int nt = dp / 4;
        double* src = X;
        double* dst = Y;
        for (int sq = 0; sq < 3; ++sq) {
            for (int tile = t; tile < nt * nt; tile += EIG_THREADS) {
                const int ti = tile / nt, tj = tile % nt;
                double acc[4][4];
                tile_gemm(src, src, dp, ld, ti, tj, acc);
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c) dst[(4 * ti + a) * ld + 4 * tj + c] = acc[a][c];
            }
            __syncthreads();
            double* tmp = src;
            src = dst;
            dst = tmp;
        }
        loc = 0.0;
        for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
            const double v = src[(idx / dp) * ld + idx % dp];
            loc += v * v;
        }
        const double n8 = sqrt(block_sum(loc, s_red));
        return s * sqrt(sqrt(sqrt(n8)));
    };
    for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
        const int r = idx / dp, c = idx % dp;
        X[r * ld + c] = (r < dim && c < dim) ? entry(r, c) : 0.0;
    }
    __syncthreads();
    const double nb = bound();
    for (int idx = t; idx < dp * dp; idx += EIG_THREADS) {
        const int r = idx / dp, c = idx % dp;
        X[r * ld + c] = (r < dim && c < dim) ? blk[(int64_t)r * dim + c] : 0.0;
    }
    __syncthreads();
    const double ni = bound();
    if (nb * ni < SPD_COND_LIMIT) {
        if (t == 0) fallback[b] = 0;
        return;
    }
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) blk[idx] = entry(idx / dim, idx % dim);
    if (t == 0) fallback[b] = 1;
}

// Cyclic Jacobi eigen-decomposition of each symmetric block in LDS, then the block is replaced by
// V f(lambda) V^T: project == 0, invert_element_block's clamped inverse (schwarz_precond.rs:329-342);
// project == 1, the positive semi-definite projection max(lambda, 0) of DefaultSemidefiniteProjection
// (fenris/src/assembly.rs:552-564).
__global__ __launch_bounds__(EIG_THREADS) void k_jacobi_inverse(const int32_t* __restrict__ msz,
                                                                const int64_t* __restrict__ val_off,
                                                                double* __restrict__ blocks,
                                                                const int* __restrict__ fallback, int ldpad,
                                                                int project) {
    if (fallback && (fallback[blockIdx.x] == 0 || fallback[blockIdx.x] >= 4)) return;
    extern __shared__ double lds[];
    const int64_t b = blockIdx.x;
    const int dim = 3 * msz[b];
    if (fallback && dim > SCHWARZ_SMALL_DIM) return;  // the LDS is sized for the small blocks
    const int nn = dim + (dim & 1);  // even number of players (one idle index when dim is odd)
    const int ld = ldpad;            // row stride (>= nn + 1)
    double* A = lds;
    double* V = lds + (size_t)nn * ld;
    double* cs = V + (size_t)nn * ld;  // [nn/2][2]
    int* perm = (int*)(cs + nn);
    __shared__ double s_red[EIG_THREADS];
    __shared__ int s_done;
    double* blk = blocks + val_off[b];
    const int t = threadIdx.x;
    for (int idx = t; idx < nn * nn; idx += EIG_THREADS) {
        const int r = idx / nn, c = idx % nn;
        A[r * ld + c] = (r < dim && c < dim) ? blk[r * dim + c] : 0.0;
        V[r * ld + c] = (r == c) ? 1.0 : 0.0;
    }
    for (int k = t; k < nn; k += EIG_THREADS) perm[k] = k;
    __syncthreads();
    // total Frobenius norm^2 for the stopping rule
    double loc = 0.0;
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const double v = A[(idx / dim) * ld + idx % dim];
        loc += v * v;
    }
    s_red[t] = loc;
    __syncthreads();
    for (int s = EIG_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) s_red[t] += s_red[t + s];
        __syncthreads();
    }
    const double total = s_red[0];
    __syncthreads();
    const int half = nn / 2;
    for (int sweep = 0; sweep < 40; ++sweep) {
        // off-diagonal norm^2
        loc = 0.0;
        for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
            const int r = idx / dim, c = idx % dim;
            if (r != c) {
                const double v = A[r * ld + c];
                loc += v * v;
            }
        }
        s_red[t] = loc;
        __syncthreads();
        for (int s = EIG_THREADS / 2; s > 0; s >>= 1) {
            if (t < s) s_red[t] += s_red[t + s];
            __syncthreads();
        }
        if (t == 0) s_done = (s_red[0] <= 1e-32 * total || s_red[0] == 0.0) ? 1 : 0;
        __syncthreads();
        if (s_done) break;
        for (int round = 0; round < nn - 1; ++round) {
            // rotation parameters for each disjoint pair
            for (int k = t; k < half; k += EIG_THREADS) {
                int p = perm[k], q = perm[nn - 1 - k];
                if (p > q) { const int tmp = p; p = q; q = tmp; }
                double c = 1.0, s = 0.0;
                if (q < dim) {
                    const double apq = A[p * ld + q];
                    if (apq != 0.0) {
                        const double app = A[p * ld + p], aqq = A[q * ld + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                    }
                }
                cs[2 * k] = c;
                cs[2 * k + 1] = s;
            }
            __syncthreads();
            // columns: A <- A R, V <- V R
            for (int idx = t; idx < half * nn; idx += EIG_THREADS) {
                const int k = idx / nn, r = idx % nn;
                int p = perm[k], q = perm[nn - 1 - k];
                if (p > q) { const int tmp = p; p = q; q = tmp; }
                const double c = cs[2 * k], s = cs[2 * k + 1];
                if (s == 0.0) continue;
                const double arp = A[r * ld + p], arq = A[r * ld + q];
                A[r * ld + p] = c * arp - s * arq;
                A[r * ld + q] = s * arp + c * arq;
                const double vrp = V[r * ld + p], vrq = V[r * ld + q];
                V[r * ld + p] = c * vrp - s * vrq;
                V[r * ld + q] = s * vrp + c * vrq;
            }
            __syncthreads();
            // rows: A <- R^T A
            for (int idx = t; idx < half * nn; idx += EIG_THREADS) {
                const int k = idx / nn, r = idx % nn;
                int p = perm[k], q = perm[nn - 1 - k];
                if (p > q) { const int tmp = p; p = q; q = tmp; }
                const double c = cs[2 * k], s = cs[2 * k + 1];
                if (s == 0.0) continue;
                const double apr = A[p * ld + r], aqr = A[q * ld + r];
                A[p * ld + r] = c * apr - s * aqr;
                A[q * ld + r] = s * apr + c * aqr;
            }
            __syncthreads();
            for (int k = t; k < half; k += EIG_THREADS) {
                int p = perm[k], q = perm[nn - 1 - k];
                if (cs[2 * k + 1] != 0.0) {
                    A[p * ld + q] = 0.0;
                    A[q * ld + p] = 0.0;
                }
            }
            __syncthreads();
            // round-robin: keep perm[0], rotate perm[1..nn-1]
            if (t == 0) {
                const int last = perm[nn - 1];
                for (int k = nn - 1; k > 1; --k) perm[k] = perm[k - 1];
                perm[1] = last;
            }
            __syncthreads();
        }
    }
    // eigenvalues on the diagonal; f(lambda) into cs[0..dim)
    loc = 0.0;
    for (int k = t; k < dim; k += EIG_THREADS) loc = fmax(loc, fabs(A[k * ld + k]));
    s_red[t] = loc;
    __syncthreads();
    for (int s = EIG_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) s_red[t] = fmax(s_red[t], s_red[t + s]);
        __syncthreads();
    }
    const double lower = 1e-12 * s_red[0];
    __syncthreads();
    for (int k = t; k < dim; k += EIG_THREADS) {
        if (project) {
            cs[k] = fmax(A[k * ld + k], 0.0);
        } else {
            const double ev = fabs(A[k * ld + k]);
            cs[k] = ev >= lower ? 1.0 / ev : 1.0 / lower;
        }
    }
    __syncthreads();
    for (int idx = t; idx < dim * dim; idx += EIG_THREADS) {
        const int r = idx / dim, c = idx % dim;
        double s = 0.0;
        for (int k = 0; k < dim; ++k) s += V[r * ld + k] * cs[k] * V[c * ld + k];
        blk[idx] = s;
    }
}

// ------------------------------------------------------------------ blocks above 96 x 96
// Blocked Gauss-Jordan inversion in global memory, 32 x 32 tiles, all large blocks of an update
// together. For panel k (rows / columns 32k ..): (1) the pivot tile A_kk is inverted in LDS
// (P = A_kk^-1, unpivoted: every pivot is a diagonal entry of a Schur complement, positive iff the
// block is positive definite); (2) the row panel A_kj <- P A_kj (j != k) and a copy C of the old
// column panel A_ik (i != k); (3) A_ij -= C_i A_kj (i, j != k), A_ik <- -C_i P, A_kk <- P. Steps 2
// and 3 multiply 32 x 32 tiles on v_mfma_f64_16x16x4 (each of four wavefronts a 16 x 16 quarter).
// The inverse is accepted when |B|_F |B^-1|_F < SPD_COND_LIMIT (an upper bound of kappa_2): the
// eigenvalue clamp is then inactive and B^+ = B^-1; it is symmetrized as in k_spd_inverse. A block
// without that certificate (indefinite or too ill-conditioned) raises KERR_SCHWARZ: large blocks have
// no eigen fallback.
constexpr int BT = 32, BTP = 33;  // tile size, LDS row stride

typedef double bgj_double4 __attribute__((ext_vector_type(4)));

// dst(32 x 32 tile in LDS) = src rows r0.., cols c0.. of a row-major dim x dim block, zero outside
__device__ inline void bgj_load(double* dst, const double* src, int dim, int r0, int c0, double sign) {
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
        const int r = idx / BT, c = idx % BT;
        dst[r * BTP + c] = (r0 + r < dim && c0 + c < dim) ? sign * src[(int64_t)(r0 + r) * dim + c0 + c] : 0.0;
    }
}

// acc (this wavefront's 16 x 16 quarter (si, sj) of a 32 x 32 tile) += A B, A and B 32 x 32 in LDS
__device__ inline void bgj_mfma(bgj_double4& acc, const double* A, const double* B, int si, int sj) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int ks = 0; ks < BT / 4; ++ks) {
        const double a = A[(16 * si + (l & 15)) * BTP + 4 * ks + (l >> 4)];
        const double b = B[(4 * ks + (l >> 4)) * BTP + 16 * sj + (l & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
}

// |B|_F^2 (which = 0) or |B^-1|_F^2 (which = 1) of every large block in BGJ_NORM_SPLIT partial sums:
// work-group (l, y) sums the y-th contiguous chunk of block l's entries into
// out[(2 l + which) BGJ_NORM_SPLIT + y]; k_bgj_finish adds the partials in order.
__global__ __launch_bounds__(EIG_THREADS) void k_bgj_norm(const int32_t* __restrict__ large,
                                                          const int32_t* __restrict__ msz,
                                                          const int64_t* __restrict__ val_off,
                                                          const double* __restrict__ blocks, double* __restrict__ out,
                                                          int which) {
    __shared__ double s_red[EIG_THREADS];
    const int b = large[blockIdx.x];
    const int dim = 3 * msz[b];
    const int64_t n = (int64_t)dim * dim;
    const int64_t chunk = (n + BGJ_NORM_SPLIT - 1) / BGJ_NORM_SPLIT;
    const int64_t c0 = blockIdx.y * chunk, c1 = min(n, c0 + chunk);
    const double* blk = blocks + val_off[b];
    double loc = 0.0;
    for (int64_t idx = c0 + threadIdx.x; idx < c1; idx += EIG_THREADS) loc += blk[idx] * blk[idx];
    const double tot = block_sum(loc, s_red);
    if (threadIdx.x == 0) out[(2 * (int64_t)blockIdx.x + which) * BGJ_NORM_SPLIT + blockIdx.y] = tot;
}

// (1) P = A_kk^-1 for every large block with a panel k
__global__ __launch_bounds__(EIG_THREADS) void k_bgj_pivot(const int32_t* __restrict__ large,
                                                           const int32_t* __restrict__ msz,
                                                           const int64_t* __restrict__ val_off,
                                                           const double* __restrict__ blocks, double* __restrict__ piv,
                                                           int* __restrict__ fallback, int* err, int k) {
    __shared__ double T[BT * BTP];
    __shared__ double s_row[BT], s_col[BT];
    const int lb = blockIdx.x, b = large[lb];
    const int dim = 3 * msz[b];
    if (BT * k >= dim || fallback[b] != 4) return;
    const int bk = min(BT, dim - BT * k);
    const double* blk = blocks + val_off[b];
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
        const int r = idx / BT, c = idx % BT;
        T[r * BTP + c] = (r < bk && c < bk) ? blk[(int64_t)(BT * k + r) * dim + BT * k + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int p = 0; p < bk; ++p) {
        if (threadIdx.x < BT) {
            s_row[threadIdx.x] = T[p * BTP + threadIdx.x];
            s_col[threadIdx.x] = T[threadIdx.x * BTP + p];
        }
        __syncthreads();
        const double pv = s_row[p];
        if (!(pv > 0.0) || !isfinite(pv)) {  // uniform: the block is not positive definite
            if (threadIdx.x == 0) {
                fallback[b] = 5;
                atomicOr(err, KERR_SCHWARZ);
            }
            return;
        }
        const double ip = 1.0 / pv;
        for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
            const int r = idx / BT, c = idx % BT;
            double v = T[r * BTP + c];
            if (r == p && c == p)
                v = ip;
            else if (r == p)
                v = s_row[c] * ip;
            else if (c == p)
                v = -s_col[r] * ip;
            else
                v -= s_col[r] * s_row[c] * ip;
            T[r * BTP + c] = v;
        }
        __syncthreads();
    }
    double* out = piv + (int64_t)lb * BT * BT;
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) out[idx] = T[(idx / BT) * BTP + idx % BT];
}

// (2) row panel A_kj <- P A_kj and the copy C_j = A_jk of the column panel, tile j != k
__global__ __launch_bounds__(EIG_THREADS) void k_bgj_panel(const int32_t* __restrict__ large,
                                                           const int32_t* __restrict__ msz,
                                                           const int64_t* __restrict__ val_off,
                                                           double* __restrict__ blocks, const double* __restrict__ piv,
                                                           double* __restrict__ colp,
                                                           const int64_t* __restrict__ colp_off,
                                                           const int* __restrict__ fallback, int k) {
    __shared__ double sP[BT * BTP], sR[BT * BTP];
    const int lb = blockIdx.x, b = large[lb], t = blockIdx.y;
    const int dim = 3 * msz[b];
    const int nt = (dim + BT - 1) / BT;
    if (BT * k >= dim || t >= nt || t == k || fallback[b] != 4) return;
    double* blk = blocks + val_off[b];
    // the column panel's tile (rows 32 t.., columns 32 k..), before step 3 overwrites it
    double* C = colp + colp_off[lb];
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
        const int r = idx / BT, c = idx % BT;
        if (BT * t + r < dim && BT * k + c < dim)
            C[(int64_t)(BT * t + r) * BT + c] = blk[(int64_t)(BT * t + r) * dim + BT * k + c];
    }
    const double* P = piv + (int64_t)lb * BT * BT;
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) sP[(idx / BT) * BTP + idx % BT] = P[idx];
    bgj_load(sR, blk, dim, BT * k, BT * t, 1.0);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, si = w >> 1, sj = w & 1;
    bgj_double4 acc = {0.0, 0.0, 0.0, 0.0};
    bgj_mfma(acc, sP, sR, si, sj);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = BT * k + 16 * si + 4 * v + (l >> 4), c = BT * t + 16 * sj + (l & 15);
        if (r < dim && c < dim) blk[(int64_t)r * dim + c] = acc[v];
    }
}

// (3) A_ij -= C_i A_kj (i, j != k), A_ik = -C_i P, A_kk = P
__global__ __launch_bounds__(EIG_THREADS) void k_bgj_update(const int32_t* __restrict__ large,
                                                            const int32_t* __restrict__ msz,
                                                            const int64_t* __restrict__ val_off,
                                                            double* __restrict__ blocks, const double* __restrict__ piv,
                                                            const double* __restrict__ colp,
                                                            const int64_t* __restrict__ colp_off,
                                                            const int* __restrict__ fallback, int k, int ntm) {
    __shared__ double sA[BT * BTP], sB[BT * BTP];
    const int lb = blockIdx.x, b = large[lb];
    const int ti = blockIdx.y / ntm, tj = blockIdx.y % ntm;
    const int dim = 3 * msz[b];
    const int nt = (dim + BT - 1) / BT;
    if (BT * k >= dim || ti >= nt || tj >= nt || (ti == k && tj != k) || fallback[b] != 4) return;
    double* blk = blocks + val_off[b];
    const double* P = piv + (int64_t)lb * BT * BT;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, si = w >> 1, sj = w & 1;
    bgj_double4 acc = {0.0, 0.0, 0.0, 0.0};
    if (ti == k) {  // tj == k: the pivot tile becomes P
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = P[(16 * si + 4 * v + (l >> 4)) * BT + 16 * sj + (l & 15)];
    } else {
        // -C_i: the column panel copy, rows 32 ti..
        const double* C = colp + colp_off[lb];
        for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
            const int r = idx / BT, c = idx % BT;
            sA[r * BTP + c] = (BT * ti + r < dim && BT * k + c < dim) ? -C[(int64_t)(BT * ti + r) * BT + c] : 0.0;
        }
        if (tj == k) {
            for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) sB[(idx / BT) * BTP + idx % BT] = P[idx];
        } else {
            bgj_load(sB, blk, dim, BT * k, BT * tj, 1.0);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = BT * ti + 16 * si + 4 * v + (l >> 4), c = BT * tj + 16 * sj + (l & 15);
                acc[v] = (r < dim && c < dim) ? blk[(int64_t)r * dim + c] : 0.0;
            }
        }
        __syncthreads();
        bgj_mfma(acc, sA, sB, si, sj);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = BT * ti + 16 * si + 4 * v + (l >> 4), c = BT * tj + 16 * sj + (l & 15);
        if (r < dim && c < dim) blk[(int64_t)r * dim + c] = acc[v];
    }
}

// certificate and symmetrization: tile pairs (ti <= tj) of every large block
__global__ __launch_bounds__(EIG_THREADS) void k_bgj_finish(const int32_t* __restrict__ large,
                                                            const int32_t* __restrict__ msz,
                                                            const int64_t* __restrict__ val_off,
                                                            double* __restrict__ blocks, const double* __restrict__ lnorm,
                                                            int* __restrict__ fallback, int* err, int ntm) {
    __shared__ double sA[BT * BTP], sB[BT * BTP];
    const int lb = blockIdx.x, b = large[lb];
    const int ti = blockIdx.y / ntm, tj = blockIdx.y % ntm;
    const int dim = 3 * msz[b];
    const int nt = (dim + BT - 1) / BT;
    if (ti >= nt || tj >= nt || ti > tj || fallback[b] != 4) return;
    double n2[2] = {0.0, 0.0};  // the partial sums of k_bgj_norm, in order
    for (int w = 0; w < 2; ++w)
        for (int y = 0; y < BGJ_NORM_SPLIT; ++y) n2[w] += lnorm[(2 * (int64_t)lb + w) * BGJ_NORM_SPLIT + y];
    const double cert = sqrt(n2[0]) * sqrt(n2[1]);
    if (!(cert < SPD_COND_LIMIT)) {  // uniform over the block's work-groups
        if (threadIdx.x == 0 && ti == 0 && tj == 0) atomicOr(err, KERR_SCHWARZ);
        return;
    }
    double* blk = blocks + val_off[b];
    bgj_load(sA, blk, dim, BT * ti, BT * tj, 1.0);
    bgj_load(sB, blk, dim, BT * tj, BT * ti, 1.0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < BT * BT; idx += EIG_THREADS) {
        const int r = idx / BT, c = idx % BT;
        const double v = 0.5 * (sA[r * BTP + c] + sB[c * BTP + r]);
        if (BT * ti + r < dim && BT * tj + c < dim) blk[(int64_t)(BT * ti + r) * dim + BT * tj + c] = v;
        if (BT * tj + c < dim && BT * ti + r < dim) blk[(int64_t)(BT * tj + c) * dim + BT * ti + r] = v;
    }
}

static void schwarz_update_large(Schwarz& s) {
    Model& m = *s.model;
    const hipStream_t st = m.stream;
    const unsigned nl = (unsigned)s.large.size();
    const int ntm = (s.large_max_dim + BT - 1) / BT;
    const int32_t* lg = s.d_large.get();
    const int32_t* msz = s.scatter.d_m.get();
    const int64_t* vo = s.scatter.d_val_off.get();
    hipLaunchKernelGGL(k_bgj_norm, dim3(nl, BGJ_NORM_SPLIT), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                       (const double*)s.d_blocks.get(), s.d_lnorm.get(), 0);
    for (int k = 0; k < ntm; ++k) {
        hipLaunchKernelGGL(k_bgj_pivot, dim3(nl), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                           (const double*)s.d_blocks.get(), s.d_piv.get(), s.d_fallback.get(), m.d_err.get(), k);
        hipLaunchKernelGGL(k_bgj_panel, dim3(nl, (unsigned)ntm), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                           s.d_blocks.get(), (const double*)s.d_piv.get(), s.d_colp.get(),
                           (const int64_t*)s.d_colp_off.get(), (const int*)s.d_fallback.get(), k);
        hipLaunchKernelGGL(k_bgj_update, dim3(nl, (unsigned)(ntm * ntm)), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                           s.d_blocks.get(), (const double*)s.d_piv.get(), (const double*)s.d_colp.get(),
                           (const int64_t*)s.d_colp_off.get(), (const int*)s.d_fallback.get(), k, ntm);
    }
    hipLaunchKernelGGL(k_bgj_norm, dim3(nl, BGJ_NORM_SPLIT), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                       (const double*)s.d_blocks.get(), s.d_lnorm.get(), 1);
    hipLaunchKernelGGL(k_bgj_finish, dim3(nl, (unsigned)(ntm * ntm)), dim3(EIG_THREADS), 0, st, lg, msz, vo,
                       s.d_blocks.get(), (const double*)s.d_lnorm.get(), s.d_fallback.get(), m.d_err.get(), ntm);
    check_launch("blocked Gauss-Jordan");
}

// untreated nodes: P_ii = 1 / J_ii (schwarz_precond.rs:195-213)
template <class VJ, class VP>
__global__ void k_schwarz_untreated(int64_t nunt, const int32_t* __restrict__ unt, const int32_t* __restrict__ prowptr,
                                    const int32_t* __restrict__ jdiag, const VJ* __restrict__ J,
                                    VP* __restrict__ P, int* err) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 3 * nunt) return;
    const int64_t row = 3 * (int64_t)unt[idx / 3] + idx % 3;
    const int32_t dp = jdiag[row];
    const double entry = dp >= 0 ? (double)J[dp] : 0.0;
    if (entry != 0.0)
        P[prowptr[row]] = (VP)(1.0 / entry);
    else
        atomicOr(err, KERR_ZERO_DIAG);
}

template <class VJ>
static void schwarz_update_t(Schwarz& s, const VJ* J) {
    Model& m = *s.model;
    const hipStream_t st = m.stream;
    if (s.nsub > 0) {
        const int64_t gmax = std::max(s.small_max_dim, s.large_max_dim);
        const unsigned chunks = (unsigned)std::max<int64_t>(1, (gmax * gmax + SG_CHUNK - 1) / SG_CHUNK);
        k_schwarz_gather<VJ><<<dim3((unsigned)s.nsub, chunks), dim3(256), 0, st>>>(
            s.scatter.d_m.get(), s.scatter.d_val_off.get(), s.scatter.d_slot_off.get(), s.d_off.get(), s.d_nodes.get(),
            s.d_jslot.get(), m.pattern.d_rowptr.get(), J, s.d_blocks.get());
        check_launch("k_schwarz_gather");
        const int dim = std::max(3, s.small_max_dim);
        // Gauss-Jordan fast path (blocks up to 96 x 96; it marks the larger ones 4), the sharper
        // certificate for the blocks it could not accept, then the eigen path for the rest; blocks
        // above 96 x 96 go through the blocked path
        // blocks up to 64 x 64 on the matrix cores (Hex20 cylinder: 93.3 vs 100.3 us per update); larger
        // ones by the register Gauss-Jordan (Hex27 cylinder's 81 x 81: 869 vs 938 us on the matrix cores,
        // whose LDS round trips and serial diagonal tiles dominate at that size)
        if (dim <= 64) {
            const size_t lds = (size_t)64 * 65 * sizeof(double);
            raise_lds_limit((const void*)k_spd_inverse_mf<4>, lds);
            hipLaunchKernelGGL(k_spd_inverse_mf<4>, dim3((unsigned)s.nsub), dim3(EIG_THREADS), lds, st,
                               s.scatter.d_m.get(), s.scatter.d_val_off.get(), s.d_blocks.get(), s.d_fallback.get());
            check_launch("k_spd_inverse_mf");
        } else {
            hipLaunchKernelGGL(k_spd_inverse<6>, dim3((unsigned)s.nsub), dim3(EIG_THREADS), 0, st, s.scatter.d_m.get(),
                               s.scatter.d_val_off.get(), s.d_blocks.get(), s.d_fallback.get());
            check_launch("k_spd_inverse");
        }
        const int dp = (dim + 3) & ~3;
        const int ld_pc = dp + 1;
        const size_t lds_pc = (size_t)2 * dp * ld_pc * sizeof(double);  // 149 KB at dim 96
        raise_lds_limit((const void*)k_spd_power_certify<VJ>, lds_pc);
        hipLaunchKernelGGL(k_spd_power_certify<VJ>, dim3((unsigned)s.nsub), dim3(EIG_THREADS), lds_pc, st,
                           s.scatter.d_m.get(), s.scatter.d_val_off.get(), s.scatter.d_slot_off.get(), s.d_off.get(),
                           s.d_nodes.get(), s.d_jslot.get(), m.pattern.d_rowptr.get(), J, s.d_blocks.get(),
                           s.d_fallback.get(), ld_pc);
        check_launch("k_spd_power_certify");
        const int nn = dim + (dim & 1);
        const int ld = nn + 1;
        const size_t lds = (size_t)2 * nn * ld * sizeof(double) + (size_t)nn * sizeof(double) + (size_t)nn * sizeof(int);
        // at least the size of this launch (147 KB for 96 x 96 blocks), raised monotonically
        raise_lds_limit((const void*)k_jacobi_inverse, lds);
        hipLaunchKernelGGL(k_jacobi_inverse, dim3((unsigned)s.nsub), dim3(EIG_THREADS), lds, st, s.scatter.d_m.get(),
                           s.scatter.d_val_off.get(), s.d_blocks.get(), (const int*)s.d_fallback.get(), ld, 0);
        check_launch("k_jacobi_inverse");
        if (!s.large.empty()) schwarz_update_large(s);
        if (s.fp32)
            launch_block_gather(s.P, s.scatter, s.d_blocks.get(), s.d_vals32.get(), 0, 0.0, nullptr, true, st);
        else
            launch_block_gather(s.P, s.scatter, s.d_blocks.get(), s.d_vals.get(), 0, 0.0, nullptr, true, st);
    }
    if (!s.untreated.empty()) {
        const int64_t n = 3 * (int64_t)s.untreated.size();
        const unsigned g = (unsigned)((n + 255) / 256);
        if (s.fp32)
            k_schwarz_untreated<VJ, float><<<dim3(g), dim3(256), 0, st>>>(
                (int64_t)s.untreated.size(), s.d_untreated.get(), s.P.d_rowptr.get(), m.pattern.d_diag_pos.get(), J,
                s.d_vals32.get(), m.d_err.get());
        else
            k_schwarz_untreated<VJ, double><<<dim3(g), dim3(256), 0, st>>>(
                (int64_t)s.untreated.size(), s.d_untreated.get(), s.P.d_rowptr.get(), m.pattern.d_diag_pos.get(), J,
                s.d_vals.get(), m.d_err.get());
        check_launch("k_schwarz_untreated");
    }
}

// DefaultSemidefiniteProjection of every element matrix in m.d_ke (row-major (3n)^2 per element),
// in place: a Cholesky certificate keeps the blocks that are positive semi-definite to within
// rounding, the Newton-Schulz sign iteration (one work-group per element) projects the rest.
void launch_element_psd_projection(Model& m) {
    if (m.ne == 0) return;
    if (m.d_elem_m.n != (size_t)m.ne || m.d_elem_m.p == nullptr) {
        std::vector<int32_t> msz((size_t)m.ne, m.n);
        std::vector<int64_t> off((size_t)m.ne);
        for (int64_t e = 0; e < m.ne; ++e) off[e] = e * 9 * (int64_t)m.n * m.n;
        m.d_elem_m.upload(msz, m.stream);
        m.d_elem_off.upload(off, m.stream);
    }
    const int dim = 3 * m.n;
    m.d_elem_need.ensure((size_t)m.ne);
    const int ld_c = dim + 1;
    const size_t lds_c = (size_t)dim * ld_c * sizeof(double);
    raise_lds_limit((const void*)k_psd_certify, 120 * 1024);
    hipLaunchKernelGGL(k_psd_certify, dim3((unsigned)m.ne), dim3(EIG_THREADS), lds_c, m.stream, m.d_elem_m.get(),
                       m.d_elem_off.get(), (const double*)m.d_ke.get(), m.d_elem_need.get(), ld_c);
    check_launch("k_psd_certify");
    const int dp = (dim + 3) & ~3;
    const int ld_ns = dp + 1;
    const size_t lds_ns = (size_t)2 * dp * ld_ns * sizeof(double);
    if (lds_ns > 120 * 1024 || (dp / 4) * (dp / 4) > 2 * EIG_THREADS)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "element matrix too large for the projection");
    raise_lds_limit((const void*)k_psd_newton_schulz, 120 * 1024);
    hipLaunchKernelGGL(k_psd_newton_schulz, dim3((unsigned)m.ne), dim3(EIG_THREADS), lds_ns, m.stream,
                       m.d_elem_m.get(), m.d_elem_off.get(), m.d_ke.get(), (const int*)m.d_elem_need.get(), ld_ns);
    check_launch("k_psd_newton_schulz");
}

void schwarz_update(Schwarz& s, const double* J) { schwarz_update_t<double>(s, J); }
void schwarz_update(Schwarz& s, const float* J) { schwarz_update_t<float>(s, J); }

void schwarz_apply(Schwarz& s, const double* x, double* y) {
    if (s.fp32)
        launch_spmv(s.P, s.d_vals32.get(), 0.0, y, 1.0, x, s.model->stream);
    else
        launch_spmv(s.P, s.d_vals.get(), 0.0, y, 1.0, x, s.model->stream);
}

}  // namespace hoefem
