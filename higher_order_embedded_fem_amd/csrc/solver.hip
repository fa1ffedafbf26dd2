// Preconditioned conjugate gradient (fenris/src/cg.rs:377-482) as two fused kernels per
// iteration, with the iteration state (zTr, alpha, beta, status) kept on the device:
//   A-kernel: [sum zTr, |r|^2 of the previous B-kernel; convergence / max-iteration status;
//             beta] p <- z + beta p (on the fly, double-buffered), Ap = A p, partial pAp
//   B-kernel: [sum pAp; indefiniteness status; alpha] x += alpha p, r <- r - alpha Ap
//             (on the fly, double-buffered), z = P r, partial zTr and |r|^2
// Every kernel runs a persistent grid (at most 1024 work-groups of 256 threads, the A-kernel 512 of
// 512) whose node ranges follow the XCDs, and writes one partial sum per work-group; the next
// kernel's work-groups each add those partials up again in the same fixed order, so all of them
// agree on the scalars bit for bit and the result is deterministic. No atomics or device-scope
// fences are needed: the kernel boundary orders the partials. The host keeps two batches of
// iterations (captured HIP graphs) in flight; the kernel that ends the solve writes its status into
// pinned host memory, which the host reads after each batch's event. Also: the residual / state
// kernels of the Backward-Euler Newton driver.
#include <algorithm>
#include <cmath>
#include <mutex>

#include <hip/hip_ext.h>

#include "internal.hpp"

namespace hoefem {

enum : int { CG_RUNNING = 0, CG_CONVERGED = 100, CG_ZERO_RHS = 101 };
constexpr int CG_THREADS = 256;
// the A-kernel runs 512-thread work-groups: half as many partial sums for the B-kernel to re-read
// (measured: 18.5 vs 19.1 us per launch on the Hex20 cylinder; the B-kernel is faster at 256)
constexpr int CG_THREADS_A = 512;
constexpr int CG_BATCH = 8;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// tot[k] = sum over b < nb of partials[b * NV + k], the same value in every thread of every
// block (fixed per-thread strides, fixed wave and cross-wave summation order), in two halves so
// that the loads can be issued early and the sums formed once they are needed. issue() loads the
// first CG_MAX_BLOCKS partials (every grid of the solver) by an unrolled loop with no branch
// (slots past nb re-read partial 0 and are not added): all of a thread's loads are in flight
// together, where a rolled or branching loop waits for each load before the next. total() adds
// them (then any beyond CG_MAX_BLOCKS) in the strided order and reduces over the work-group.
template <int NV, int BT = CG_THREADS>
struct Partials {
    static constexpr int PER = (CG_MAX_BLOCKS + BT - 1) / BT;
    double v[PER][NV];
    const double* src;
    int nb;

    __device__ inline void issue(const double* __restrict__ partials, int n) {
        src = partials;
        nb = n;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int b = (int)threadIdx.x + i * BT;
            const int64_t bc = b < n ? b : 0;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[i][k] = partials[bc * NV + k];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the loads together, ahead of any use
    }

    __device__ inline void total(double (&tot)[NV]) const {
        __builtin_amdgcn_sched_barrier(0);  // loads placed before this call are issued before any add
        double acc[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] = 0.0;
#pragma unroll
        for (int i = 0; i < PER; ++i)
            if ((int)threadIdx.x + i * BT < nb)
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[k] += v[i][k];
        for (int b = (int)threadIdx.x + PER * BT; b < nb; b += BT)
#pragma unroll
            for (int k = 0; k < NV; ++k) acc[k] += src[(int64_t)b * NV + k];
        __shared__ double s_w[BT / 64][NV];
        const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double s = wave_sum(acc[k]);
            if (lane == 0) s_w[wid][k] = s;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < BT / 64; ++w) s += s_w[w][k];
            tot[k] = s;
        }
    }
};

template <int NV, int BT = CG_THREADS>
__device__ inline void sum_partials(const double* __restrict__ partials, int nb, double (&tot)[NV]) {
    Partials<NV, BT> pt;
    pt.issue(partials, nb);
    pt.total(tot);
}

// partials[bid * NV + k] = sum of v[k] over the block's threads. BT is the block size the kernel is
// launched with (its launch bound, a multiple of 64): one LDS slot per wave.
template <int NV, int BT = CG_THREADS>
__device__ inline void block_partials(const double (&v)[NV], double* __restrict__ partials, int64_t bid) {
    static_assert(BT % 64 == 0 && BT <= 1024, "block size must be a multiple of 64, at most 1024");
    __shared__ double s_w[BT / 64][NV];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    constexpr int nw = BT / 64;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) s_w[wid][k] = s;
    }
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += s_w[w][k];
            partials[bid * NV + k] = s;
        }
    }
}

template <class VT>
struct Csr3 {  // node-blocked CSR view
    int64_t nn;
    const int32_t* nptr;
    const int32_t* ncol;
    const int32_t* rowptr;
    const uint8_t* diag_only;
    const VT* vals;
    const NodeEntry* ent;  // block nodes [0, n_block), diagonal-only nodes after them
    int64_t n_block, n_diag;
};

// acc_i = sum_j A(3v+i, 3b+c) * x_c(b) over the node row block of v (G lanes, reduced into
// all lanes); xnode(b, x0, x1, x2) yields the three components of node b. Products and
// accumulation in the matrix value type VT (FP64, or FP32 for the mixed-precision operator);
// the result is returned in FP64.
template <int G, class VT, class XN>
__device__ inline void row_block_dot(const Csr3<VT>& A, int64_t v, int lane, bool valid, XN xnode, double& r0,
                                     double& r1, double& r2) {
    VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
    if (valid) {
        if (A.diag_only[v]) {
            if (lane == 0) {
                double x0, x1, x2;
                xnode(v, x0, x1, x2);
                a0 = A.vals[A.rowptr[3 * v]] * VT(x0);
                a1 = A.vals[A.rowptr[3 * v + 1]] * VT(x1);
                a2 = A.vals[A.rowptr[3 * v + 2]] * VT(x2);
            }
        } else {
            const int kb = A.nptr[v], k = A.nptr[v + 1] - kb;
            const VT* R0 = A.vals + A.rowptr[3 * v];
            const VT* R1 = R0 + 3 * k;
            const VT* R2 = R1 + 3 * k;
            for (int j = lane; j < k; j += G) {
                const int64_t b = A.ncol[kb + j];
                double y0, y1, y2;
                xnode(b, y0, y1, y2);
                const VT x0 = VT(y0), x1 = VT(y1), x2 = VT(y2);
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
    r0 = double(a0);
    r1 = double(a1);
    r2 = double(a2);
}

// a_i += sum over the lane's blocks j = j0, j0 + G, ... < k of R_i[3j..3j+2] . x(nc[j]); two blocks
// per step with all their loads issued before the products, so a lane with two blocks waits for one
// index -> gather chain instead of two.
template <int G, class VT, class XN>
__device__ inline void row_blocks(const VT* __restrict__ R0, const VT* __restrict__ R1, const VT* __restrict__ R2,
                                  const int32_t* __restrict__ nc, int j0, int k, XN xnode, VT& a0, VT& a1, VT& a2) {
    int j = j0;
    for (; j + G < k; j += 2 * G) {
        const int32_t c0 = nc[j], c1 = nc[j + G];
        VT u[9], w[9];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u[c] = R0[3 * j + c];
            u[3 + c] = R1[3 * j + c];
            u[6 + c] = R2[3 * j + c];
            w[c] = R0[3 * (j + G) + c];
            w[3 + c] = R1[3 * (j + G) + c];
            w[6 + c] = R2[3 * (j + G) + c];
        }
        double y0, y1, y2, z0, z1, z2;
        xnode((int64_t)c0, y0, y1, y2);
        xnode((int64_t)c1, z0, z1, z2);
        const VT x0 = VT(y0), x1 = VT(y1), x2 = VT(y2), v0 = VT(z0), v1 = VT(z1), v2 = VT(z2);
        a0 += u[0] * x0 + u[1] * x1 + u[2] * x2 + (w[0] * v0 + w[1] * v1 + w[2] * v2);
        a1 += u[3] * x0 + u[4] * x1 + u[5] * x2 + (w[3] * v0 + w[4] * v1 + w[5] * v2);
        a2 += u[6] * x0 + u[7] * x1 + u[8] * x2 + (w[6] * v0 + w[7] * v1 + w[8] * v2);
    }
    if (j < k) {
        double y0, y1, y2;
        xnode((int64_t)nc[j], y0, y1, y2);
        const VT x0 = VT(y0), x1 = VT(y1), x2 = VT(y2);
        a0 += R0[3 * j] * x0 + R0[3 * j + 1] * x1 + R0[3 * j + 2] * x2;
        a1 += R1[3 * j] * x0 + R1[3 * j + 1] * x1 + R1[3 * j + 2] * x2;
        a2 += R2[3 * j] * x0 + R2[3 * j + 1] * x1 + R2[3 * j + 2] * x2;
    }
}

// acc_i = sum over the neighbours b of entry e of A(3v+i, 3b+c) x_c(b) (G lanes, reduced into
// all lanes of the group). Products and accumulation in VT; the result is returned in FP64.
template <int G, class VT, class XN>
__device__ inline void entry_row_dot(const Csr3<VT>& A, const NodeEntry& e, int lane, bool valid, XN xnode,
                                     double& r0, double& r1, double& r2) {
    VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
    if (valid) {
        const int k = e.count;
        const VT* R0 = A.vals + e.rowoff;
        const VT* R1 = R0 + 3 * k;
        const VT* R2 = R1 + 3 * k;
        const int32_t* nc = A.ncol + e.first;
        row_blocks<G>(R0, R1, R2, nc, lane, k, xnode, a0, a1, a2);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, G);
        a1 += __shfl_xor(a1, off, G);
        a2 += __shfl_xor(a2, off, G);
    }
    r0 = double(a0);
    r1 = double(a1);
    r2 = double(a2);
}

// Group ranges of a work-group. Work-groups are dispatched to the 8 XCDs round robin (work-group b
// on XCD b mod 8), and each XCD has its own L2. With the grid a multiple of 8, XCD x walks the
// x-th contiguous eighth of the groups with its own work-groups in a grid-stride loop, so its L2
// holds that slab's gathered vector entries (plus a halo) instead of a share of every node's, and
// each XCD still streams through one compact window.
struct GroupRange {
    int64_t begin, end, step;
};
__device__ inline GroupRange group_range(int64_t groups, int64_t b, int64_t n) {
    if (n % 8) return {b, groups, n};
    const int64_t x = b % 8;
    return {x * groups / 8 + b / 8, (x + 1) * groups / 8, n / 8};
}

// Loops over the entries of a pattern: block nodes in groups of BT / G (G lanes each;
// block(e, lane, valid) is called the same number of times by every thread of a work-group, as
// the G-lane shuffles need whole groups) over the work-group's group_range, the next group's
// entry prefetched; then diagonal-only nodes one per thread (diag(e)). Entries are read at a
// clamped index with no branch (an entry read for an invalid slot is not used), so no entry load
// waits where it is issued: the first one is in flight while pro() runs (the kernel's prologue:
// its own loads, status tests, reductions; false stops the kernel, uniformly over the grid).
template <int G, int BT = CG_THREADS, class FP, class FB, class FD>
__device__ inline bool for_entries(const NodeEntry* __restrict__ ent, int64_t nb, int64_t nd, FP pro, FB block,
                                   FD diag, int64_t bid, int64_t nblk) {
    constexpr int per = BT / G;
    const int lane = (int)(threadIdx.x % G);
    const int64_t groups = (nb + per - 1) / per;
    const int64_t sub = threadIdx.x / G;
    const GroupRange R = group_range(groups, bid, nblk);
    const int64_t last = nb + nd - 1;
    if (last < 0) return pro();  // an empty pattern
    auto load = [&](int64_t g) {
        const int64_t i = g * per + sub;
        return ent[i < last ? i : last];
    };
    int64_t g = R.begin;
    NodeEntry e = load(g);
    if (!pro()) {
        // the first entry is used on this path too (an empty asm reads it), so the compiler keeps
        // its load ahead of pro()'s status test instead of sinking it past the test
        asm volatile("" ::"v"(e.count));
        return false;
    }
    for (; g < R.end; g += R.step) {
        const NodeEntry cur = e;
        const bool valid = g * per + sub < nb;
        e = load(g + R.step);
        block(cur, lane, valid);
    }
    // diagonal-only nodes from the last work-group backwards: those have the fewest block groups
    const int64_t rb = nblk - 1 - bid;
    for (int64_t i = rb * BT + threadIdx.x; i < nd; i += nblk * BT) diag(ent[nb + i]);
    return true;
}

// The lane's first 3x3 block of a row (j = lane) and its column, loaded a pass ahead.
template <class VT>
struct RowPf {
    VT a[9];
    int32_t col;
};

template <class VT>
__device__ inline void row_pf_issue(const Csr3<VT>& A, const NodeEntry& e, int lane, bool valid, RowPf<VT>& pf) {
    if (valid && lane < e.count) {
        const int k = e.count;
        const VT* R0 = A.vals + e.rowoff + 3 * lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pf.a[c] = R0[c];
            pf.a[3 + c] = R0[3 * k + c];
            pf.a[6 + c] = R0[6 * k + c];
        }
        pf.col = A.ncol[e.first + lane];
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) pf.a[c] = VT(0);
        pf.col = -1;
    }
}

// entry_row_dot with the first block of each lane taken from pf.
template <int G, class VT, class XN>
__device__ inline void entry_row_dot_pf(const Csr3<VT>& A, const NodeEntry& e, int lane, bool valid,
                                        const RowPf<VT>& pf, XN xnode, double& r0, double& r1, double& r2) {
    VT a0 = VT(0), a1 = VT(0), a2 = VT(0);
    if (valid && pf.col >= 0) {
        double y0, y1, y2;
        xnode((int64_t)pf.col, y0, y1, y2);
        const VT x0 = VT(y0), x1 = VT(y1), x2 = VT(y2);
        a0 = pf.a[0] * x0 + pf.a[1] * x1 + pf.a[2] * x2;
        a1 = pf.a[3] * x0 + pf.a[4] * x1 + pf.a[5] * x2;
        a2 = pf.a[6] * x0 + pf.a[7] * x1 + pf.a[8] * x2;
        const int k = e.count;
        const VT* R0 = A.vals + e.rowoff;
        const VT* R1 = R0 + 3 * k;
        const VT* R2 = R1 + 3 * k;
        const int32_t* nc = A.ncol + e.first;
        for (int j = lane + G; j < k; j += G) {
            xnode((int64_t)nc[j], y0, y1, y2);
            const VT z0 = VT(y0), z1 = VT(y1), z2 = VT(y2);
            a0 += R0[3 * j] * z0 + R0[3 * j + 1] * z1 + R0[3 * j + 2] * z2;
            a1 += R1[3 * j] * z0 + R1[3 * j + 1] * z1 + R1[3 * j + 2] * z2;
            a2 += R2[3 * j] * z0 + R2[3 * j + 1] * z1 + R2[3 * j + 2] * z2;
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, G);
        a1 += __shfl_xor(a1, off, G);
        a2 += __shfl_xor(a2, off, G);
    }
    r0 = double(a0);
    r1 = double(a1);
    r2 = double(a2);
}

// for_entries with the memory latency of the iteration's start taken off the critical path:
// the first entry is loaded before pro_issue() (loads that need no entry, e.g. the partial
// sums), the first pass's row data right after it, and pro_finish() (status tests, reductions;
// returns false to stop, uniformly over the grid) runs while those are in flight. In the loop
// the row data of pass n + 1 and the entry of pass n + 2 are loaded before pass n is computed.
template <int G, class VT, class PI, class PF, class FB, class FD>
__device__ inline bool for_entries_pf(const Csr3<VT>& A, PI pro_issue, PF pro_finish, FB block, FD diag, int64_t bid,
                                      int64_t nblk) {
    constexpr int per = CG_THREADS / G;
    const int lane = (int)(threadIdx.x % G);
    const int64_t nb = A.n_block, nd = A.n_diag;
    const int64_t groups = (nb + per - 1) / per;
    const int64_t sub = threadIdx.x / G;
    const GroupRange R = group_range(groups, bid, nblk);
    const int64_t stride = R.step;
    auto has = [&](int64_t g) { return g < R.end && g * per + sub < nb; };
    const int64_t last = nb + nd - 1;
    if (last < 0) {  // an empty pattern
        pro_issue();
        return pro_finish();
    }
    auto load = [&](int64_t g) {  // clamped, with no branch, as in for_entries
        const int64_t i = g * per + sub;
        return A.ent[i < last ? i : last];
    };
    int64_t g = R.begin;
    NodeEntry e = load(g);
    pro_issue();
    RowPf<VT> pf;
    row_pf_issue(A, e, lane, has(g), pf);
    NodeEntry e1 = load(g + stride);
    if (!pro_finish()) return false;
    for (; g < R.end; g += stride) {
        const NodeEntry cur = e;
        const bool valid = has(g);
        RowPf<VT> pn;
        row_pf_issue(A, e1, lane, has(g + stride), pn);
        e = e1;
        e1 = load(g + 2 * stride);
        block(cur, lane, valid, pf);
        pf = pn;
    }
    const int64_t rb = nblk - 1 - bid;  // as in for_entries
    for (int64_t i = rb * CG_THREADS + threadIdx.x; i < nd; i += nblk * CG_THREADS)
        diag(A.ent[nb + i]);
    return true;
}

inline unsigned cg_grid(int64_t n_block, int64_t n_diag, int G, int bt = CG_THREADS) {
    const int64_t groups = (n_block + bt / G - 1) / (bt / G) + (n_diag + bt - 1) / bt;
    // 4 waves per SIMD: 1024 work-groups of 256 threads or 512 of 512
    const int64_t cap = (int64_t)CG_MAX_BLOCKS * CG_THREADS / bt;
    int64_t grid = std::max<int64_t>(1, std::min<int64_t>(groups, cap));
    if (grid >= 64) grid -= grid % 8;  // a multiple of the 8 XCDs (group_range)
    return (unsigned)grid;
}

// The two vectors each SpMV gathers are stored interleaved per node, six doubles (48 B, three
// 16-byte loads): zp = [z | p] and ra = [r | Ap], each double-buffered. Iteration k (c = k & 1):
//   A_k gathers zp[c] = (z_k, p_{k-1}), writes p_k -> zp[1-c].p and Ap_k -> ra[c].a
//   B_k gathers ra[c] = (r_k, Ap_k), reads p_k from zp[1-c], writes r_{k+1} -> ra[1-c].r and
//       z_{k+1} -> zp[1-c].z
__device__ inline void load6(const double* __restrict__ base, int64_t b, double (&q)[6]) {
    const double2* s = reinterpret_cast<const double2*>(base + 6 * b);
    const double2 u0 = s[0], u1 = s[1], u2 = s[2];
    q[0] = u0.x;
    q[1] = u0.y;
    q[2] = u1.x;
    q[3] = u1.y;
    q[4] = u2.x;
    q[5] = u2.y;
}

// One linear system of a (possibly batched) CG launch, in device memory: its matrices, vectors,
// partial sums and iteration state, and where its work-groups sit in the launches. A single solve
// is a batch of one segment; a batch of independent systems (independent scenes) runs every
// iteration of all of them in one A and one B launch, each work-group working on its own segment
// exactly as it would alone (same local index, same local grid), so the results are bit-identical.
template <class VT>
struct CgSeg {
    Csr3<VT> A, Pm;      // system matrix; preconditioner (the Jacobi variant reads dinv instead)
    double* x;
    const double* b;
    const double* dinv;
    double* zp[2];       // interleaved [z | p], double-buffered
    double* ra[2];       // interleaved [r | Ap], double-buffered
    double* part_pre;
    double* part_a;
    double* part_b;
    CgScalars* sc;
    int* host_status;    // pinned host copy of sc->status (or any device word)
    int nb_a, nb_b;      // work-groups of the segment in the A-side (preamble 1, A) and B-side launches
    int base_a, base_b;  // its first work-group in those launches
    int nparts_b;        // partial sums the B side writes (nb_b, or the aggregates of the learned apply)
    int pad;
    double tol;          // relative tolerance and iteration limit (read by the batched start kernel)
    int64_t max_iter;
};

struct CgPtrs {
    double* x;
    const double* zp_old;  // A_k: zp[c]                       (pre2: unused)
    double* zp_new;        // A_k and B_k: zp[1-c]              (pre2: zp[0])
    double* ap_out;        // A_k: ra[c] (its Ap half is written)
    const double* ra_old;  // B_k: ra[c]                        (pre2: ra[0])
    double* ra_new;        // B_k: ra[1-c]                      (pre1: ra[0])
    const double* b;
    const double* dinv;
    double* part_pre;
    double* part_a;
    double* part_b;
    int nb_a;  // grid of the preamble-1 / A-kernels
    int nb_b;  // grid of the preamble-2 / B-kernels
    CgScalars* sc;
    int* host_status;  // pinned host copy of sc->status, written when a kernel ends the solve
};

// The kernel view of segment S: stage 0 preamble 1, 1 preamble 2, 2 the A and B kernels of an
// iteration whose buffer parity is cur.
template <class VT>
__device__ inline CgPtrs seg_ptrs(const CgSeg<VT>& S, int stage, int cur) {
    CgPtrs P{};
    P.x = S.x;
    P.b = S.b;
    P.dinv = S.dinv;
    P.part_pre = S.part_pre;
    P.part_a = S.part_a;
    P.part_b = S.part_b;
    P.nb_a = S.nb_a;
    P.nb_b = S.nparts_b;
    P.sc = S.sc;
    P.host_status = S.host_status;
    if (stage == 0) {
        P.ra_new = S.ra[0];
    } else if (stage == 1) {
        P.ra_old = S.ra[0];
        P.zp_new = S.zp[0];
    } else {
        P.zp_old = S.zp[cur];
        P.zp_new = S.zp[1 - cur];
        P.ap_out = S.ra[cur];
        P.ra_old = S.ra[cur];
        P.ra_new = S.ra[1 - cur];
    }
    return P;
}

// the segment of this work-group and its index / count within the segment's A-side (a) or B-side launch
struct SegIdx {
    int s;
    int64_t bid, nblk;
};
// (s = -1: a padding work-group between two segments of a batched launch, which returns at once).
// A single solve (SEG false) takes its one segment by value, in the kernel arguments: read by the
// dispatch with the launch, no dependent load of a descriptor from memory at every kernel's start.
template <bool SEG, class VT>
__device__ inline SegIdx seg_index(const CgSeg<VT>& one, const CgSeg<VT>* segs, const int* seg_of, bool a_side) {
    if (!SEG) return a_side ? SegIdx{0, (int64_t)blockIdx.x, one.nb_a} : SegIdx{0, (int64_t)blockIdx.x, one.nb_b};
    const int s = seg_of[blockIdx.x];
    if (s < 0) return SegIdx{-1, 0, 0};
    const CgSeg<VT>& S = segs[s];
    return a_side ? SegIdx{s, (int64_t)blockIdx.x - S.base_a, S.nb_a}
                  : SegIdx{s, (int64_t)blockIdx.x - S.base_b, S.nb_b};
}

// preamble 1: r = b - A x -> ra_new.r ; partial |b|^2
template <int G, class VT, bool SEG>
__global__ __launch_bounds__(CG_THREADS) void k_cg_pre1(const CgSeg<VT> one, const CgSeg<VT>* __restrict__ segs,
                                                        const int* __restrict__ seg_of) {
    const SegIdx I = seg_index<SEG>(one, segs, seg_of, true);
    if (I.s < 0) return;
    const CgSeg<VT>& S = SEG ? segs[I.s] : one;
    if (S.sc->status != CG_RUNNING) return;  // a segment without a solve in this batch
    const Csr3<VT> A = S.A;
    const CgPtrs P = seg_ptrs(S, 0, 0);
    const double* x = P.x;
    double part[1] = {0.0};
    auto finish = [&](int64_t v, double a0, double a1, double a2) {
        const double b0 = P.b[3 * v], b1 = P.b[3 * v + 1], b2 = P.b[3 * v + 2];
        P.ra_new[6 * v] = b0 - a0;
        P.ra_new[6 * v + 1] = b1 - a1;
        P.ra_new[6 * v + 2] = b2 - a2;
        part[0] += b0 * b0 + b1 * b1 + b2 * b2;
    };
    for_entries<G>(
        A.ent, A.n_block, A.n_diag, [] { return true; },
        [&](const NodeEntry& e, int lane, bool valid) {
            double a0, a1, a2;
            entry_row_dot<G>(A, e, lane, valid, [&](int64_t b, double& x0, double& x1, double& x2) {
                x0 = x[3 * b];
                x1 = x[3 * b + 1];
                x2 = x[3 * b + 2];
            }, a0, a1, a2);
            if (valid && lane == 0) finish(e.node, a0, a1, a2);
        },
        [&](const NodeEntry& e) {
            const int64_t v = e.node;
            const VT* d = A.vals + e.rowoff;
            finish(v, double(d[0] * VT(x[3 * v])), double(d[1] * VT(x[3 * v + 1])), double(d[2] * VT(x[3 * v + 2])));
        },
        I.bid, I.nblk);
    block_partials<1>(part, P.part_pre, I.bid);
}

// preamble 2: |b| ; z = P r (r from ra_old) -> zp_new.z ; partial zTr, |r|^2
template <int G, bool JAC, bool FLEX, class VT, bool SEG>
__global__ __launch_bounds__(CG_THREADS) void k_cg_pre2(const CgSeg<VT> one, const CgSeg<VT>* __restrict__ segs,
                                                        const int* __restrict__ seg_of) {
    const SegIdx I = seg_index<SEG>(one, segs, seg_of, false);
    if (I.s < 0) return;
    const CgSeg<VT>& S = SEG ? segs[I.s] : one;
    if (S.sc->status != CG_RUNNING) return;
    const Csr3<VT> Pm = S.Pm;
    const CgPtrs P = seg_ptrs(S, 1, 0);
    double bb[1];
    sum_partials<1>(P.part_pre, P.nb_a, bb);
    if (I.bid == 0 && threadIdx.x == 0) P.sc->bnorm = sqrt(bb[0]);
    const double* ra = P.ra_old;
    double part[2] = {0.0, 0.0};
    auto finish = [&](int64_t v, double z0, double z1, double z2) {
        const double r0 = ra[6 * v], r1 = ra[6 * v + 1], r2 = ra[6 * v + 2];
        if (JAC) {
            z0 = P.dinv[3 * v] * r0;
            z1 = P.dinv[3 * v + 1] * r1;
            z2 = P.dinv[3 * v + 2] * r2;
        }
        P.zp_new[6 * v] = z0;
        P.zp_new[6 * v + 1] = z1;
        P.zp_new[6 * v + 2] = z2;
        part[0] += z0 * r0 + z1 * r1 + z2 * r2;
        part[1] += r0 * r0 + r1 * r1 + r2 * r2;
    };
    if (JAC) {
        for (int64_t v = I.bid * CG_THREADS + threadIdx.x; v < S.A.nn; v += I.nblk * CG_THREADS) finish(v, 0.0, 0.0, 0.0);
    } else {
        for_entries<G>(
            Pm.ent, Pm.n_block, Pm.n_diag, [] { return true; },
            [&](const NodeEntry& e, int lane, bool valid) {
                double z0, z1, z2;
                entry_row_dot<G>(Pm, e, lane, valid, [&](int64_t b, double& x0, double& x1, double& x2) {
                    x0 = ra[6 * b];
                    x1 = ra[6 * b + 1];
                    x2 = ra[6 * b + 2];
                }, z0, z1, z2);
                if (valid && lane == 0) finish(e.node, z0, z1, z2);
            },
            [&](const NodeEntry& e) {
                const int64_t v = e.node;
                const VT* d = Pm.vals + e.rowoff;
                finish(v, double(d[0] * VT(ra[6 * v])), double(d[1] * VT(ra[6 * v + 1])),
                       double(d[2] * VT(ra[6 * v + 2])));
            },
            I.bid, I.nblk);
    }
    if (!FLEX) block_partials<2>(part, P.part_b, I.bid);
}

// A-kernel of iteration k: stopping tests on r_k, beta_k = zTr_k / zTr_{k-1},
// p_k = z_k + beta p_{k-1} ; Ap_k = A p_k ; partial pAp. The partial sums are reduced before the
// first entry is issued: issuing the entry early (as the B-kernel does) measured 0.25 us slower
// per launch with 9 MB more HBM fetch on the Hex20 cylinder, and loading the entries two passes and
// the first column indices one pass ahead 1.0 us slower (19.3 vs 18.3 us, 126 VGPRs).
template <int G, bool FLEX, class VT, bool SEG>
__global__ __launch_bounds__(CG_THREADS_A) void k_cg_a(const CgSeg<VT> one, const CgSeg<VT>* __restrict__ segs,
                                                       const int* __restrict__ seg_of, int cur) {
    const SegIdx I = seg_index<SEG>(one, segs, seg_of, true);
    if (I.s < 0) return;
    const CgSeg<VT>& S = SEG ? segs[I.s] : one;
    CgScalars* sc = S.sc;
    const Csr3<VT> A = S.A;
    const CgPtrs P = seg_ptrs(S, 2, cur);
    constexpr int NV = FLEX ? 3 : 2;
    const bool lead = I.bid == 0 && threadIdx.x == 0;
    double beta = 0.0;
    // the status, the iteration index (written by the previous B-kernel, 0 after the preamble), that
    // kernel's partial sums and (for_entries) the first row entry are loaded together, not one
    // memory latency after the other
    auto pro = [&]() -> bool {
        const int status = sc->status;
        const int64_t k = sc->iters;
        Partials<NV, CG_THREADS_A> pt;
        pt.issue(P.part_b, P.nb_b);
        if (status != CG_RUNNING) return false;
        double tot[NV];
        pt.total(tot);
        // z_k.r_k, |r_k|^2 and (flexible variant) z_k.r_{k-1}
        const double zTr = tot[0], zTr_cross = FLEX ? tot[1] : 0.0, rr = tot[NV - 1];
        int decided = CG_RUNNING;
        if (sc->bnorm == 0.0)
            decided = CG_ZERO_RHS;
        else if (sqrt(rr) <= sc->tol * sc->bnorm)
            decided = CG_CONVERGED;
        else if (sc->max_iter >= 0 && k >= sc->max_iter)
            decided = HOEFEM_ERROR_CG_MAX_ITERATIONS;
        if (decided != CG_RUNNING) {
            if (lead) {
                sc->zTr = zTr;
                sc->rr = rr;
                sc->status = decided;
                *P.host_status = decided;
            }
            return false;
        }
        // Fletcher-Reeves form for a fixed preconditioner; Polak-Ribiere z_k.(r_k - r_{k-1}) when
        // the preconditioner depends on its input (Notay, SIAM J. Sci. Comput. 22 (2000) 1444)
        beta = k == 0 ? 0.0 : (FLEX ? (zTr - zTr_cross) : zTr) / sc->zrec[(k - 1) & 1];
        if (lead) {
            sc->zrec[k & 1] = zTr;
            sc->zTr = zTr;
            sc->rr = rr;
            sc->beta = beta;
            sc->kcur = k;
        }
        return true;
    };
    const double* zp = P.zp_old;
    double part[1] = {0.0};
    auto pvec = [&](int64_t b, double& x0, double& x1, double& x2) {
        double q[6];
        load6(zp, b, q);
        x0 = q[0] + beta * q[3];
        x1 = q[1] + beta * q[4];
        x2 = q[2] + beta * q[5];
    };
    auto finish = [&](int64_t v, double p0, double p1, double p2, double a0, double a1, double a2) {
        P.zp_new[6 * v + 3] = p0;
        P.zp_new[6 * v + 4] = p1;
        P.zp_new[6 * v + 5] = p2;
        P.ap_out[6 * v + 3] = a0;
        P.ap_out[6 * v + 4] = a1;
        P.ap_out[6 * v + 5] = a2;
        part[0] += p0 * a0 + p1 * a1 + p2 * a2;
    };
    const bool ran = for_entries<G, CG_THREADS_A>(
        A.ent, A.n_block, A.n_diag, pro,
        [&](const NodeEntry& e, int lane, bool valid) {
            double a0, a1, a2;
            entry_row_dot<G>(A, e, lane, valid, pvec, a0, a1, a2);
            if (valid && lane == 0) {
                double p0, p1, p2;
                pvec(e.node, p0, p1, p2);
                finish(e.node, p0, p1, p2, a0, a1, a2);
            }
        },
        [&](const NodeEntry& e) {
            const int64_t v = e.node;
            const VT* d = A.vals + e.rowoff;
            double p0, p1, p2;
            pvec(v, p0, p1, p2);
            finish(v, p0, p1, p2, double(d[0] * VT(p0)), double(d[1] * VT(p1)), double(d[2] * VT(p2)));
        },
        I.bid, I.nblk);
    if (!ran) return;
    block_partials<1, CG_THREADS_A>(part, P.part_a, I.bid);
}

// B-kernel of iteration k: indefiniteness tests, alpha = zTr / pAp ; x += alpha p_k ;
// r_{k+1} = r_k - alpha Ap_k ; z_{k+1} = P r_{k+1} ; partial zTr, |r|^2
template <int G, bool JAC, bool FLEX, class VT, bool SEG>
__global__ __launch_bounds__(CG_THREADS) void k_cg_b(const CgSeg<VT> one, const CgSeg<VT>* __restrict__ segs,
                                                     const int* __restrict__ seg_of, int cur) {
    const SegIdx I = seg_index<SEG>(one, segs, seg_of, false);
    if (I.s < 0) return;
    const CgSeg<VT>& S = SEG ? segs[I.s] : one;
    const Csr3<VT> Pm = S.Pm;
    const CgPtrs P = seg_ptrs(S, 2, cur);
    CgScalars* sc = P.sc;
    Partials<1> pt;
    double alpha = 0.0;
    const bool lead = I.bid == 0 && threadIdx.x == 0;
    int status = CG_RUNNING;
    int64_t k = 0;
    double zrec[2];
    // the partial sums, the status, the iteration index (written by this iteration's A-kernel) and
    // both zTr records are loaded together; zTr of iteration k is selected afterwards
    auto pro_issue = [&] {
        pt.issue(P.part_a, P.nb_a);
        status = sc->status;
        k = sc->kcur;
        zrec[0] = sc->zrec[0];
        zrec[1] = sc->zrec[1];
    };
    auto pro_finish = [&]() -> bool {
        if (status != CG_RUNNING) return false;
        double tot[1];
        pt.total(tot);
        const double pAp = tot[0];
        const double zTr = zrec[k & 1];
        int decided = CG_RUNNING;
        if (pAp <= 0.0)
            decided = HOEFEM_ERROR_CG_INDEFINITE_OPERATOR;
        else if (zTr <= 0.0)
            decided = HOEFEM_ERROR_CG_INDEFINITE_PRECONDITIONER;
        if (decided != CG_RUNNING) {
            if (lead) {
                sc->pAp = pAp;
                sc->status = decided;
                *P.host_status = decided;
            }
            return false;
        }
        alpha = zTr / pAp;
        if (lead) {
            sc->pAp = pAp;
            sc->alpha = alpha;
            sc->iters = k + 1;
        }
        return true;
    };
    const double* ra = P.ra_old;
    double part[2] = {0.0, 0.0};
    auto rvec = [&](int64_t b, double& x0, double& x1, double& x2) {
        double q[6];
        load6(ra, b, q);
        x0 = q[0] - alpha * q[3];
        x1 = q[1] - alpha * q[4];
        x2 = q[2] - alpha * q[5];
    };
    auto finish = [&](int64_t v, double r0, double r1, double r2, double z0, double z1, double z2) {
        if (JAC) {
            z0 = P.dinv[3 * v] * r0;
            z1 = P.dinv[3 * v + 1] * r1;
            z2 = P.dinv[3 * v + 2] * r2;
        }
        const double* pk = P.zp_new + 6 * v + 3;
        P.x[3 * v] = P.x[3 * v] + alpha * pk[0];
        P.x[3 * v + 1] = P.x[3 * v + 1] + alpha * pk[1];
        P.x[3 * v + 2] = P.x[3 * v + 2] + alpha * pk[2];
        P.ra_new[6 * v] = r0;
        P.ra_new[6 * v + 1] = r1;
        P.ra_new[6 * v + 2] = r2;
        P.zp_new[6 * v] = z0;
        P.zp_new[6 * v + 1] = z1;
        P.zp_new[6 * v + 2] = z2;
        part[0] += z0 * r0 + z1 * r1 + z2 * r2;
        part[1] += r0 * r0 + r1 * r1 + r2 * r2;
    };
    if (JAC) {
        pro_issue();
        if (!pro_finish()) return;
        for (int64_t v = I.bid * CG_THREADS + threadIdx.x; v < S.A.nn; v += I.nblk * CG_THREADS) {
            double r0, r1, r2;
            rvec(v, r0, r1, r2);
            finish(v, r0, r1, r2, 0.0, 0.0, 0.0);
        }
    } else {
        const bool ran = for_entries_pf<G>(
            Pm, pro_issue, pro_finish,
            [&](const NodeEntry& e, int lane, bool valid, const RowPf<VT>& pf) {
                // the row's own r_k, Ap_k, p_k and x, component c on lane c < 3 (lanes 3.. read
                // component 0 and do not use it), loaded ahead of the product: x and p_k are not
                // among the product's loads, so reading them after the reduction cost one more
                // memory latency per pass (12.7 -> 11.9 us per launch on the Hex20 cylinder)
                const int64_t v = e.node;
                const int c = lane < 3 ? lane : 0;
                const double orr = ra[6 * v + c], oap = ra[6 * v + 3 + c];
                const double opk = P.zp_new[6 * v + 3 + c], ox = P.x[3 * v + c];
                double z0, z1, z2;
                entry_row_dot_pf<G>(Pm, e, lane, valid, pf, rvec, z0, z1, z2);
                if (valid && lane < 3) {
                    const double r = orr - alpha * oap;
                    const double z = lane == 0 ? z0 : (lane == 1 ? z1 : z2);
                    P.x[3 * v + lane] = ox + alpha * opk;
                    P.ra_new[6 * v + lane] = r;
                    P.zp_new[6 * v + lane] = z;
                    part[0] += z * r;
                    part[1] += r * r;
                }
            },
            [&](const NodeEntry& e) {
                const int64_t v = e.node;
                const VT* d = Pm.vals + e.rowoff;
                double r0, r1, r2;
                rvec(v, r0, r1, r2);
                finish(v, r0, r1, r2, double(d[0] * VT(r0)), double(d[1] * VT(r1)), double(d[2] * VT(r2)));
            },
            I.bid, I.nblk);
        if (!ran) return;
    }
    if (!FLEX) block_partials<2>(part, P.part_b, I.bid);
}

// Learned coarse correction of an iteration, in four launches: k_lc_project_cg (one work-group
// per aggregate) forms the network input x_a = bf16(g / |g|), g = Phi_a^T r_new, and |g|; the network
// maps every x_a to y_a (launch_mlp_forward) and the sign guard (launch_lc_guard) removes a component
// of y_a against g_a; k_lc_apply adds z += Phi_a (kappa |g| y_a) on the aggregate's DOFs (z holds P r_new)
// and writes the partial sums z.r_new, z.r_old and |r_new|^2 of the iteration. Every DOF belongs to exactly one aggregate, so the work-groups' partials add up to
// the full dot products.
__global__ __launch_bounds__(256) void k_lc_project_cg(LearnedView lv, const double* __restrict__ ra_new) {
    __shared__ double s_r[LC_MAX_AGG_NODES * 3];
    __shared__ double s_g[LC_IN];
    const int a = blockIdx.x;
    const int o = lv.agg_ptr[a], nd = lv.agg_ptr[a + 1] - o;
    const int t = threadIdx.x;
    for (int d = t; d < nd; d += 256) {
        const int dof = lv.agg_dofs[o + d];
        s_r[d] = ra_new[6 * (dof / 3) + dof % 3];
    }
    __syncthreads();
    const float* B = lv.basis + (int64_t)LC_IN * o;
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
        const float xn = nrm > 0.0 ? (float)(s_g[t] / nrm) : 0.0f;
        lv.x[(int64_t)a * LC_IN + t] = bf16_round(xn);
        lv.xf[(int64_t)a * LC_IN + t] = xn;
        if (t == 0) lv.gn[a] = (float)nrm;
    }
}

__global__ __launch_bounds__(256) void k_lc_apply(LearnedView lv, const double* __restrict__ ra_new,
                                                  const double* __restrict__ ra_old, double* __restrict__ zp,
                                                  double* __restrict__ partials) {
    __shared__ double s_c[LC_IN];
    const int a = blockIdx.x;
    const int o = lv.agg_ptr[a], nd = lv.agg_ptr[a + 1] - o;
    const int t = threadIdx.x;
    if (t < LC_IN) s_c[t] = (double)(lv.kappa * lv.gn[a] * lv.y[(int64_t)a * LC_OUT + t]);
    __syncthreads();
    const float* B = lv.basis + (int64_t)LC_IN * o;
    double part[3] = {0.0, 0.0, 0.0};
    for (int d = t; d < nd; d += 256) {
        double corr = 0.0;
        for (int j = 0; j < LC_IN; ++j) corr += (double)B[j * nd + d] * s_c[j];
        const int dof = lv.agg_dofs[o + d];
        const int64_t zi = 6 * (int64_t)(dof / 3) + dof % 3;
        const double z = zp[zi] + corr;
        zp[zi] = z;
        const double rn = ra_new[zi];
        const double ro = ra_old ? ra_old[zi] : 0.0;
        part[0] += z * rn;
        part[1] += z * ro;
        part[2] += rn * rn;
    }
    block_partials<3>(part, partials, blockIdx.x);
}

static void launch_lc_correction(const LearnedView& lv, const double* ra_new, const double* ra_old, double* zp,
                                 double* partials, hipStream_t s) {
    k_lc_project_cg<<<dim3((unsigned)lv.na), dim3(256), 0, s>>>(lv, ra_new);
    launch_mlp_forward(lv, lv.na, lv.x, lv.y, s);
    launch_lc_guard(lv.na, lv.xf, lv.y, nullptr, s);
    k_lc_apply<<<dim3((unsigned)lv.na), dim3(256), 0, s>>>(lv, ra_new, ra_old, zp, partials);
}

__global__ void k_cg_init(CgScalars* sc, double tol, int64_t max_iter) {
    sc->tol = tol;
    sc->max_iter = max_iter;
    sc->status = CG_RUNNING;
    sc->iters = 0;
    sc->kcur = 0;
    sc->beta = 0.0;
    sc->zrec[0] = sc->zrec[1] = 0.0;
}

// DiagonalPreconditioner::update (simulation_toolbox/src/fem/newton_cg.rs:558-575)
template <class VT>
__global__ void k_jacobi_diag(int64_t ndof, const int32_t* __restrict__ diag_pos, const VT* __restrict__ A,
                              double* __restrict__ dinv, int* err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ndof) return;
    const int32_t dp = diag_pos[r];
    const double a = dp >= 0 ? (double)A[dp] : 0.0;
    if (a != 0.0)
        dinv[r] = 1.0 / a;
    else {
        dinv[r] = 0.0;
        atomicOr(err, KERR_ZERO_DIAG);
    }
}

// single block: *out = sum of nb partials (the residual norm)
__global__ __launch_bounds__(CG_THREADS) void k_sum_partials(const double* __restrict__ partials, int nb,
                                                             double* __restrict__ out) {
    double tot[1];
    sum_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) *out = tot[0];
}

void CgWork::ensure(int64_t nn_dof, int nblocks) {
    if (n != nn_dof) {
        n = nn_dof;
        for (int k = 0; k < 2; ++k) {
            zp[k].alloc(2 * n);
            ra[k].alloc(2 * n);
        }
        diag_inv.alloc(n);
    }
    if (!part_pre.p) {
        part_pre.alloc(CG_MAX_BLOCKS);
        part_a.alloc(CG_MAX_BLOCKS);
        part_b.alloc(3 * CG_MAX_BLOCKS);
    }
    if ((int64_t)partials.n < (int64_t)nblocks) partials.alloc((size_t)nblocks);
    if (!scalars.p) {
        scalars.alloc(1);
        HIP_CHECK(hipMemset(scalars.get(), 0, sizeof(CgScalars)));
    }
    if (!ev[0]) {
        HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&join[0], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&join[1], hipEventDisableTiming));
        HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    }
}

CgWork::~CgWork() {
    if (batch_exec) (void)hipGraphExecDestroy(batch_exec);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : join)
        if (e) (void)hipEventDestroy(e);
    if (cs) (void)hipStreamDestroy(cs);
}

template <class VT>
static Csr3<VT> view(const NodePattern& p, const VT* vals) {
    return Csr3<VT>{p.num_nodes, p.d_nptr.get(), p.d_ncol.get(), p.d_rowptr.get(), p.d_diag_only.get(), vals,
                    p.d_entries.get(), p.n_block, p.n_diag};
}

// xs = x, rs = r (from the interleaved [r | Ap] buffer)
__global__ void k_cg_snapshot(int64_t n, const double* __restrict__ x, const double* __restrict__ ra,
                              double* __restrict__ xs, double* __restrict__ rs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xs[i] = x[i];
    rs[i] = ra[6 * (i / 3) + i % 3];
}

template <int GA, int GB, bool JAC, bool FLEX, class VT>
static void cg_launch_all(Model& m, CgWork& w, const Csr3<VT>& A, const Csr3<VT>& Pm, const double* b, double* x,
                          double tol, int64_t max_iter, int64_t* iterations, hoefem_status* out_status,
                          CgHarvest* h, const LearnedView* lv) {
    const hipStream_t s = w.cs;
    HIP_CHECK(hipEventRecord(w.join[0], m.stream));
    HIP_CHECK(hipStreamWaitEvent(s, w.join[0], 0));
    const unsigned blkA = cg_grid(A.n_block, A.n_diag, GA, CG_THREADS_A);
    const unsigned blkB = JAC ? cg_grid(0, Pm.nn, GB) : cg_grid(Pm.n_block, Pm.n_diag, GB);
    // the solve's one segment, passed by value in the kernel arguments
    CgSeg<VT> seg{};
    seg.A = A;
    seg.Pm = Pm;
    seg.x = x;
    seg.b = b;
    seg.dinv = w.diag_inv.get();
    for (int k = 0; k < 2; ++k) {
        seg.zp[k] = w.zp[k].get();
        seg.ra[k] = w.ra[k].get();
    }
    seg.part_pre = w.part_pre.get();
    seg.part_a = w.part_a.get();
    seg.part_b = w.part_b.get();
    seg.sc = w.scalars.get();
    seg.host_status = w.h_status.p;
    seg.nb_a = (int)blkA;
    seg.nb_b = (int)blkB;
    seg.nparts_b = FLEX ? lv->na : (int)blkB;
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, s, w.scalars.get(), tol, max_iter);
    HIP_CHECK(hipMemsetAsync(w.zp[0].get(), 0, sizeof(double) * 2 * w.n, s));  // p_{-1} = 0
    k_cg_pre1<GA, VT, false><<<dim3(blkA), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr);
    k_cg_pre2<GB, JAC, FLEX, VT, false><<<dim3(blkB), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr);
    if (FLEX) launch_lc_correction(*lv, w.ra[0].get(), nullptr, w.zp[0].get(), w.part_b.get(), s);
    check_launch("cg preamble");
    std::vector<int64_t> snap_k;  // iteration index of every snapshot taken
    int cur = 0;
    int64_t k_next = 0;  // iteration index of the next A/B pair
    size_t nev = 0;      // events used by this solve (profiling only)
    // graph replay of whole batches: every argument of a batch is fixed (the iteration index
    // lives on the device and CG_BATCH is even, so the buffer parity repeats); per-iteration
    // hooks (kernel timing, snapshots) need the direct launches
    const bool graph = w.use_graphs && !m.prof && !h;
    // the captured batch holds the kernels' arguments: the descriptor (matrices, vectors, grids),
    // the tile shapes and the learned correction's arguments
    std::vector<const void*> key = {(const void*)(intptr_t)GA, (const void*)(intptr_t)GB,
                                    (const void*)(intptr_t)JAC, (const void*)(intptr_t)FLEX,
                                    (const void*)(intptr_t)(sizeof(VT))};
    if (lv) {
        static_assert(sizeof(LearnedView) % sizeof(void*) == 0, "view words");
        const void* const* words = reinterpret_cast<const void* const*>(lv);
        key.insert(key.end(), words, words + sizeof(LearnedView) / sizeof(void*));
    }
    {
        static_assert(sizeof(CgSeg<VT>) % sizeof(void*) == 0, "descriptor words");
        const void* const* words = reinterpret_cast<const void* const*>(&seg);
        key.insert(key.end(), words, words + sizeof(CgSeg<VT>) / sizeof(void*));
    }
    auto launch_batch = [&]() {
        for (int q = 0; q < CG_BATCH; ++q) {
            if (m.prof) {
                // the events take the kernel's own start / end timestamps (not a marker
                // before and after the launch), so dispatch gaps are not counted
                while (m.prof_ev.size() < nev + 2) {
                    hipEvent_t e;
                    HIP_CHECK(hipEventCreate(&e));
                    m.prof_ev.push_back(e);
                }
                hipExtLaunchKernelGGL(k_cg_a<GA, FLEX, VT, false>, dim3(blkA), dim3(CG_THREADS_A), 0, s,
                                      m.prof_ev[nev], m.prof_ev[nev + 1], 0, seg, (const CgSeg<VT>*)nullptr,
                                      (const int*)nullptr, cur);
                nev += 2;
            } else {
                k_cg_a<GA, FLEX, VT, false><<<dim3(blkA), dim3(CG_THREADS_A), 0, s>>>(seg, nullptr, nullptr, cur);
            }
            k_cg_b<GB, JAC, FLEX, VT, false><<<dim3(blkB), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr, cur);
            if (FLEX)
                launch_lc_correction(*lv, w.ra[1 - cur].get(), w.ra[cur].get(), w.zp[1 - cur].get(), w.part_b.get(),
                                     s);
            cur ^= 1;
            ++k_next;
            if (h && h->taken + (int)snap_k.size() < h->max && k_next % h->every == 0) {
                const int q = (int)snap_k.size();
                const int64_t nd = w.n;
                k_cg_snapshot<<<dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s>>>(
                    nd, x, w.ra[cur].get(), h->xs + (int64_t)q * nd, h->rs + (int64_t)q * nd);
                snap_k.push_back(k_next);
            }
        }
        check_launch("cg iteration");
    };
    auto enqueue_batch = [&]() {
        if (!graph) {
            launch_batch();
            return;
        }
        if (!w.batch_exec || w.batch_key != key) {
            if (w.batch_exec) {
                HIP_CHECK(hipGraphExecDestroy(w.batch_exec));
                w.batch_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int cur0 = cur;
            const int64_t k0 = k_next;
            launch_batch();
            HIP_CHECK(hipStreamEndCapture(s, &g));
            HIP_CHECK(hipGraphInstantiate(&w.batch_exec, g, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphDestroy(g));
            w.batch_key = key;
            cur = cur0;  // the capture only recorded the batch
            k_next = k0;
        }
        HIP_CHECK(hipGraphLaunch(w.batch_exec, s));
        k_next += CG_BATCH;  // CG_BATCH is even: the buffer parity is back where it started
    };
    CgScalars* hs = w.h_scalars.p + 2;
    // two batches in flight; poll the older one
    int64_t nbatches = 0;
    const int64_t max_batches = (max_iter >= 0 ? max_iter / CG_BATCH + 2 : INT64_MAX);
    // the kernel that ends the solve also writes its status into pinned host memory, so polling
    // needs no copy per batch: the status read after a batch's event is that batch's or a later one's
    volatile int* hstat = w.h_status.p;
    *hstat = CG_RUNNING;
    auto submit = [&](int slot) {
        enqueue_batch();
        HIP_CHECK(hipEventRecord(w.ev[slot], s));
        ++nbatches;
    };
    submit(0);
    if (nbatches < max_batches) submit(1);
    int poll = 0;
    while (true) {
        HIP_CHECK(hipEventSynchronize(w.ev[poll]));
        if (*hstat != CG_RUNNING || nbatches >= max_batches) break;
        submit(poll);
        poll ^= 1;
    }
    HIP_CHECK(hipMemcpyAsync(hs, w.scalars.get(), sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventRecord(w.join[1], s));
    HIP_CHECK(hipStreamWaitEvent(m.stream, w.join[1], 0));
    *iterations = hs->iters;
    if (h) {  // keep the snapshots of iterations the solve actually reached
        int keep = 0;
        for (int64_t kk : snap_k)
            if (kk < hs->iters) ++keep;
        h->taken = keep;
    }
    if (m.prof) {
        // launches [0, iters) did a full iteration; the rest found the solve finished and returned
        for (size_t k = 0; k + 1 < nev; k += 2) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, m.prof_ev[k], m.prof_ev[k + 1]));
            m.prof_all_ms += ms;
            m.prof_all_launches += 1;
            if ((int64_t)(k / 2) < hs->iters) {
                m.prof_full_ms += ms;
                m.prof_full_launches += 1;
            }
        }
    }
    switch (hs->status) {
        case CG_CONVERGED: *out_status = HOEFEM_OK; break;
        case CG_ZERO_RHS:
            HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * w.n, m.stream));  // after the join
            *out_status = HOEFEM_OK;
            break;
        case CG_RUNNING: *out_status = HOEFEM_ERROR_CG_MAX_ITERATIONS; break;
        default: *out_status = (hoefem_status)hs->status; break;
    }
}

template <class VT>
static hoefem_status pcg_solve_t(Model& m, CgWork& w, const VT* Avals, const VT* Pvals, Schwarz* Pre, const double* b,
                                 double* x, double tol, int64_t max_iter, int64_t* iterations, CgHarvest* h = nullptr,
                                 const LearnedView* lv = nullptr) {
    const NodePattern& pat = m.pattern;
    const int64_t nn = pat.num_nodes;
    const int GA = spmv_group_size(pat);
    const int GB = Pre ? spmv_group_size(Pre->P) : 8;
    const int64_t blocks = (nn * 32 + CG_THREADS - 1) / CG_THREADS + 1;
    w.ensure(pat.ndof(), (int)blocks);
    const Csr3<VT> A = view(pat, Avals);
    const Csr3<VT> Pm = Pre ? view(Pre->P, Pvals) : A;
    if (!Pre) {
        k_jacobi_diag<VT><<<dim3((unsigned)((pat.ndof() + 255) / 256)), dim3(256), 0, m.stream>>>(
            pat.ndof(), pat.d_diag_pos.get(), Avals, w.diag_inv.get(), m.d_err.get());
        check_launch("k_jacobi_diag");
        if (read_and_clear_err(m) & KERR_ZERO_DIAG) throw Error(HOEFEM_ERROR_SINGULAR, "Diagonal matrix is not invertible.");
    }
    hoefem_status st = HOEFEM_OK;
    const bool JAC = (Pre == nullptr);
    const bool FLEX = (lv != nullptr);
    if (FLEX && JAC) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "the learned correction needs the Schwarz preconditioner");
    if (FLEX && (lv->na < 1 || lv->na > CG_MAX_BLOCKS))
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "number of aggregates out of range");
#define CG_CASE(ga, gb, j, f)                                                                            \
    if (GA == ga && GB == gb && JAC == j && FLEX == f) {                                                 \
        cg_launch_all<ga, gb, j, f, VT>(m, w, A, Pm, b, x, tol, max_iter, iterations, &st, h, lv);        \
        return st;                                                                                       \
    }
    CG_CASE(8, 8, true, false) CG_CASE(16, 8, true, false) CG_CASE(32, 8, true, false)
    CG_CASE(8, 8, false, false) CG_CASE(8, 16, false, false) CG_CASE(8, 32, false, false)
    CG_CASE(16, 8, false, false) CG_CASE(16, 16, false, false) CG_CASE(16, 32, false, false)
    CG_CASE(32, 8, false, false) CG_CASE(32, 16, false, false) CG_CASE(32, 32, false, false)
    CG_CASE(8, 8, false, true) CG_CASE(8, 16, false, true) CG_CASE(8, 32, false, true)
    CG_CASE(16, 8, false, true) CG_CASE(16, 16, false, true) CG_CASE(16, 32, false, true)
    CG_CASE(32, 8, false, true) CG_CASE(32, 16, false, true) CG_CASE(32, 32, false, true)
#undef CG_CASE
    throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unsupported CG configuration");
}

__global__ void k_fill(int64_t n, double* __restrict__ a, double value);

// ================================================================== allocation-free PCG
// hoefem_pcg_noalloc runs the same kernels as pcg_solve with every work buffer (the interleaved CG
// vectors, the Jacobi diagonal, the per-block partial sums and the iteration state) in `scratch`,
// a fixed array in device memory that exists once per GPU for the life of the process. The host
// side allocates nothing: it launches on the model stream and polls the status with a small copy
// after each batch of iterations. Calls on one GPU share the array, so calls that overlap in time
// on one GPU race (different models included).
constexpr int64_t NOALLOC_MAX_DOFS = 262144;
constexpr int64_t NOALLOC_VEC = 2 * NOALLOC_MAX_DOFS;  // one interleaved [z | p] or [r | Ap] buffer
// layout in doubles: iteration state, the kernels' status word, partial sums, four CG vectors, the
// Jacobi diagonal (every region starts on a 256-byte boundary)
constexpr int64_t NA_SCALARS = 0, NA_STATUS = 32, NA_PART_PRE = 64;
constexpr int64_t NA_PART_A = NA_PART_PRE + CG_MAX_BLOCKS, NA_PART_B = NA_PART_A + CG_MAX_BLOCKS;
constexpr int64_t NA_VEC0 = NA_PART_B + 3 * CG_MAX_BLOCKS;
constexpr int64_t NA_DIAG = NA_VEC0 + 4 * NOALLOC_VEC;
constexpr int64_t NA_TOTAL = NA_DIAG + NOALLOC_MAX_DOFS;  // 9 * 262144 + 5248 doubles = 18.9 MB
static_assert(sizeof(CgScalars) <= 32 * sizeof(double), "iteration state does not fit its region");
static_assert(NA_PART_PRE % 32 == 0 && NA_VEC0 % 32 == 0 && NOALLOC_VEC % 32 == 0, "region alignment");

static __device__ __align__(256) double scratch[NA_TOTAL];

// Device address of `scratch` on the current device, looked up once per device (each GPU holds
// its own copy of the array). The array's name is not a device address in host code.
static double* scratch_address() {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double* addr[MAX_DEVICES] = {};
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) throw Error(HOEFEM_ERROR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(HOEFEM_ERROR_DEVICE, "scratch array has no device address");
        addr[dev] = (double*)p;
    }
    return addr[dev];
}

template <int GA, int GB, bool JAC>
static hoefem_status cg_launch_noalloc(Model& m, double* base, const Csr3<double>& A, const Csr3<double>& Pm,
                                       const double* b, double* x, double tol, int64_t max_iter,
                                       int64_t* iterations) {
    const hipStream_t s = m.stream;
    const int64_t n = 3 * m.nv;
    const unsigned blkA = cg_grid(A.n_block, A.n_diag, GA, CG_THREADS_A);
    const unsigned blkB = JAC ? cg_grid(0, Pm.nn, GB) : cg_grid(Pm.n_block, Pm.n_diag, GB);
    double* zp[2] = {base + NA_VEC0, base + NA_VEC0 + NOALLOC_VEC};
    double* ra[2] = {base + NA_VEC0 + 2 * NOALLOC_VEC, base + NA_VEC0 + 3 * NOALLOC_VEC};
    CgScalars* sc = (CgScalars*)(base + NA_SCALARS);
    CgSeg<double> seg{};  // passed by value in the kernel arguments
    seg.A = A;
    seg.Pm = Pm;
    seg.x = x;
    seg.b = b;
    seg.dinv = base + NA_DIAG;
    for (int k = 0; k < 2; ++k) {
        seg.zp[k] = zp[k];
        seg.ra[k] = ra[k];
    }
    seg.part_pre = base + NA_PART_PRE;
    seg.part_a = base + NA_PART_A;
    seg.part_b = base + NA_PART_B;
    seg.sc = sc;
    seg.host_status = (int*)(base + NA_STATUS);  // a device word here: the status is polled from sc
    seg.nb_a = (int)blkA;
    seg.nb_b = seg.nparts_b = (int)blkB;
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, s, sc, tol, max_iter);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s, 2 * n, zp[0], 0.0);  // p_{-1}
    k_cg_pre1<GA, double, false><<<dim3(blkA), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr);
    k_cg_pre2<GB, JAC, false, double, false><<<dim3(blkB), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr);
    check_launch("cg preamble");
    int cur = 0, status = CG_RUNNING;
    int64_t launched = 0;
    const int64_t max_batches = max_iter >= 0 ? max_iter / CG_BATCH + 2 : INT64_MAX;
    for (int64_t batch = 0; batch < max_batches && status == CG_RUNNING; ++batch) {
        for (int q = 0; q < CG_BATCH; ++q) {
            k_cg_a<GA, false, double, false><<<dim3(blkA), dim3(CG_THREADS_A), 0, s>>>(seg, nullptr, nullptr, cur);
            k_cg_b<GB, JAC, false, double, false><<<dim3(blkB), dim3(CG_THREADS), 0, s>>>(seg, nullptr, nullptr, cur);
            cur ^= 1;
        }
        launched += CG_BATCH;
        check_launch("cg iteration");
        HIP_CHECK(hipMemcpyAsync(&status, &sc->status, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    CgScalars hs;
    HIP_CHECK(hipMemcpyAsync(&hs, sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    *iterations = hs.iters;
    switch (hs.status) {
        case CG_CONVERGED: return HOEFEM_OK;
        case CG_ZERO_RHS:
            hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, 0.0);
            check_launch("k_fill");
            return HOEFEM_OK;
        case CG_RUNNING: return HOEFEM_ERROR_CG_MAX_ITERATIONS;
        default: return (hoefem_status)hs.status;
    }
}

hoefem_status pcg_solve_noalloc(Model& m, const double* Avals, Schwarz* Pre, const double* b, double* x, double tol,
                                int64_t max_iter, int64_t* iterations) {
    const NodePattern& pat = m.pattern;
    *iterations = 0;
    if (pat.ndof() > NOALLOC_MAX_DOFS)
        throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "hoefem_pcg_noalloc supports at most " +
                                                       std::to_string(NOALLOC_MAX_DOFS) + " DOFs (model has " +
                                                       std::to_string(pat.ndof()) + ")");
    if (Pre && Pre->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "FP32 preconditioner with an FP64 operator");
    if (pat.num_nodes == 0) return HOEFEM_OK;
    double* base = scratch_address();
    const int GA = spmv_group_size(pat);
    const int GB = Pre ? spmv_group_size(Pre->P) : 8;
    const Csr3<double> A = view(pat, Avals);
    const Csr3<double> Pm = Pre ? view(Pre->P, (const double*)Pre->d_vals.get()) : A;
    if (!Pre) {
        k_jacobi_diag<double><<<dim3((unsigned)((pat.ndof() + 255) / 256)), dim3(256), 0, m.stream>>>(
            pat.ndof(), pat.d_diag_pos.get(), Avals, base + NA_DIAG, m.d_err.get());
        check_launch("k_jacobi_diag");
        if (read_and_clear_err(m) & KERR_ZERO_DIAG) throw Error(HOEFEM_ERROR_SINGULAR, "Diagonal matrix is not invertible.");
    }
#define NA_CASE(ga, gb, j) \
    if (GA == ga && GB == gb && (Pre == nullptr) == j) return cg_launch_noalloc<ga, gb, j>(m, base, A, Pm, b, x, tol, max_iter, iterations);
    NA_CASE(8, 8, true) NA_CASE(16, 8, true) NA_CASE(32, 8, true)
    NA_CASE(8, 8, false) NA_CASE(8, 16, false) NA_CASE(8, 32, false)
    NA_CASE(16, 8, false) NA_CASE(16, 16, false) NA_CASE(16, 32, false)
    NA_CASE(32, 8, false) NA_CASE(32, 16, false) NA_CASE(32, 32, false)
#undef NA_CASE
    throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unsupported CG configuration");
}

hoefem_status pcg_solve(Model& m, CgWork& w, const double* A, Schwarz* P, const double* b, double* x, double tol,
                        int64_t max_iter, int64_t* iterations) {
    if (P && P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "FP32 preconditioner with an FP64 operator");
    return pcg_solve_t<double>(m, w, A, P ? P->d_vals.get() : nullptr, P, b, x, tol, max_iter, iterations);
}

hoefem_status pcg_solve(Model& m, CgWork& w, const float* A, Schwarz* P, const double* b, double* x, double tol,
                        int64_t max_iter, int64_t* iterations) {
    if (P && !P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "FP64 preconditioner with an FP32 operator");
    return pcg_solve_t<float>(m, w, A, P ? P->d_vals32.get() : nullptr, P, b, x, tol, max_iter, iterations);
}

hoefem_status pcg_solve_harvest(Model& m, CgWork& w, const double* A, Schwarz* P, const double* b, double* x,
                                double tol, int64_t max_iter, int64_t* iterations, CgHarvest* h) {
    if (P && P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "FP32 preconditioner with an FP64 operator");
    return pcg_solve_t<double>(m, w, A, P ? P->d_vals.get() : nullptr, P, b, x, tol, max_iter, iterations, h);
}

hoefem_status pcg_solve_harvest(Model& m, CgWork& w, const float* A, Schwarz* P, const double* b, double* x,
                                double tol, int64_t max_iter, int64_t* iterations, CgHarvest* h) {
    if (P && !P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "FP64 preconditioner with an FP32 operator");
    return pcg_solve_t<float>(m, w, A, P ? P->d_vals32.get() : nullptr, P, b, x, tol, max_iter, iterations, h);
}

hoefem_status pcg_solve_learned(Model& m, CgWork& w, const double* A, Schwarz* P, const LearnedView& lv,
                                const double* b, double* x, double tol, int64_t max_iter, int64_t* iterations) {
    if (!P || P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "learned solve needs the FP64 Schwarz operator");
    return pcg_solve_t<double>(m, w, A, P->d_vals.get(), P, b, x, tol, max_iter, iterations, nullptr, &lv);
}

hoefem_status pcg_solve_learned(Model& m, CgWork& w, const float* A, Schwarz* P, const LearnedView& lv,
                                const double* b, double* x, double tol, int64_t max_iter, int64_t* iterations) {
    if (!P || !P->fp32) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "learned solve needs the FP32 Schwarz operator");
    return pcg_solve_t<float>(m, w, A, P->d_vals32.get(), P, b, x, tol, max_iter, iterations, nullptr, &lv);
}

// ================================================================== batched PCG
// Independent systems solved together. Systems are grouped by tile shape (lanes per node row of A
// and of P, Jacobi or Schwarz); each group has one launch per kernel, whose work-groups are its
// systems' grids laid end to end (the segments), every segment starting on a multiple of 8
// work-groups, so a work-group sits on the XCD (b mod 8) it has when its system runs alone; the
// work-groups in between map to segment -1 and return. The work-group -> segment maps and the
// descriptors are uploaded before each solve, so a captured batch of iterations depends only on the
// launch shapes and is replayed for any systems of the same shapes.

template <class VT>
struct SegOps {
    void (*a)(const CgSeg<VT>*, const int*, unsigned, int, hipStream_t);
    void (*b)(const CgSeg<VT>*, const int*, unsigned, int, hipStream_t);
    // the preamble of one system, its descriptor by value (the kernels of a single solve)
    void (*pre)(const CgSeg<VT>&, hipStream_t);
};

template <int GA, int GB, bool JAC, class VT>
struct SegLaunch {
    static void a(const CgSeg<VT>* s, const int* ma, unsigned na, int cur, hipStream_t st) {
        k_cg_a<GA, false, VT, true><<<dim3(na), dim3(CG_THREADS_A), 0, st>>>(CgSeg<VT>{}, s, ma, cur);
    }
    static void b(const CgSeg<VT>* s, const int* mb, unsigned nb, int cur, hipStream_t st) {
        k_cg_b<GB, JAC, false, VT, true><<<dim3(nb), dim3(CG_THREADS), 0, st>>>(CgSeg<VT>{}, s, mb, cur);
    }
    static void pre(const CgSeg<VT>& one, hipStream_t st) {
        k_cg_pre1<GA, VT, false><<<dim3((unsigned)one.nb_a), dim3(CG_THREADS), 0, st>>>(one, nullptr, nullptr);
        k_cg_pre2<GB, JAC, false, VT, false><<<dim3((unsigned)one.nb_b), dim3(CG_THREADS), 0, st>>>(one, nullptr,
                                                                                                    nullptr);
    }
};

template <class VT>
static SegOps<VT> seg_ops(int GA, int GB, bool JAC) {
#define SEG_CASE(ga, gb, j) \
    if (GA == ga && GB == gb && JAC == j) \
        return SegOps<VT>{&SegLaunch<ga, gb, j, VT>::a, &SegLaunch<ga, gb, j, VT>::b, &SegLaunch<ga, gb, j, VT>::pre};
    SEG_CASE(8, 8, true) SEG_CASE(16, 8, true) SEG_CASE(32, 8, true)
    SEG_CASE(8, 8, false) SEG_CASE(8, 16, false) SEG_CASE(8, 32, false)
    SEG_CASE(16, 8, false) SEG_CASE(16, 16, false) SEG_CASE(16, 32, false)
    SEG_CASE(32, 8, false) SEG_CASE(32, 16, false) SEG_CASE(32, 32, false)
#undef SEG_CASE
    throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "unsupported CG configuration");
}

// work-groups of the two kernels of one system, padded to the next multiple of 8 (upper bound)
constexpr int64_t CG_SEG_MAX_BLOCKS = CG_MAX_BLOCKS + (int64_t)CG_MAX_BLOCKS * CG_THREADS / CG_THREADS_A + 16;

void CgBatch::ensure(size_t seg_bytes, size_t map_ints) {
    if (!cs) {
        HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&upload, hipEventDisableTiming));
    }
    if (seg_bytes > seg_cap || map_ints > map_cap) {
        HIP_CHECK(hipStreamSynchronize(cs));  // no launch reads the buffers being replaced
        for (Graph& g : graphs)                // and no captured batch refers to them afterwards
            if (g.exec) HIP_CHECK(hipGraphExecDestroy(g.exec));
        graphs.clear();
        if (seg_bytes > seg_cap) {
            seg_cap = std::max(seg_bytes, 2 * seg_cap);
            h_seg.reset(new PinnedBuf<char>(seg_cap));
            d_seg.alloc(seg_cap);
        }
        if (map_ints > map_cap) {
            map_cap = std::max(map_ints, 2 * map_cap);
            h_map.reset(new PinnedBuf<int>(map_cap));
            d_map.alloc(map_cap);
        }
    }
}

hipEvent_t CgBatch::take_event() {
    if (!spare.empty()) {
        hipEvent_t e = spare.back();
        spare.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

CgBatch::~CgBatch() {
    if (cs) (void)hipStreamSynchronize(cs);
    for (Graph& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (upload) (void)hipEventDestroy(upload);
    if (cs) (void)hipStreamDestroy(cs);
}

// The pool of a batch: systems join when their inputs are ready and leave when their solve ends.
// Each system's preamble runs as in its own solve (its descriptor by value); its iterations run in
// the batched launches of the current layout (groups of equal tile shape, segments on multiples of 8
// work-groups), so its x, iteration count and status equal pcg_solve's bit for bit. The host keeps
// two batches of CG_BATCH iterations in flight and never blocks: it queries the batches' events, reads
// the statuses the kernels write into pinned memory, and re-lays the launches out when systems join
// (at once) or leave (once the running ones hold at most 3/4 of the launched work-groups). The staging
// buffers are rewritten only after their previous upload has executed; the device copies are ordered
// after every batch enqueued before them.
template <class VT>
struct CgPoolT final : CgPool {
    struct Item {
        CgBatchSystem q;
        int ga = 8, gb = 8;
        bool jac = false;
        unsigned blkA = 0, blkB = 0;
        Csr3<VT> A, Pm;
        SegOps<VT> ops{};
        int state = 0;  // 0 waiting to join, 1 running, 2 stopped (result copy pending), 3 ended
    };
    struct Group {
        int ga, gb;
        bool jac;
        SegOps<VT> ops;
        std::vector<int> items;
        int seg0 = 0;
        int64_t mapA = 0, nA = 0, mapB = 0, nB = 0;
    };
    struct Layout {
        std::vector<Group> groups;
        int nseg = 0;
        int64_t map_total = 0, blocks = 0;
        std::vector<int64_t> key;
    };
    struct Copy {
        hipEvent_t ev;
        std::vector<int> ids;
    };
    CgBatch& B;
    std::vector<Item> items;
    std::vector<int> waiting, running;
    std::vector<Copy> copies;
    Layout lay;
    bool layout_dirty = false, uploaded = false;
    int inflight = 0, oldest = 0, next_slot = 0, cap = 1;

    CgPoolT(CgBatch& b, int capacity) : B(b) {
        // at most `capacity` systems run at once (one solve per scene at a time)
        B.ensure(sizeof(CgSeg<VT>) * (size_t)std::max(capacity, 1), (size_t)std::max(capacity, 1) * CG_SEG_MAX_BLOCKS);
        cap = std::max(capacity, 1);
    }
    ~CgPoolT() override {
        if (B.cs) (void)hipStreamSynchronize(B.cs);
        for (Copy& c : copies) B.spare.push_back(c.ev);
    }

    int join(const CgBatchSystem& s) override {
        Item it;
        it.q = s;
        it.q.iterations = 0;
        it.q.status = HOEFEM_OK;
        Model& m = *s.m;
        const NodePattern& pat = m.pattern;
        if (s.P && s.P->fp32 != (sizeof(VT) == sizeof(float)))
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "preconditioner precision differs from the operator's");
        const int64_t blocks = (pat.num_nodes * 32 + CG_THREADS - 1) / CG_THREADS + 1;
        s.w->ensure(pat.ndof(), (int)blocks);
        it.jac = s.P == nullptr;
        it.ga = spmv_group_size(pat);
        it.gb = s.P ? spmv_group_size(s.P->P) : 8;
        it.A = view(pat, (const VT*)s.A);
        if (s.P) {
            const VT* pv;
            if constexpr (sizeof(VT) == sizeof(double))
                pv = (const VT*)s.P->d_vals.get();
            else
                pv = (const VT*)s.P->d_vals32.get();
            it.Pm = view(s.P->P, pv);
        } else {
            it.Pm = it.A;  // the Jacobi diagonal is already in w->diag_inv
        }
        it.blkA = cg_grid(it.A.n_block, it.A.n_diag, it.ga, CG_THREADS_A);
        it.blkB = it.jac ? cg_grid(0, it.Pm.nn, it.gb) : cg_grid(it.Pm.n_block, it.Pm.n_diag, it.gb);
        it.ops = seg_ops<VT>(it.ga, it.gb, it.jac);
        if ((int)(waiting.size() + running.size()) >= cap)
            throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "more running systems than the CG pool holds");
        items.push_back(it);
        waiting.push_back((int)items.size() - 1);
        return (int)items.size() - 1;
    }

    const CgBatchSystem& system(int id) const override { return items[id].q; }
    bool idle() const override { return waiting.empty() && running.empty() && copies.empty(); }

    CgSeg<VT> descriptor(const Item& it) const {
        CgWork& w = *it.q.w;
        CgSeg<VT> S{};
        S.A = it.A;
        S.Pm = it.Pm;
        S.x = it.q.x;
        S.b = it.q.b;
        S.dinv = w.diag_inv.get();
        for (int k = 0; k < 2; ++k) {
            S.zp[k] = w.zp[k].get();
            S.ra[k] = w.ra[k].get();
        }
        S.part_pre = w.part_pre.get();
        S.part_a = w.part_a.get();
        S.part_b = w.part_b.get();
        S.sc = w.scalars.get();
        S.host_status = w.h_status.p;
        S.nb_a = (int)it.blkA;
        S.nb_b = S.nparts_b = (int)it.blkB;
        S.tol = it.q.tol;
        S.max_iter = it.q.max_iter;
        return S;
    }

    // pcg_solve's start of one system on the solver stream: x = 0, iteration state, p_{-1} = 0, preambles
    void preamble(Item& it) {
        const hipStream_t cs = B.cs;
        CgWork& w = *it.q.w;
        *(volatile int*)w.h_status.p = CG_RUNNING;
        HIP_CHECK(hipMemsetAsync(it.q.x, 0, sizeof(double) * 3 * it.A.nn, cs));
        hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, cs, w.scalars.get(), it.q.tol, it.q.max_iter);
        HIP_CHECK(hipMemsetAsync(w.zp[0].get(), 0, sizeof(double) * 2 * w.n, cs));
        it.ops.pre(descriptor(it), cs);
        check_launch("cg preamble");
    }

    static int64_t round8(int64_t v) { return (v + 7) / 8 * 8; }

    Layout build_layout(const std::vector<int>& live) const {
        Layout L;
        for (int k : live) {
            const Item& it = items[k];
            size_t g = 0;
            while (g < L.groups.size() &&
                   !(L.groups[g].ga == it.ga && L.groups[g].gb == it.gb && L.groups[g].jac == it.jac))
                ++g;
            if (g == L.groups.size()) L.groups.push_back(Group{it.ga, it.gb, it.jac, it.ops, {}, 0, 0, 0, 0, 0});
            L.groups[g].items.push_back(k);
        }
        for (Group& g : L.groups) {
            g.seg0 = L.nseg;
            L.nseg += (int)g.items.size();
            int64_t a = 0, b = 0;
            for (int k : g.items) {
                a = round8(a) + items[k].blkA;
                b = round8(b) + items[k].blkB;
            }
            g.mapA = L.map_total;
            g.nA = a;
            g.mapB = L.map_total + a;
            g.nB = b;
            L.map_total += a + b;
            L.blocks += a + b;
        }
        L.key = {(int64_t)sizeof(VT)};
        for (const Group& g : L.groups)
            for (int64_t v : {(int64_t)g.ga, (int64_t)g.gb, (int64_t)g.jac, (int64_t)g.seg0, g.mapA, g.nA, g.mapB, g.nB})
                L.key.push_back(v);
        return L;
    }

    void upload(const Layout& L) {
        CgSeg<VT>* hs = reinterpret_cast<CgSeg<VT>*>(B.h_seg->p);
        int* hm = B.h_map->p;
        for (const Group& g : L.groups) {
            int64_t a = 0, b = 0;
            for (size_t j = 0; j < g.items.size(); ++j) {
                const Item& it = items[g.items[j]];
                const int64_t ba = round8(a), bb = round8(b);
                for (int64_t t = a; t < ba; ++t) hm[g.mapA + t] = -1;
                for (int64_t t = b; t < bb; ++t) hm[g.mapB + t] = -1;
                for (int64_t t = 0; t < it.blkA; ++t) hm[g.mapA + ba + t] = (int)j;
                for (int64_t t = 0; t < it.blkB; ++t) hm[g.mapB + bb + t] = (int)j;
                a = ba + it.blkA;
                b = bb + it.blkB;
                CgSeg<VT> S = descriptor(it);
                S.base_a = (int)ba;
                S.base_b = (int)bb;
                hs[g.seg0 + j] = S;
            }
        }
        const hipStream_t cs = B.cs;
        HIP_CHECK(hipMemcpyAsync(B.d_seg.get(), hs, sizeof(CgSeg<VT>) * L.nseg, hipMemcpyHostToDevice, cs));
        HIP_CHECK(hipMemcpyAsync(B.d_map.get(), hm, sizeof(int) * L.map_total, hipMemcpyHostToDevice, cs));
        HIP_CHECK(hipEventRecord(B.upload, cs));
        uploaded = true;
        B.relayouts += 1;
    }

    bool upload_free() const { return !uploaded || hipEventQuery(B.upload) == hipSuccess; }

    void launch_batch(const Layout& L) {
        const hipStream_t cs = B.cs;
        const CgSeg<VT>* segs = reinterpret_cast<const CgSeg<VT>*>(B.d_seg.get());
        const int* dmap = B.d_map.get();
        int cur = 0;
        for (int q = 0; q < CG_BATCH; ++q) {
            for (const Group& g : L.groups) g.ops.a(segs + g.seg0, dmap + g.mapA, (unsigned)g.nA, cur, cs);
            for (const Group& g : L.groups) g.ops.b(segs + g.seg0, dmap + g.mapB, (unsigned)g.nB, cur, cs);
            cur ^= 1;  // CG_BATCH is even: every batch starts at parity 0
        }
        check_launch("batched cg iteration");
    }

    // a layout seen before replays its captured batch; the first sight launches directly
    void submit() {
        const hipStream_t cs = B.cs;
        hipGraphExec_t exec = nullptr;
        if (B.use_graphs) {
            CgBatch::Graph* gr = nullptr;
            for (CgBatch::Graph& g : B.graphs)
                if (g.key == lay.key) gr = &g;
            if (!gr) {
                constexpr size_t MAX_GRAPHS = 256;
                if (B.graphs.size() >= MAX_GRAPHS) {
                    size_t old = 0;
                    for (size_t k = 1; k < B.graphs.size(); ++k)
                        if (B.graphs[k].used < B.graphs[old].used) old = k;
                    if (B.graphs[old].exec) HIP_CHECK(hipGraphExecDestroy(B.graphs[old].exec));
                    B.graphs.erase(B.graphs.begin() + (std::ptrdiff_t)old);
                }
                B.graphs.push_back(CgBatch::Graph{});
                gr = &B.graphs.back();
                gr->key = lay.key;
            }
            gr->used = ++B.clock;
            if (!gr->exec && ++gr->seen >= 2) {
                hipGraph_t gph = nullptr;
                HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                launch_batch(lay);
                HIP_CHECK(hipStreamEndCapture(cs, &gph));
                HIP_CHECK(hipGraphInstantiate(&gr->exec, gph, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(gph));
            }
            exec = gr->exec;
        }
        if (exec)
            HIP_CHECK(hipGraphLaunch(exec, cs));
        else
            launch_batch(lay);
        HIP_CHECK(hipEventRecord(B.ev[next_slot], cs));
        next_slot ^= 1;
        ++inflight;
    }

    void progress(std::vector<int>& ended) override {
        const hipStream_t cs = B.cs;
        // 1. completed batches: the statuses the kernels wrote
        bool batch_done = false;
        while (inflight > 0 && hipEventQuery(B.ev[oldest]) == hipSuccess) {
            --inflight;
            oldest ^= 1;
            batch_done = true;
        }
        if (batch_done) {
            std::vector<int> stopped;
            for (size_t k = 0; k < running.size();) {
                Item& it = items[running[k]];
                if (*(volatile int*)it.q.w->h_status.p != CG_RUNNING) {
                    it.state = 2;
                    stopped.push_back(running[k]);
                    CgWork& w = *it.q.w;
                    HIP_CHECK(hipMemcpyAsync(w.h_scalars.p + 2, w.scalars.get(), sizeof(CgScalars),
                                             hipMemcpyDeviceToHost, cs));
                    running[k] = running.back();
                    running.pop_back();
                } else {
                    ++k;
                }
            }
            if (!stopped.empty()) {
                Copy c{B.take_event(), std::move(stopped)};
                HIP_CHECK(hipEventRecord(c.ev, cs));
                copies.push_back(std::move(c));
                layout_dirty = true;
            }
        }
        // 2. finished result copies
        for (size_t k = 0; k < copies.size();) {
            if (hipEventQuery(copies[k].ev) != hipSuccess) {
                ++k;
                continue;
            }
            for (int id : copies[k].ids) {
                Item& it = items[id];
                const CgScalars& r = it.q.w->h_scalars.p[2];
                it.q.iterations = r.iters;
                switch (r.status) {
                    case CG_CONVERGED:
                    case CG_ZERO_RHS: it.q.status = HOEFEM_OK; break;  // x was zeroed by the preamble
                    case CG_RUNNING: it.q.status = HOEFEM_ERROR_CG_MAX_ITERATIONS; break;
                    default: it.q.status = (hoefem_status)r.status; break;
                }
                it.state = 3;
                ended.push_back(id);
            }
            B.spare.push_back(copies[k].ev);
            copies.erase(copies.begin() + (std::ptrdiff_t)k);
        }
        // 3. joins, 4. a new layout (the staging buffers must be free for either)
        if ((!waiting.empty() || layout_dirty) && upload_free()) {
            bool relayout = !waiting.empty();
            for (int id : waiting) {
                preamble(items[id]);
                items[id].state = 1;
                running.push_back(id);
            }
            waiting.clear();
            if (!relayout && layout_dirty) {
                int64_t live = 0;
                for (int id : running) live += round8(items[id].blkA) + round8(items[id].blkB);
                relayout = running.empty() || live * 4 <= lay.blocks * 3;
            }
            if (relayout) {
                std::sort(running.begin(), running.end());
                lay = build_layout(running);
                if (!running.empty()) upload(lay);
                layout_dirty = false;
            }
        }
        // 5. keep two batches in flight while systems run
        while (!running.empty() && inflight < 2) submit();
    }

};

std::unique_ptr<CgPool> make_cg_pool(CgBatch& B, bool fp32, int capacity) {
    if (fp32) return std::unique_ptr<CgPool>(new CgPoolT<float>(B, capacity));
    return std::unique_ptr<CgPool>(new CgPoolT<double>(B, capacity));
}

void launch_jacobi_diag(Model& m, CgWork& w, const void* A, bool fp32) {
    const NodePattern& pat = m.pattern;
    const int64_t blocks = (pat.num_nodes * 32 + CG_THREADS - 1) / CG_THREADS + 1;
    w.ensure(pat.ndof(), (int)blocks);
    const unsigned g = (unsigned)((pat.ndof() + 255) / 256);
    if (fp32)
        k_jacobi_diag<float><<<dim3(g), dim3(256), 0, m.stream>>>(pat.ndof(), pat.d_diag_pos.get(), (const float*)A,
                                                                 w.diag_inv.get(), m.d_err.get());
    else
        k_jacobi_diag<double><<<dim3(g), dim3(256), 0, m.stream>>>(pat.ndof(), pat.d_diag_pos.get(), (const double*)A,
                                                                  w.diag_inv.get(), m.d_err.get());
    check_launch("k_jacobi_diag");
}

// ================================================================== Newton / BE kernels
// The vector updates of the Backward-Euler Newton loop round every product and sum on its own
// (no fused multiply-add), as the reference's nalgebra arithmetic does on the host: a caller that
// composes the same step on the host from the dynamic-system entry points (eval_f, apply_mass,
// solve_jacobian_combination) then gets bit-identical iterates.

// uu = u_p + w * dt with Dirichlet DOFs zeroed (BackwardEulerImplicitFunction::eval_into,
// hamilton2/src/integrators/euler.rs:97-114; apply_dirichlet_bcs_unknowns,
// simulation_toolbox/src/fem/system_assembly.rs:78-104)
__global__ void k_state(int64_t ndof, const double* __restrict__ up, const double* __restrict__ w, double dt,
                        const uint8_t* __restrict__ dirmask, const double* __restrict__ bcu, double* __restrict__ uu) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndof) return;
    const double val = up[i] + w[i] * dt;
    uu[i] = dirmask[i / 3] ? (bcu ? bcu[i] : 0.0) : val;
}

// dst = src with the Dirichlet values on the Dirichlet DOFs (bc: a DOF vector of prescribed values,
// nullptr for zeros); the copies NewtonCgDynamicSystem::eval_f and set_state make,
// newton_cg.rs:449-455, :662-671
__global__ void k_bc_copy(int64_t ndof, const double* __restrict__ src, const uint8_t* __restrict__ dirmask,
                          const double* __restrict__ bc, double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndof) return;
    const double val = src[i];
    dst[i] = dirmask[i / 3] ? (bc ? bc[i] : 0.0) : val;
}

// f = M g + f_int (element forces in color order) - D v, Dirichlet rows zeroed
// (NewtonCgDynamicSystem::eval_f, newton_cg.rs:441-478); one thread per node
__global__ void k_node_force(int64_t nn, const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_blk,
                             const int32_t* __restrict__ inc_loc, int n, const double* __restrict__ fe,
                             const double* __restrict__ gforce, const double* __restrict__ dv,
                             const uint8_t* __restrict__ dirmask, double* __restrict__ f) {
#pragma clang fp contract(off)
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double f0 = gforce[3 * v], f1 = gforce[3 * v + 1], f2 = gforce[3 * v + 2];
    for (int it = inc_ptr[v]; it < inc_ptr[v + 1]; ++it) {
        const double* src = fe + ((int64_t)inc_blk[it] * n + inc_loc[it]) * 3;
        f0 += src[0];
        f1 += src[1];
        f2 += src[2];
    }
    if (dv) {  // f += -D v (spmv_csr with beta = 1, alpha = -1)
        f0 = f0 - dv[3 * v];
        f1 = f1 - dv[3 * v + 1];
        f2 = f2 - dv[3 * v + 2];
    }
    if (dirmask[v]) f0 = f1 = f2 = 0.0;
    f[3 * v] = f0;
    f[3 * v + 1] = f1;
    f[3 * v + 2] = f2;
}

// dst = src * sc  (F <- -dt f, euler.rs:110)
__global__ void k_scale(int64_t n, const double* __restrict__ src, double sc, double* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] * sc;
}

// y <- y + M x (apply_mass_matrix, newton_cg.rs:428-431: spmv_csr(1, y, 1, M, x)); with
// partials, also the per-block partial sums of |y|^2 (the Newton residual norm)
template <int G>
__global__ __launch_bounds__(CG_THREADS) void k_mass_add(Csr3<double> M, const double* __restrict__ x,
                                                         double* __restrict__ y, double* __restrict__ partials) {
#pragma clang fp contract(off)
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t v = gid / G;
    const int lane = (int)(gid % G);
    const bool valid = v < M.nn;
    double m0, m1, m2;
    row_block_dot<G>(M, v, lane, valid, [&](int64_t b, double& x0, double& x1, double& x2) {
        x0 = x[3 * b];
        x1 = x[3 * b + 1];
        x2 = x[3 * b + 2];
    }, m0, m1, m2);
    double part[1] = {0.0};
    if (valid && lane == 0) {
        const double y0 = y[3 * v] + m0, y1 = y[3 * v + 1] + m1, y2 = y[3 * v + 2] + m2;
        y[3 * v] = y0;
        y[3 * v + 1] = y1;
        y[3 * v + 2] = y2;
        part[0] = y0 * y0 + y1 * y1 + y2 * y2;
    }
    if (partials) block_partials<1>(part, partials, blockIdx.x);
}

// w <- w + delta * (-sol)  (line-search step along p = -sol, hamilton2/src/newton.rs:113-116, :219)
__global__ void k_axpy_neg(int64_t n, double* __restrict__ w, double delta, const double* __restrict__ sol) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    w[i] = w[i] + delta * (-sol[i]);
}

// v_n = v ; u_p = u + v * dt
__global__ void k_be_begin(int64_t n, const double* __restrict__ u, const double* __restrict__ v, double dt,
                           double* __restrict__ up, double* __restrict__ vn) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vn[i] = v[i];
    up[i] = u[i] + v[i] * dt;
}

// v += w ; u += v * dt ; then the Dirichlet values on the Dirichlet unknowns (euler.rs:195-196,
// newton_cg.rs:293)
__global__ void k_be_end(int64_t n, double* __restrict__ u, double* __restrict__ v, const double* __restrict__ w,
                         double dt, const uint8_t* __restrict__ dirmask, const double* __restrict__ bcu,
                         const double* __restrict__ bcv) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vv = v[i] + w[i];
    const double uu = u[i] + vv * dt;
    const bool d = dirmask[i / 3] != 0;
    v[i] = d ? (bcv ? bcv[i] : 0.0) : vv;
    u[i] = d ? (bcu ? bcu[i] : 0.0) : uu;
}

// the Dirichlet values on the Dirichlet DOFs of a and b (apply_dirichlet_bcs_unknowns,
// simulation_toolbox/src/fem/system_assembly.rs:78-104)
__global__ void k_set_dirichlet(int64_t n, double* __restrict__ a, double* __restrict__ b,
                                const uint8_t* __restrict__ dirmask, const double* __restrict__ bca,
                                const double* __restrict__ bcb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dirmask[i / 3]) {
        a[i] = bca ? bca[i] : 0.0;
        if (b) b[i] = bcb ? bcb[i] : 0.0;
    }
}

__global__ void k_finite(int64_t n, const double* __restrict__ a, const double* __restrict__ b, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!isfinite(a[i]) || !isfinite(b[i])) atomicOr(err, KERR_NONFINITE);
}

__global__ void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_fill(int64_t n, double* __restrict__ a, double value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = value;
}

// partial sums of a.b over a grid-stride loop (one per block)
__global__ __launch_bounds__(CG_THREADS) void k_dot_partials(int64_t n, const double* __restrict__ a,
                                                             const double* __restrict__ b, double* __restrict__ partials) {
    double part[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * CG_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * CG_THREADS)
        part[0] += a[i] * b[i];
    block_partials<1>(part, partials, blockIdx.x);
}

// a.b on the model stream (deterministic), returned to the host
double device_dot(Model& m, CgWork& w, const double* a, const double* b, double* d_out, double* h_out) {
    const int64_t n = 3 * m.nv;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + CG_THREADS - 1) / CG_THREADS, 512));
    w.ensure(n, (int)g + 1);
    k_dot_partials<<<dim3(g), dim3(CG_THREADS), 0, m.stream>>>(n, a, b, w.partials.get());
    k_sum_partials<<<dim3(1), dim3(CG_THREADS), 0, m.stream>>>(w.partials.get(), (int)g, d_out);
    check_launch("device_dot");
    HIP_CHECK(hipMemcpyAsync(h_out, d_out, sizeof(double), hipMemcpyDeviceToHost, m.stream));
    HIP_CHECK(hipStreamSynchronize(m.stream));
    return *h_out;
}

__global__ void k_add(int64_t n, double* __restrict__ a, const double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

void launch_add(Model& m, double* a, const double* b, int64_t n) {
    if (n <= 0) return;
    k_add<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(n, a, b);
    check_launch("k_add");
}

void launch_f64_to_f32(Model& m, const double* a, float* b, int64_t n) {
    if (n <= 0) return;
    k_f64_to_f32<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(n, a, b);
    check_launch("k_f64_to_f32");
}

// out = (0 + cd D) + cm M, the damping and mass terms of compute_jacobian_combination_into
// (simulation_toolbox/src/fem/system_assembly.rs:197-235) when the stiffness term is absent;
// D may be null
template <class T>
__global__ void k_comb(int64_t n, T* __restrict__ out, double cd, const double* __restrict__ D, double cm,
                       const double* __restrict__ M) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    if (D) v += cd * D[i];
    v += cm * M[i];
    out[i] = (T)v;
}

template <class T>
void launch_comb(Model& m, int64_t n, T* out, double cd, const double* D, double cm, const double* M) {
    if (n <= 0) return;
    k_comb<T><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(n, out, cd, D, cm, M);
    check_launch("k_comb");
}
template void launch_comb<double>(Model&, int64_t, double*, double, const double*, double, const double*);
template void launch_comb<float>(Model&, int64_t, float*, double, const double*, double, const double*);

void launch_fill(Model& m, double* a, int64_t n, double value) {
    if (n <= 0) return;
    k_fill<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(n, a, value);
    check_launch("k_fill");
}

void launch_f32_to_f64(Model& m, const float* a, double* b, int64_t n) {
    if (n <= 0) return;
    k_f32_to_f64<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream>>>(n, a, b);
    check_launch("k_f32_to_f64");
}

// v = v_n + w with the Dirichlet velocities applied (zero): the velocity eval_f sees
__global__ void k_state_v(int64_t ndof, const double* __restrict__ vn, const double* __restrict__ w,
                          const uint8_t* __restrict__ dirmask, const double* __restrict__ bcv, double* __restrict__ vv) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndof) return;
    const double val = vn[i] + w[i];
    vv[i] = dirmask[i / 3] ? (bcv ? bcv[i] : 0.0) : val;
}

void launch_state_v(Model& m, const double* vn, const double* w, const uint8_t* dirmask, const double* bcv,
                    double* vv) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_state_v, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, vn, w, dirmask, bcv,
                       vv);
    check_launch("k_state_v");
}

void launch_state(Model& m, const double* up, const double* w, double dt, const uint8_t* dirmask, const double* bcu,
                  double* uu) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, up, w, dt, dirmask, bcu,
                       uu);
    check_launch("k_state");
}

void launch_bc_copy(Model& m, const double* src, const uint8_t* dirmask, const double* bc, double* dst) {
    const int64_t n = 3 * m.nv;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bc_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, src, dirmask, bc, dst);
    check_launch("k_bc_copy");
}

void launch_node_force(Model& m, const double* gforce, const double* dv, const uint8_t* dirmask, double* f) {
    const int64_t nn = m.nv;
    if (nn <= 0) return;
    launch_timed(prof_slot(m), k_node_force, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, m.stream, nn,
                 (const int32_t*)m.scatter.d_inc_ptr.get(), (const int32_t*)m.scatter.d_inc_blk.get(),
                 (const int32_t*)m.scatter.d_inc_loc.get(), m.n, (const double*)m.d_fe.get(), gforce, dv, dirmask, f);
    check_launch("k_node_force");
}

void launch_scale(Model& m, const double* src, double sc, double* dst) {
    const int64_t n = 3 * m.nv;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, src, sc, dst);
    check_launch("k_scale");
}

// y += M x on the model pattern; with norm2 (device), also |y|^2 (deterministic partial sums)
void launch_mass_add(Model& m, const double* Mvals, const double* x, double* y, CgWork& cw, double* norm2) {
    const NodePattern& pat = m.pattern;
    if (pat.num_nodes == 0) return;
    const int G = spmv_group_size(pat);
    const Csr3<double> Mv = view(pat, Mvals);
    const unsigned blocks = (unsigned)((pat.num_nodes * G + CG_THREADS - 1) / CG_THREADS);
    double* part = nullptr;
    if (norm2) {
        cw.ensure(pat.ndof(), (int)blocks + 1);
        part = cw.partials.get();
    }
#define MASS_CASE(GG) \
    hipLaunchKernelGGL(k_mass_add<GG>, dim3(blocks), dim3(CG_THREADS), 0, m.stream, Mv, x, y, part)
    if (G == 32)
        MASS_CASE(32);
    else if (G == 16)
        MASS_CASE(16);
    else
        MASS_CASE(8);
#undef MASS_CASE
    check_launch("k_mass_add");
    if (norm2) {
        k_sum_partials<<<dim3(1), dim3(CG_THREADS), 0, m.stream>>>(part, (int)blocks, norm2);
        check_launch("k_sum_partials");
    }
}

void launch_axpy_neg(Model& m, double* w, double delta, const double* sol) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_axpy_neg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, w, delta, sol);
    check_launch("k_axpy_neg");
}

void launch_be_begin(Model& m, const double* u, const double* v, double dt, double* up, double* vn) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_be_begin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, u, v, dt, up, vn);
    check_launch("k_be_begin");
}

void launch_be_end(Model& m, double* u, double* v, const double* w, double dt, const uint8_t* dirmask,
                   const double* bcu, const double* bcv) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_be_end, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, u, v, w, dt, dirmask,
                       bcu, bcv);
    check_launch("k_be_end");
}

void launch_set_dirichlet(Model& m, double* a, double* b, const uint8_t* dirmask, const double* bca,
                          const double* bcb) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_set_dirichlet, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, a, b, dirmask,
                       bca, bcb);
    check_launch("k_set_dirichlet");
}

void launch_finite_check(Model& m, const double* a, const double* b) {
    const int64_t n = 3 * m.nv;
    hipLaunchKernelGGL(k_finite, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m.stream, n, a, b, m.d_err.get());
    check_launch("k_finite");
}

}  // namespace hoefem
