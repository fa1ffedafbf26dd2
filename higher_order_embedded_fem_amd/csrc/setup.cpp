// Host-side setup of the hoefem library: node-level CSR patterns, first-fit coloring and
// the color-ordered block incidence used by the atomic-free gather kernels.
#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>

#include "internal.hpp"

namespace hoefem {

void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(HOEFEM_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void KernelProf::next(hipEvent_t& a, hipEvent_t& b) {
    while (ev.size() < 2 * (pending + 1)) {
        hipEvent_t x;
        HIP_CHECK(hipEventCreate(&x));
        ev.push_back(x);
    }
    a = ev[2 * pending];
    b = ev[2 * pending + 1];
    ++pending;
}

void KernelProf::resolve() {
    for (size_t k = 0; k < pending; ++k) {
        float t = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
        ms += t;
        launches += 1;
    }
    pending = 0;
}

KernelProf::~KernelProf() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

void raise_lds_limit(const void* func, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> limit;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    size_t& cur = limit[{dev, func}];
    if (bytes <= cur) return;
    HIP_CHECK(hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
}

int nodes_per_element(int t) {
    switch (t) {
        case TET4: return 4;
        case TET10: return 10;
        case TET20: return 20;
        case HEX8: return 8;
        case HEX20: return 20;
        case HEX27: return 27;
    }
    return -1;
}

int geometry_nodes(int t) {
    switch (t) {
        case TET4: case TET10: case TET20: return 4;
        case HEX8: case HEX20: case HEX27: return 8;
    }
    return -1;
}

int32_t NodePattern::find(int32_t a, int32_t b) const {
    const int32_t* lo = ncol.data() + nptr[a];
    const int32_t* hi = ncol.data() + nptr[a + 1];
    const int32_t* it = std::lower_bound(lo, hi, b);
    if (it == hi || *it != b) return -1;
    return (int32_t)(it - lo);
}

void NodePattern::upload(hipStream_t s) {
    d_nptr.upload(nptr, s);
    d_ncol.upload(ncol, s);
    d_rowptr.upload(rowptr, s);
    d_col.upload(col, s);
    d_diag_pos.upload(diag_pos, s);
    d_diag_only.upload(diag_only, s);
    entries.clear();
    for (int64_t v = 0; v < num_nodes; ++v)
        if (!diag_only[v])
            entries.push_back(NodeEntry{(int32_t)v, nptr[v], nptr[v + 1] - nptr[v], rowptr[3 * v]});
    n_block = (int64_t)entries.size();
    for (int64_t v = 0; v < num_nodes; ++v)
        if (diag_only[v]) entries.push_back(NodeEntry{(int32_t)v, 0, 0, rowptr[3 * v]});
    n_diag = (int64_t)entries.size() - n_block;
    d_entries.upload(entries, s);
}

// The set of node pairs of every node set, expanded to 3x3 blocks, sorted by row and column:
// the same set of coordinates CsrParAssembler::assemble_pattern and CooMatrix::to_csr produce
// (fenris/src/assembly.rs:177-249, fenris/src/sparse.rs:102-165). Neighbour lists are built per
// node by stamping the nodes of its incident sets.
void build_node_pattern(int64_t num_nodes, int64_t nsets, const int64_t* off, const int32_t* nodes,
                        const std::vector<int32_t>& diag_only_nodes, NodePattern& out) {
    out.num_nodes = num_nodes;
    std::vector<int64_t> sptr(num_nodes + 1, 0);
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t k = off[s]; k < off[s + 1]; ++k) {
            if (nodes[k] < 0 || nodes[k] >= num_nodes) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "node index out of range");
            sptr[nodes[k] + 1]++;
        }
    for (int64_t a = 0; a < num_nodes; ++a) sptr[a + 1] += sptr[a];
    std::vector<int32_t> sets(sptr[num_nodes]);
    {
        std::vector<int64_t> fill(sptr.begin(), sptr.end() - 1);
        for (int64_t s = 0; s < nsets; ++s)
            for (int64_t k = off[s]; k < off[s + 1]; ++k) sets[fill[nodes[k]]++] = (int32_t)s;
    }
    out.diag_only.assign(num_nodes, 0);
    for (int32_t v : diag_only_nodes) {
        if (v < 0 || v >= num_nodes) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "node index out of range");
        if (sptr[v + 1] == sptr[v]) out.diag_only[v] = 1;
    }
    std::vector<int64_t> stamp(num_nodes, -1);
    out.nptr.assign(num_nodes + 1, 0);
    out.ncol.clear();
    std::vector<int32_t> tmp;
    out.max_nbr = 0;
    for (int64_t a = 0; a < num_nodes; ++a) {
        tmp.clear();
        for (int64_t k = sptr[a]; k < sptr[a + 1]; ++k) {
            const int32_t s = sets[k];
            for (int64_t j = off[s]; j < off[s + 1]; ++j) {
                const int32_t b = nodes[j];
                if (stamp[b] != a) {
                    stamp[b] = a;
                    tmp.push_back(b);
                }
            }
        }
        if (out.diag_only[a]) tmp.push_back((int32_t)a);
        std::sort(tmp.begin(), tmp.end());
        out.ncol.insert(out.ncol.end(), tmp.begin(), tmp.end());
        out.nptr[a + 1] = (int32_t)out.ncol.size();
        out.max_nbr = std::max(out.max_nbr, (int)tmp.size());
    }
    const int64_t ndof = 3 * num_nodes;
    out.rowptr.assign(ndof + 1, 0);
    out.diag_pos.assign(ndof, -1);
    int64_t nnz = 0;
    for (int64_t a = 0; a < num_nodes; ++a) {
        const int64_t k = out.nptr[a + 1] - out.nptr[a];
        for (int i = 0; i < 3; ++i) {
            nnz += out.diag_only[a] ? 1 : 3 * k;
            if (nnz > INT32_MAX) throw Error(HOEFEM_ERROR_INVALID_ARGUMENT, "nnz exceeds int32");
            out.rowptr[3 * a + i + 1] = (int32_t)nnz;
        }
    }
    out.col.resize(nnz);
    for (int64_t a = 0; a < num_nodes; ++a)
        for (int i = 0; i < 3; ++i) {
            const int64_t row = 3 * a + i;
            int64_t pos = out.rowptr[row];
            if (out.diag_only[a]) {
                out.col[pos] = (int32_t)row;
                out.diag_pos[row] = (int32_t)pos;
                continue;
            }
            for (int32_t k = out.nptr[a]; k < out.nptr[a + 1]; ++k)
                for (int j = 0; j < 3; ++j) {
                    const int32_t c = 3 * out.ncol[k] + j;
                    if (c == row) out.diag_pos[row] = (int32_t)pos;
                    out.col[pos++] = c;
                }
        }
}

// reorder_mesh_par (fenris/src/reorder.rs:52-88): a reverse Cuthill-McKee vertex permutation of
// the node adjacency pattern (cuthill_mckee, :169-235: every connected component is started
// from the first unvisited vertex of minimum degree and walked breadth first, the neighbours of
// each vertex visited by increasing degree, ties in increasing vertex index; then reversed), and
// the elements stably sorted by their smallest new vertex index. Permutations map target index
// to source index: new_vertices[i] = vertices[vperm[i]].
void mesh_reorder(int64_t nv, int64_t ne, int n, const int32_t* conn, int64_t* vperm, int64_t* eperm) {
    std::vector<int64_t> eoff((size_t)ne + 1);
    for (int64_t e = 0; e <= ne; ++e) eoff[e] = e * n;
    NodePattern pat;
    build_node_pattern(nv, ne, eoff.data(), conn, {}, pat);
    std::vector<int32_t> deg((size_t)nv);
    for (int64_t v = 0; v < nv; ++v) deg[v] = pat.nptr[v + 1] - pat.nptr[v];
    std::vector<uint8_t> visited((size_t)nv, 0);
    std::vector<int64_t> order;
    order.reserve((size_t)nv);
    std::vector<int32_t> adj;
    int64_t scan = 0;  // every vertex before scan is visited
    while ((int64_t)order.size() < nv) {
        while (scan < nv && visited[scan]) ++scan;
        int64_t start = scan;
        for (int64_t v = scan; v < nv; ++v)
            if (!visited[v] && deg[v] < deg[start]) start = v;
        size_t head = order.size();
        order.push_back(start);
        visited[start] = 1;
        while (head < order.size()) {
            const int64_t v = order[head++];
            adj.assign(pat.ncol.begin() + pat.nptr[v], pat.ncol.begin() + pat.nptr[v + 1]);
            std::stable_sort(adj.begin(), adj.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
            for (int32_t a : adj)
                if (!visited[a]) {
                    visited[a] = 1;
                    order.push_back(a);
                }
        }
    }
    std::reverse(order.begin(), order.end());
    std::vector<int64_t> inv((size_t)nv);
    for (int64_t i = 0; i < nv; ++i) {
        vperm[i] = order[i];
        inv[order[i]] = i;
    }
    std::vector<int64_t> key((size_t)ne);
    for (int64_t e = 0; e < ne; ++e) {
        int64_t k = INT64_MAX;
        for (int j = 0; j < n; ++j) k = std::min(k, inv[conn[e * n + j]]);
        key[e] = k;
    }
    std::vector<int64_t> ep((size_t)ne);
    std::iota(ep.begin(), ep.end(), 0);
    std::stable_sort(ep.begin(), ep.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    std::copy(ep.begin(), ep.end(), eperm);
}

// First fit over colors kept as per-node bitmasks: bit c of node v is set when color c
// already contains a set touching v, so the first zero bit of the OR over a set's nodes is
// the first color whose node set is disjoint from it.
std::vector<std::vector<int32_t>> greedy_coloring(int64_t nsets, const int64_t* off, const int32_t* nodes) {
    std::vector<std::vector<int32_t>> colors;
    int32_t max_node = -1;
    for (int64_t k = 0; k < (nsets ? off[nsets] : 0); ++k) max_node = std::max(max_node, nodes[k]);
    const int64_t nn = (int64_t)max_node + 1;
    int words = 1;
    std::vector<uint64_t> mask((size_t)nn * words, 0);
    std::vector<uint64_t> acc;
    for (int64_t s = 0; s < nsets; ++s) {
        acc.assign(words, 0);
        for (int64_t k = off[s]; k < off[s + 1]; ++k)
            for (int w = 0; w < words; ++w) acc[w] |= mask[(size_t)nodes[k] * words + w];
        int c = -1;
        for (int w = 0; w < words && c < 0; ++w)
            if (~acc[w]) c = 64 * w + __builtin_ctzll(~acc[w]);
        if (c < 0) c = 64 * words;
        if (c >= (int)colors.size()) colors.resize(c + 1);
        if (c >= 64 * words) {
            const int nw = words * 2;
            std::vector<uint64_t> m2((size_t)nn * nw, 0);
            for (int64_t v = 0; v < nn; ++v)
                for (int w = 0; w < words; ++w) m2[(size_t)v * nw + w] = mask[(size_t)v * words + w];
            mask.swap(m2);
            words = nw;
        }
        colors[c].push_back((int32_t)s);
        for (int64_t k = off[s]; k < off[s + 1]; ++k)
            mask[(size_t)nodes[k] * words + c / 64] |= (1ull << (c % 64));
    }
    return colors;
}

void BlockScatter::build(const NodePattern& pat, int64_t nb, const int64_t* off, const int32_t* nodes,
                         const std::vector<std::vector<int32_t>>* color_lists, int ncolor_lists) {
    nblocks = nb;
    m.resize(nb);
    val_off.resize(nb + 1);
    slot_off.resize(nb + 1);
    val_off[0] = slot_off[0] = 0;
    max_m = 0;
    for (int64_t b = 0; b < nb; ++b) {
        m[b] = (int32_t)(off[b + 1] - off[b]);
        max_m = std::max(max_m, m[b]);
        val_off[b + 1] = val_off[b] + 9 * (int64_t)m[b] * m[b];
        slot_off[b + 1] = slot_off[b] + (int64_t)m[b] * m[b];
    }
    slot.resize(slot_off[nb]);
    for (int64_t b = 0; b < nb; ++b) {
        const int32_t* g = nodes + off[b];
        for (int a = 0; a < m[b]; ++a)
            for (int c = 0; c < m[b]; ++c) {
                const int32_t k = pat.find(g[a], g[c]);
                if (k < 0) throw Error(HOEFEM_ERROR_PATTERN, "block entry missing from CSR pattern");
                slot[slot_off[b] + (int64_t)a * m[b] + c] = k;
            }
    }
    inc_ptr.assign(pat.num_nodes + 1, 0);
    for (int l = 0; l < ncolor_lists; ++l)
        for (const auto& color : color_lists[l])
            for (int32_t b : color)
                for (int a = 0; a < m[b]; ++a) inc_ptr[nodes[off[b] + a] + 1]++;
    max_inc = 0;
    for (int64_t v = 0; v < pat.num_nodes; ++v) max_inc = std::max(max_inc, inc_ptr[v + 1]);
    for (int64_t v = 0; v < pat.num_nodes; ++v) inc_ptr[v + 1] += inc_ptr[v];
    inc_blk.resize(inc_ptr[pat.num_nodes]);
    inc_loc.resize(inc_ptr[pat.num_nodes]);
    std::vector<int32_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int l = 0; l < ncolor_lists; ++l)
        for (const auto& color : color_lists[l])
            for (int32_t b : color)
                for (int a = 0; a < m[b]; ++a) {
                    const int32_t v = nodes[off[b] + a];
                    inc_blk[fill[v]] = b;
                    inc_loc[fill[v]] = a;
                    fill[v]++;
                }
}

void BlockScatter::upload(hipStream_t s) {
    d_m.upload(m, s);
    d_slot.upload(slot, s);
    d_inc_ptr.upload(inc_ptr, s);
    d_inc_blk.upload(inc_blk, s);
    d_inc_loc.upload(inc_loc, s);
    d_val_off.upload(val_off, s);
    d_slot_off.upload(slot_off, s);
}

}  // namespace hoefem
