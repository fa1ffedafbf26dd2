// CPU FP64 oracle (test infrastructure only): a plain restatement of the reference's
// implicit elastodynamics step. Every function cites the reference file:line it follows.
// Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to verify the HIP library;
// the product library never links it.
//
// Parity status of each part (how it is pinned; see tests/test_oracle_*.py):
//  - SNH stress / dP/dF / contraction: finite differences, batch-vs-single, analytic
//    eigen-decomposition, errata Lame triples (reference tests).
//  - element bases: Lagrange property, partition of unity, FD gradients, exact
//    polynomial reproduction.
//  - assembly: brute-force dense assembly, constant-displacement zero force,
//    K = -df/du by FD, mass * 1 = rho * volume, exact pattern vs brute force.
//  - coloring: first-fit rule checked by a brute-force re-implementation + disjointness.
//  - Schwarz block inverse: numpy.linalg.eigh reconstruction (the reference's LAPACK
//    ?syev call is third-party and has no reference test).
//  - CG / Newton / Backward Euler: the reference's 3x3 CG cases, single-iteration
//    Newton on a linear system, closed-form Backward Euler on the linear mock.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- 3x3 helpers
// column-major: m[r + 3c]
inline double det3(const double* m) {
    // nalgebra's 3x3 determinant by cofactor expansion along the first row
    const double m11 = m[0], m12 = m[3], m13 = m[6];
    const double m21 = m[1], m22 = m[4], m23 = m[7];
    const double m31 = m[2], m32 = m[5], m33 = m[8];
    const double minor_1 = m22 * m33 - m32 * m23;
    const double minor_2 = m21 * m33 - m31 * m23;
    const double minor_3 = m21 * m32 - m31 * m22;
    return m11 * minor_1 - m12 * minor_2 + m13 * minor_3;
}

inline bool inv3(const double* m, double* out) {
    const double m11 = m[0], m12 = m[3], m13 = m[6];
    const double m21 = m[1], m22 = m[4], m23 = m[7];
    const double m31 = m[2], m32 = m[5], m33 = m[8];
    const double minor_m12_m23 = m22 * m33 - m32 * m23;
    const double minor_m11_m23 = m21 * m33 - m31 * m23;
    const double minor_m11_m22 = m21 * m32 - m31 * m22;
    const double det = m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
    if (det == 0.0) return false;
    out[0 + 3 * 0] = minor_m12_m23 / det;
    out[0 + 3 * 1] = (m13 * m32 - m33 * m12) / det;
    out[0 + 3 * 2] = (m12 * m23 - m22 * m13) / det;
    out[1 + 3 * 0] = -minor_m11_m23 / det;
    out[1 + 3 * 1] = (m11 * m33 - m31 * m13) / det;
    out[1 + 3 * 2] = (m13 * m21 - m23 * m11) / det;
    out[2 + 3 * 0] = minor_m11_m22 / det;
    out[2 + 3 * 1] = (m12 * m31 - m32 * m11) / det;
    out[2 + 3 * 2] = (m11 * m22 - m21 * m12) / det;
    return true;
}

inline void transpose3(const double* a, double* out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[c + 3 * r] = a[r + 3 * c];
}

inline void matmul3(const double* a, const double* b, double* out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[r + 3 * k] * b[k + 3 * c];
            out[r + 3 * c] = s;
        }
}

// ---------------------------------------------------------------- materials
// fenris/src/solid/materials.rs:243-253 (d = 3)
inline void snh_reparam(double mu, double lambda, double* mu_r, double* lambda_r) {
    const double d = 3.0;
    *mu_r = (d + 1.0) / d * mu;
    *lambda_r = lambda + (1.0 - 2.0 / (d * (d + 1.0))) * mu;
}

// fenris/src/solid/materials.rs:299-322 (reparametrized mu/lambda passed in)
bool snh_stress_r(double mu, double lambda, const double* F, double* P) {
    const double d = 3.0;
    const double alpha = 1.0 + mu / lambda * (d / (d + 1.0));
    double Ft[9], C[9];
    transpose3(F, Ft);
    matmul3(Ft, F, C);
    const double I_C = C[0] + C[4] + C[8];
    const double J = det3(F);
    double Finv[9];
    if (!inv3(F, Finv)) return false;
    // dJ_dF = F^{-T} * J
    const double s1 = mu * (1.0 - 1.0 / (I_C + 1.0));
    const double s2 = lambda * (J - alpha);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double dJ = Finv[c + 3 * r] * J;
            P[r + 3 * c] = F[r + 3 * c] * s1 + dJ * s2;
        }
    return true;
}

// fenris/src/solid/materials.rs:506-557 + add_volume_hessian_3d :763-776
bool snh_dpdf_r(double mu, double lambda, const double* F, double* dp) {
    const double d = 3.0;
    const double alpha = 1.0 + mu / lambda * (d / (d + 1.0));
    double Ft[9], C[9];
    transpose3(F, Ft);
    matmul3(Ft, F, C);
    const double I_C = C[0] + C[4] + C[8];
    const double J = det3(F);
    const double beta = I_C + 1.0;
    const double beta2 = beta * beta;
    double Finv[9];
    if (!inv3(F, Finv)) return false;
    double f[9], g[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            f[3 * c + r] = F[r + 3 * c];
            g[3 * c + r] = Finv[c + 3 * r] * J;  // vec(J F^{-T})
        }
    for (int i = 0; i < 81; ++i) dp[i] = 0.0;
    const double diag = mu * (1.0 - 1.0 / beta);
    for (int i = 0; i < 9; ++i) dp[i + 9 * i] = diag;
    const double cf = mu * 2.0 / beta2;
    // ger: dp += a * x y^T (nalgebra ger with beta = 1)
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r) dp[r + 9 * c] += cf * f[r] * f[c];
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r) dp[r + 9 * c] += lambda * g[r] * g[c];
    const double scale = lambda * (J - alpha);
    // cross_product_matrix(x) = [[0,-x2,x1],[x2,0,-x0],[-x1,x0,0]] (fenris/src/util.rs:258-260)
    double hat[3][9];
    for (int k = 0; k < 3; ++k) {
        const double x0 = F[0 + 3 * k], x1 = F[1 + 3 * k], x2 = F[2 + 3 * k];
        const double m[9] = {0.0, x2, -x1, -x2, 0.0, x0, x1, -x0, 0.0};  // column-major
        for (int i = 0; i < 9; ++i) hat[k][i] = m[i] * scale;
    }
    auto add_block = [&](int r0, int c0, const double* h, double sgn) {
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                if (sgn > 0)
                    dp[(r0 + r) + 9 * (c0 + c)] += h[r + 3 * c];
                else
                    dp[(r0 + r) + 9 * (c0 + c)] -= h[r + 3 * c];
            }
    };
    add_block(3, 0, hat[2], +1);
    add_block(6, 0, hat[1], -1);
    add_block(0, 3, hat[2], -1);
    add_block(6, 3, hat[0], +1);
    add_block(0, 6, hat[1], +1);
    add_block(3, 6, hat[0], -1);
    return true;
}

// fenris/src/solid/materials.rs:690-743 (adds into out, column-major (3n)^2)
void contract_stiffness(const double* dp, int n, const double* a, double* out) {
    const int dim = 3 * n;
    for (int J = 0; J < n; ++J) {
        for (int I = J; I < n; ++I) {
            double res[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const double* aI = a + 3 * I;
            const double* aJ = a + 3 * J;
            for (int l = 0; l < 3; ++l)
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j)
                        for (int i = 0; i < 3; ++i) {
                            const int col = 3 * l + j;
                            const int row = 3 * k + i;
                            res[i + 3 * j] += dp[row + 9 * col] * aI[k] * aJ[l];
                        }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) out[(3 * I + i) + dim * (3 * J + j)] += res[i + 3 * j];
            if (I != J) {
                for (int j = 0; j < 3; ++j)
                    for (int i = 0; i < 3; ++i) out[(3 * J + j) + dim * (3 * I + i)] += res[i + 3 * j];
            }
        }
    }
}

// ---------------------------------------------------------------- elements
int nodes_per(int t) {
    switch (t) {
        case ORC_TET4: return 4;
        case ORC_TET10: return 10;
        case ORC_TET20: return 20;
        case ORC_HEX8: return 8;
        case ORC_HEX20: return 20;
        case ORC_HEX27: return 27;
    }
    return -1;
}
int geom_nodes(int t) {
    switch (t) {
        case ORC_TET4: case ORC_TET10: case ORC_TET20: return 4;
        case ORC_HEX8: case ORC_HEX20: case ORC_HEX27: return 8;
    }
    return -1;
}
bool is_tet(int t) { return t == ORC_TET4 || t == ORC_TET10 || t == ORC_TET20; }

// fenris/src/element.rs:986-1007 (Tet4 on the [-1,1] reference simplex)
inline void tet4_basis(const double* xi, double* psi) {
    psi[0] = -0.5 * xi[0] - 0.5 * xi[1] - 0.5 * xi[2] - 0.5;
    psi[1] = 0.5 * xi[0] + 0.5;
    psi[2] = 0.5 * xi[1] + 0.5;
    psi[3] = 0.5 * xi[2] + 0.5;
}
const double TET4_G[4][3] = {{-0.5, -0.5, -0.5}, {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}};

const double HEX8_NODES[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
// fenris/src/element.rs:1250-1298 (Hex27 / Hex20 reference node coordinates)
const double HEX27_NODES[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {-1, 0, -1}, {-1, -1, 0}, {1, 0, -1}, {1, -1, 0}, {0, 1, -1},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1},  {-1, 0, 1},  {1, 0, 1},  {0, 1, 1},  {0, 0, -1},
    {0, -1, 0},   {-1, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},  {0, 0, 0}};

inline double lin1(double a, double x) { return (1.0 + a * x) / 2.0; }
inline double lin1g(double a) { return a / 2.0; }
// fenris/src/element.rs:1089-1121
inline double quad1(double a, double x) {
    const double a2 = a * a;
    const double x2 = x * x;
    return (3.0 / 2.0 * a2 - 1.0) * x2 + 0.5 * a * x + 1.0 - a2;
}
inline double quad1g(double a, double x) {
    const double a2 = a * a;
    return 2.0 * (3.0 / 2.0 * a2 - 1.0) * x + 0.5 * a;
}

void basis(int t, const double* xi, double* phi) {
    switch (t) {
        case ORC_TET4: tet4_basis(xi, phi); return;
        case ORC_TET10: {
            // fenris/src/element.rs:1906-1922
            double p[4];
            tet4_basis(xi, p);
            phi[0] = p[0] * (2.0 * p[0] - 1.0);
            phi[1] = p[1] * (2.0 * p[1] - 1.0);
            phi[2] = p[2] * (2.0 * p[2] - 1.0);
            phi[3] = p[3] * (2.0 * p[3] - 1.0);
            phi[4] = 4.0 * p[0] * p[1];
            phi[5] = 4.0 * p[1] * p[2];
            phi[6] = 4.0 * p[0] * p[2];
            phi[7] = 4.0 * p[0] * p[3];
            phi[8] = 4.0 * p[2] * p[3];
            phi[9] = 4.0 * p[1] * p[3];
            return;
        }
        case ORC_TET20: {
            // fenris/src/element.rs:2061-2111
            double p[4];
            tet4_basis(xi, p);
            auto edge = [&](int c, int o) { return (9.0 / 2.0) * p[c] * p[o] * (3.0 * p[c] - 1.0); };
            auto face = [&](int a, int b, int c) { return 27.0 * p[a] * p[b] * p[c]; };
            for (int i = 0; i < 4; ++i) phi[i] = 0.5 * p[i] * (3.0 * p[i] - 1.0) * (3.0 * p[i] - 2.0);
            phi[4] = edge(0, 1); phi[5] = edge(1, 0);
            phi[6] = edge(0, 2); phi[7] = edge(2, 0);
            phi[8] = edge(0, 3); phi[9] = edge(3, 0);
            phi[10] = edge(1, 2); phi[11] = edge(2, 1);
            phi[12] = edge(1, 3); phi[13] = edge(3, 1);
            phi[14] = edge(2, 3); phi[15] = edge(3, 2);
            phi[16] = face(0, 1, 2); phi[17] = face(0, 1, 3);
            phi[18] = face(0, 2, 3); phi[19] = face(1, 2, 3);
            return;
        }
        case ORC_HEX8:
            for (int i = 0; i < 8; ++i)
                phi[i] = lin1(HEX8_NODES[i][0], xi[0]) * lin1(HEX8_NODES[i][1], xi[1]) * lin1(HEX8_NODES[i][2], xi[2]);
            return;
        case ORC_HEX27:
            // fenris/src/element.rs:1309-1355
            for (int i = 0; i < 27; ++i)
                phi[i] = quad1(HEX27_NODES[i][0], xi[0]) * quad1(HEX27_NODES[i][1], xi[1]) *
                         quad1(HEX27_NODES[i][2], xi[2]);
            return;
        case ORC_HEX20: {
            // fenris/src/element.rs:1503-1552
            for (int i = 0; i < 8; ++i) {
                const double a = HEX27_NODES[i][0], b = HEX27_NODES[i][1], c = HEX27_NODES[i][2];
                phi[i] = (1.0 / 8.0) * (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]) *
                         (a * xi[0] + b * xi[1] + c * xi[2] - 2.0);
            }
            for (int i = 8; i < 20; ++i) {
                const double a = HEX27_NODES[i][0], b = HEX27_NODES[i][1], c = HEX27_NODES[i][2];
                const double a2 = a * a, b2 = b * b, c2 = c * c;
                phi[i] = (1.0 / 4.0) * (1.0 - (1.0 - a2) * xi[0] * xi[0]) * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
                         (1.0 - (1.0 - c2) * xi[2] * xi[2]) * (1.0 + a * xi[0]) * (1.0 + b * xi[1]) *
                         (1.0 + c * xi[2]);
            }
            return;
        }
    }
}

// gradients: grad[3*I + k] = d phi_I / d xi_k  (a 3 x n column-major matrix)
void gradients(int t, const double* xi, double* grad) {
    switch (t) {
        case ORC_TET4:
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 3; ++k) grad[3 * i + k] = TET4_G[i][k];
            return;
        case ORC_TET10: {
            // fenris/src/element.rs:1925-1951
            double p[4];
            tet4_basis(xi, p);
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 3; ++k) grad[3 * i + k] = TET4_G[i][k] * (4.0 * p[i] - 1.0);
            const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
            for (int e = 0; e < 6; ++e) {
                const int i = edges[e][0], j = edges[e][1];
                for (int k = 0; k < 3; ++k)
                    grad[3 * (4 + e) + k] = TET4_G[i][k] * (4.0 * p[j]) + TET4_G[j][k] * (4.0 * p[i]);
            }
            return;
        }
        case ORC_TET20: {
            // fenris/src/element.rs:2114-2175
            double p[4];
            tet4_basis(xi, p);
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 3; ++k) grad[3 * i + k] = TET4_G[i][k] * 0.5 * (27.0 * p[i] * p[i] - 18.0 * p[i] + 2.0);
            const int ed[12][2] = {{0, 1}, {1, 0}, {0, 2}, {2, 0}, {0, 3}, {3, 0},
                                   {1, 2}, {2, 1}, {1, 3}, {3, 1}, {2, 3}, {3, 2}};
            for (int e = 0; e < 12; ++e) {
                const int a = ed[e][0], b = ed[e][1];
                const double pa = p[a], pb = p[b];
                for (int k = 0; k < 3; ++k)
                    grad[3 * (4 + e) + k] =
                        (TET4_G[a][k] * (pb * (6.0 * pa - 1.0)) + TET4_G[b][k] * (pa * (3.0 * pa - 1.0))) * (9.0 / 2.0);
            }
            const int fc[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
            for (int f = 0; f < 4; ++f) {
                const int a = fc[f][0], b = fc[f][1], c = fc[f][2];
                for (int k = 0; k < 3; ++k)
                    grad[3 * (16 + f) + k] =
                        (TET4_G[a][k] * p[b] * p[c] + TET4_G[b][k] * p[a] * p[c] + TET4_G[c][k] * p[a] * p[b]) * 27.0;
            }
            return;
        }
        case ORC_HEX8:
            for (int i = 0; i < 8; ++i) {
                const double a = HEX8_NODES[i][0], b = HEX8_NODES[i][1], c = HEX8_NODES[i][2];
                grad[3 * i + 0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
                grad[3 * i + 1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
                grad[3 * i + 2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
            }
            return;
        case ORC_HEX27:
            for (int i = 0; i < 27; ++i) {
                const double a = HEX27_NODES[i][0], b = HEX27_NODES[i][1], c = HEX27_NODES[i][2];
                grad[3 * i + 0] = quad1g(a, xi[0]) * quad1(b, xi[1]) * quad1(c, xi[2]);
                grad[3 * i + 1] = quad1(a, xi[0]) * quad1g(b, xi[1]) * quad1(c, xi[2]);
                grad[3 * i + 2] = quad1(a, xi[0]) * quad1(b, xi[1]) * quad1g(c, xi[2]);
            }
            return;
        case ORC_HEX20: {
            // fenris/src/element.rs:1556-1625
            for (int i = 0; i < 8; ++i) {
                const double a = HEX27_NODES[i][0], b = HEX27_NODES[i][1], c = HEX27_NODES[i][2];
                const double f = a * xi[0] + b * xi[1] + c * xi[2] - 2.0;
                const double g = (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]);
                const double s = 1.0 / 8.0;
                grad[3 * i + 0] = s * (a * g + f * a * (1.0 + b * xi[1]) * (1.0 + c * xi[2]));
                grad[3 * i + 1] = s * (b * g + f * b * (1.0 + a * xi[0]) * (1.0 + c * xi[2]));
                grad[3 * i + 2] = s * (c * g + f * c * (1.0 + a * xi[0]) * (1.0 + b * xi[1]));
            }
            for (int i = 8; i < 20; ++i) {
                const double a = HEX27_NODES[i][0], b = HEX27_NODES[i][1], c = HEX27_NODES[i][2];
                const double a2 = a * a, b2 = b * b, c2 = c * c;
                const double h = (1.0 - (1.0 - a2) * xi[0] * xi[0]) * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
                                 (1.0 - (1.0 - c2) * xi[2] * xi[2]);
                const double g = (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]);
                const double s = 1.0 / 4.0;
                const double dh0 = -2.0 * (1.0 - a2) * xi[0] * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
                                   (1.0 - (1.0 - c2) * xi[2] * xi[2]);
                const double dh1 = -2.0 * (1.0 - b2) * xi[1] * (1.0 - (1.0 - a2) * xi[0] * xi[0]) *
                                   (1.0 - (1.0 - c2) * xi[2] * xi[2]);
                const double dh2 = -2.0 * (1.0 - c2) * xi[2] * (1.0 - (1.0 - a2) * xi[0] * xi[0]) *
                                   (1.0 - (1.0 - b2) * xi[1] * xi[1]);
                grad[3 * i + 0] = s * (dh0 * g + h * a * (1.0 + b * xi[1]) * (1.0 + c * xi[2]));
                grad[3 * i + 1] = s * (dh1 * g + h * b * (1.0 + a * xi[0]) * (1.0 + c * xi[2]));
                grad[3 * i + 2] = s * (dh2 * g + h * c * (1.0 + a * xi[0]) * (1.0 + b * xi[1]));
            }
            return;
        }
    }
}

// Tet4Element::reference_jacobian / Hex8Element::reference_jacobian: J = X * G^T
// (fenris/src/element.rs:1013-1019, :1189-1195). nodes: [g*3] geometry node coordinates.
void reference_jacobian(int t, const double* nodes, const double* xi, double* J) {
    double g[8 * 3];
    const int gn = geom_nodes(t);
    gradients(is_tet(t) ? ORC_TET4 : ORC_HEX8, xi, g);
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < gn; ++k) s += nodes[3 * k + i] * g[3 * k + c];
            J[i + 3 * c] = s;
        }
}

// ---------------------------------------------------------------- greedy coloring
// paradis/src/coloring.rs:48-72: element e joins the first color whose node set is
// disjoint from e's nodes, otherwise opens a new color.
std::vector<std::vector<int32_t>> greedy_coloring(int64_t nsets, const int64_t* off, const int32_t* nodes) {
    std::vector<std::vector<int32_t>> colors;        // labels per color
    std::vector<std::vector<int32_t>> color_nodes;   // sorted node set per color
    for (int64_t s = 0; s < nsets; ++s) {
        std::vector<int32_t> sub(nodes + off[s], nodes + off[s + 1]);
        std::sort(sub.begin(), sub.end());
        bool placed = false;
        for (size_t c = 0; c < colors.size() && !placed; ++c) {
            auto& cn = color_nodes[c];
            bool disjoint = true;
            for (int32_t v : sub)
                if (std::binary_search(cn.begin(), cn.end(), v)) {
                    disjoint = false;
                    break;
                }
            if (disjoint) {
                std::vector<int32_t> merged;
                merged.reserve(cn.size() + sub.size());
                std::merge(cn.begin(), cn.end(), sub.begin(), sub.end(), std::back_inserter(merged));
                cn.swap(merged);
                colors[c].push_back((int32_t)s);
                placed = true;
            }
        }
        if (!placed) {
            colors.push_back({(int32_t)s});
            color_nodes.push_back(sub);
        }
    }
    return colors;
}

// ---------------------------------------------------------------- CSR pattern
struct Csr {
    int64_t n = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<double> val;
    int64_t find(int64_t r, int64_t c) const {
        auto b = col.begin() + rowptr[r], e = col.begin() + rowptr[r + 1];
        auto it = std::lower_bound(b, e, (int32_t)c);
        if (it == e || *it != c) return -1;
        return it - col.begin();
    }
};

// CsrParAssembler::assemble_pattern (fenris/src/assembly.rs:177-249): sorted unique
// coordinates of all element node pairs x sdim^2. Done at node level and expanded,
// which yields the same set.
void build_pattern(int64_t num_nodes, int64_t nsets, const int64_t* off, const int32_t* nodes, int sdim,
                   std::vector<int32_t>& diag_nodes, Csr& out) {
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (int64_t s = 0; s < nsets; ++s)
        for (int64_t a = off[s]; a < off[s + 1]; ++a)
            for (int64_t b = off[s]; b < off[s + 1]; ++b) pairs.emplace_back(nodes[a], nodes[b]);
    for (int32_t v : diag_nodes) pairs.emplace_back(v, -1 - v);  // marker: diagonal-only
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    // per node: list of neighbor nodes (full 3x3 blocks) and diagonal-only flag
    std::vector<std::vector<int32_t>> nbr(num_nodes);
    std::vector<char> diag_only(num_nodes, 0);
    for (auto& p : pairs) {
        if (p.second < 0)
            diag_only[p.first] = 1;
        else
            nbr[p.first].push_back(p.second);
    }
    out.n = num_nodes * sdim;
    out.rowptr.assign(out.n + 1, 0);
    out.col.clear();
    for (int64_t a = 0; a < num_nodes; ++a) {
        for (int i = 0; i < sdim; ++i) {
            const int64_t row = a * sdim + i;
            std::vector<int32_t> cols;
            for (int32_t b : nbr[a])
                for (int j = 0; j < sdim; ++j) cols.push_back((int32_t)(b * sdim + j));
            if (diag_only[a]) cols.push_back((int32_t)row);
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            out.col.insert(out.col.end(), cols.begin(), cols.end());
            out.rowptr[row + 1] = (int32_t)out.col.size();
        }
    }
    out.val.assign(out.col.size(), 0.0);
}

// add_element_row_to_csr_row (fenris/src/assembly.rs:1042-1075): forward linear search
// through the (sorted) CSR row for each element column in sorted-node order.
bool add_element_row(Csr& csr, int64_t row, const int32_t* gnodes, const std::vector<int>& perm, int sdim,
                     const double* Ke, int dim, int local_row) {
    int64_t pos = csr.rowptr[row];
    const int64_t end = csr.rowptr[row + 1];
    for (int nl : perm) {
        const int64_t g = gnodes[nl];
        for (int i = 0; i < sdim; ++i) {
            const int local_col = sdim * nl + i;
            const int64_t gcol = sdim * g + i;
            while (pos < end && csr.col[pos] != gcol) ++pos;
            if (pos == end) return false;
            csr.val[pos] += Ke[local_row + (int64_t)dim * local_col];
            ++pos;
        }
    }
    return true;
}

// ---------------------------------------------------------------- symmetric eigen (cyclic Jacobi)
// Textbook cyclic Jacobi rotations (Golub & Van Loan, "Matrix Computations", sec. 8.5).
// a: column-major m x m (overwritten copy), evecs column-major.
void jacobi_eigen(int m, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& v) {
    v.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) v[i + (size_t)m * i] = 1.0;
    auto A = [&](int r, int c) -> double& { return a[r + (size_t)m * c]; };
    double total = 0.0;
    for (int i = 0; i < m * m; ++i) total += a[i] * a[i];
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 1; q < m; ++q)
            for (int p = 0; p < q; ++p) off += A(p, q) * A(p, q);
        if (off <= 1e-32 * total || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p) {
            for (int q = p + 1; q < m; ++q) {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double app = A(p, p), aqq = A(q, q);
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0);
                const double s = t * c;
                // A <- R^T A R with R the rotation in the (p, q) plane
                for (int k = 0; k < m; ++k) {
                    const double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - s * aqk;
                    A(q, k) = s * apk + c * aqk;
                }
                A(p, q) = 0.0;
                A(q, p) = 0.0;
                for (int k = 0; k < m; ++k) {
                    const double vkp = v[k + (size_t)m * p], vkq = v[k + (size_t)m * q];
                    v[k + (size_t)m * p] = c * vkp - s * vkq;
                    v[k + (size_t)m * q] = s * vkp + c * vkq;
                }
            }
        }
    }
    evals.resize(m);
    for (int i = 0; i < m; ++i) evals[i] = A(i, i);
}

// ---------------------------------------------------------------- symmetric eigen (tridiagonal QR)
// For large blocks (cyclic Jacobi costs ~10 m^3 per sweep): Householder reduction to tridiagonal
// form, then the implicit symmetric QR iteration with Wilkinson shifts and deflation (Golub & Van
// Loan, "Matrix Computations", algorithms 8.3.1 and 8.3.3), rotations accumulated into the
// eigenvectors. a: column-major m x m (overwritten copy), evecs column-major.
void tridiagonal_qr_eigen(int m, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& q) {
    const size_t M = (size_t)m;
    auto A = [&](int r, int c) -> double& { return a[r + M * c]; };
    q.assign(M * M, 0.0);
    for (int i = 0; i < m; ++i) q[i + M * i] = 1.0;
    std::vector<double> v(M), p(M), w(M);
    // Householder: for column k, zero A(k+2:, k) with H = I - 2 v v^T acting on rows/cols k+1..m-1
    for (int k = 0; k + 2 < m; ++k) {
        double xn = 0.0;
        for (int i = k + 1; i < m; ++i) xn += A(i, k) * A(i, k);
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const double x0 = A(k + 1, k);
        const double alpha = x0 > 0.0 ? -xn : xn;
        double vn = 0.0;
        for (int i = k + 1; i < m; ++i) {
            v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
            vn += v[i] * v[i];
        }
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < m; ++i) v[i] /= vn;
        // trailing block: A <- H A H = A - 2 (v q^T + q v^T), p = A v, q = p - (v^T p) v
        double K = 0.0;
        for (int i = k + 1; i < m; ++i) {
            double sum = 0.0;
            for (int j = k + 1; j < m; ++j) sum += A(i, j) * v[j];
            p[i] = sum;
            K += v[i] * sum;
        }
        for (int i = k + 1; i < m; ++i) w[i] = p[i] - K * v[i];
        for (int j = k + 1; j < m; ++j)
            for (int i = k + 1; i < m; ++i) A(i, j) -= 2.0 * (v[i] * w[j] + w[i] * v[j]);
        A(k + 1, k) = alpha;
        A(k, k + 1) = alpha;
        for (int i = k + 2; i < m; ++i) A(i, k) = A(k, i) = 0.0;
        // Q <- Q H
        for (int r = 0; r < m; ++r) {
            double sum = 0.0;
            for (int i = k + 1; i < m; ++i) sum += q[r + M * i] * v[i];
            for (int i = k + 1; i < m; ++i) q[r + M * i] -= 2.0 * sum * v[i];
        }
    }
    std::vector<double> d(M), e(M > 0 ? M - 1 : 0);
    for (int i = 0; i < m; ++i) d[i] = A(i, i);
    for (int i = 0; i + 1 < m; ++i) e[i] = A(i + 1, i);
    const double eps = 2.220446049250313e-16;
    for (int iter = 0; iter < 60 * m; ++iter) {
        for (int i = 0; i + 1 < m; ++i)
            if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
        int hi = m - 1;
        while (hi > 0 && e[hi - 1] == 0.0) --hi;
        if (hi == 0) break;
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        // Wilkinson shift from the trailing 2 x 2 of the unreduced block [lo, hi]
        const double dd = 0.5 * (d[hi - 1] - d[hi]), bb = e[hi - 1];
        const double mu = d[hi] - bb * bb / (dd + (dd >= 0.0 ? 1.0 : -1.0) * std::hypot(dd, bb));
        double x = d[lo] - mu, z = e[lo];
        for (int k = lo; k < hi; ++k) {
            const double r = std::hypot(x, z);
            const double c = r == 0.0 ? 1.0 : x / r, sn = r == 0.0 ? 0.0 : z / r;
            if (k > lo) e[k - 1] = r;
            const double ak = d[k], ak1 = d[k + 1], bk = e[k];
            d[k] = c * c * ak + 2.0 * c * sn * bk + sn * sn * ak1;
            d[k + 1] = sn * sn * ak - 2.0 * c * sn * bk + c * c * ak1;
            e[k] = c * sn * (ak1 - ak) + (c * c - sn * sn) * bk;
            if (k + 1 < hi) {
                z = sn * e[k + 1];
                e[k + 1] = c * e[k + 1];
                x = e[k];
            }
            for (int r2 = 0; r2 < m; ++r2) {
                const double qk = q[r2 + M * k], qk1 = q[r2 + M * (k + 1)];
                q[r2 + M * k] = c * qk + sn * qk1;
                q[r2 + M * (k + 1)] = -sn * qk + c * qk1;
            }
        }
    }
    evals = d;
}

// invert_element_block (simulation_toolbox/src/fem/schwarz_precond.rs:329-342):
// V * diag(1/max(|lambda|, 1e-12 max|lambda|)) * V^T (negative eigenvalues flipped).
// Blocks above 96 x 96 use the tridiagonal QR eigensolver, smaller ones cyclic Jacobi.
constexpr int ORC_JACOBI_MAX_DIM = 96;

void invert_block(int m, const double* block, double* out) {
    std::vector<double> a(block, block + (size_t)m * m), ev, v;
    if (m > ORC_JACOBI_MAX_DIM)
        tridiagonal_qr_eigen(m, a, ev, v);
    else
        jacobi_eigen(m, a, ev, v);
    double maxabs = 0.0;
    for (double e : ev) maxabs = std::max(maxabs, std::fabs(e));
    const double lower = 1e-12 * maxabs;
    std::vector<double> inv(m);
    for (int i = 0; i < m; ++i) inv[i] = std::fabs(ev[i]) >= lower ? 1.0 / std::fabs(ev[i]) : 1.0 / lower;
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += v[r + (size_t)m * k] * inv[k] * v[c + (size_t)m * k];
            out[r + (size_t)m * c] = s;
        }
}

// ---------------------------------------------------------------- SVD-based material variants
// F = U diag(sigma) V^T with sigma >= 0 and U, V orthogonal (determinant +-1), by one-sided
// Jacobi (Hestenes) rotations on the columns of F. It stands in for nalgebra 0.21.1's
// Matrix::svd(true, true) (Golub-Kahan bidiagonalization + implicit QR), a dependency of the
// reference that is not vendored: the factors of both are unique up to the order and the
// joint signs of the singular triplets, and every use below is invariant under both (ties
// between singular values aside).
void svd3(const double* F, double* U, double* sigma, double* V) {
    double B[9];
    for (int k = 0; k < 9; ++k) {
        B[k] = F[k];
        V[k] = (k % 4 == 0) ? 1.0 : 0.0;
    }
    const int PP[3] = {0, 0, 1}, QQ[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int k = 0; k < 3; ++k) {
            const int p = PP[k], q = QQ[k];
            double a = 0.0, b = 0.0, g = 0.0;
            for (int r = 0; r < 3; ++r) {
                a += B[r + 3 * p] * B[r + 3 * p];
                b += B[r + 3 * q] * B[r + 3 * q];
                g += B[r + 3 * p] * B[r + 3 * q];
            }
            if (g == 0.0 || std::fabs(g) <= 1e-16 * std::sqrt(a * b)) continue;
            rotated = true;
            const double zeta = (b - a) / (2.0 * g);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
            for (int r = 0; r < 3; ++r) {
                const double bp = B[r + 3 * p], bq = B[r + 3 * q];
                B[r + 3 * p] = c * bp - sn * bq;
                B[r + 3 * q] = sn * bp + c * bq;
                const double vp = V[r + 3 * p], vq = V[r + 3 * q];
                V[r + 3 * p] = c * vp - sn * vq;
                V[r + 3 * q] = sn * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    bool have[3];
    int nz = 0;
    for (int i = 0; i < 3; ++i) {
        sigma[i] = std::sqrt(B[3 * i] * B[3 * i] + B[3 * i + 1] * B[3 * i + 1] + B[3 * i + 2] * B[3 * i + 2]);
        have[i] = sigma[i] > 0.0;
        if (have[i]) {
            ++nz;
            for (int r = 0; r < 3; ++r) U[r + 3 * i] = B[r + 3 * i] / sigma[i];
        }
    }
    // complete U to an orthonormal basis where F has a null space
    if (nz == 0) {
        for (int k = 0; k < 9; ++k) U[k] = (k % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    for (int i = 0; i < 3 && nz < 3; ++i) {
        if (have[i]) continue;
        int a = -1, b = -1;
        for (int j = 0; j < 3; ++j)
            if (have[j]) (a < 0 ? a : b) = j;
        double* u = U + 3 * i;
        if (b >= 0) {
            const double* x = U + 3 * a;
            const double* y = U + 3 * b;
            u[0] = x[1] * y[2] - x[2] * y[1];
            u[1] = x[2] * y[0] - x[0] * y[2];
            u[2] = x[0] * y[1] - x[1] * y[0];
        } else {
            const double* x = U + 3 * a;
            int ax = 0;
            for (int r = 1; r < 3; ++r)
                if (std::fabs(x[r]) < std::fabs(x[ax])) ax = r;
            for (int r = 0; r < 3; ++r) u[r] = (r == ax ? 1.0 : 0.0) - x[ax] * x[r];
            const double n = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int r = 0; r < 3; ++r) u[r] /= n;
        }
        have[i] = true;
        ++nz;
    }
}

int imin3(const double* s) {
    int k = 0;
    for (int i = 1; i < 3; ++i)
        if (s[i] < s[k]) k = i;
    return k;
}

// rotation_svd (fenris/src/util.rs:68-94): U, V^T proper rotations, the sign moved to the
// smallest singular value.
void rotation_svd3(const double* F, double* U, double* s, double* Vt) {
    double V[9];
    svd3(F, U, s, V);
    transpose3(V, Vt);
    const int k = imin3(s);
    if (det3(U) < 0.0) {
        for (int r = 0; r < 3; ++r) U[r + 3 * k] = -U[r + 3 * k];
        s[k] = -s[k];
    }
    if (det3(Vt) < 0.0) {
        for (int c = 0; c < 3; ++c) Vt[k + 3 * c] = -Vt[k + 3 * c];
        s[k] = -s[k];
    }
}

// project_deformation_gradient (fenris/src/solid/materials.rs:876-915)
void project_deformation_gradient(const double* F, double threshold, double* Fout) {
    double U[9], s[3], V[9];
    svd3(F, U, s, V);
    const bool inversion = (det3(U) < 0.0) != (det3(V) < 0.0);  // det V = det V^T
    if (inversion) {
        const int k = imin3(s);
        for (int r = 0; r < 3; ++r) U[r + 3 * k] = -U[r + 3 * k];
        s[k] = -s[k];
    }
    bool correction = false;
    for (int i = 0; i < 3; ++i)
        if (s[i] < threshold) {
            s[i] = threshold;
            correction = true;
        }
    if (!correction) {
        for (int k = 0; k < 9; ++k) Fout[k] = F[k];
        return;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            double v = 0.0;
            for (int i = 0; i < 3; ++i) v += U[r + 3 * i] * s[i] * V[c + 3 * i];
            Fout[r + 3 * c] = v;
        }
}

// build_stable_neohookean_dp_df_eigen_3d (fenris/src/solid/materials.rs:561-687):
// evals[9], evecs column-major 9x9 (eigenvectors as columns, vec() ordering).
void snh_dpdf_eigen(double mu, double lambda, const double* F, double* evals, double* evecs) {
    double U[9], s[3], Vt[9];
    rotation_svd3(F, U, s, Vt);
    const double J = det3(F);
    double Ft[9], C[9];
    transpose3(F, Ft);
    matmul3(Ft, F, C);
    const double I_C = C[0] + C[4] + C[8];
    const double alpha = 1.0 + mu / lambda * (3.0 / 4.0);
    const double mu_t = mu * (1.0 - 1.0 / (I_C + 1.0));
    const double sv_scale = lambda * (J - alpha);
    for (int k = 0; k < 3; ++k) {
        evals[k] = s[k] * sv_scale + mu_t;
        evals[3 + k] = -s[k] * sv_scale + mu_t;
    }
    // D = (u_i v_j^T -+ u_j v_i^T) / sqrt(2); v_j = row j of V^T
    const double h = 1.0 / std::sqrt(2.0);
    auto outer = [&](int i, int j, double* out) {
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) out[r + 3 * c] = U[r + 3 * i] * (Vt[j + 3 * c] * h);
    };
    double u0v1[9], u0v2[9], u1v0[9], u1v2[9], u2v0[9], u2v1[9];
    outer(0, 1, u0v1);
    outer(0, 2, u0v2);
    outer(1, 0, u1v0);
    outer(1, 2, u1v2);
    outer(2, 0, u2v0);
    outer(2, 1, u2v1);
    for (int k = 0; k < 9; ++k) {
        evecs[k + 9 * 0] = u1v2[k] - u2v1[k];
        evecs[k + 9 * 1] = u0v2[k] - u2v0[k];
        evecs[k + 9 * 2] = u0v1[k] - u1v0[k];
        evecs[k + 9 * 3] = u1v2[k] + u2v1[k];
        evecs[k + 9 * 4] = u0v2[k] + u2v0[k];
        evecs[k + 9 * 5] = u0v1[k] + u1v0[k];
    }
    const double beta = I_C + 1.0, beta2 = beta * beta;
    const double gamma = lambda * (2.0 * J - alpha);
    std::vector<double> A(9), ev, E;
    A[0] = mu_t + 2.0 * s[0] * s[0] * mu / beta2 + s[1] * s[1] * s[2] * s[2] * lambda;
    A[4] = mu_t + 2.0 * s[1] * s[1] * mu / beta2 + s[0] * s[0] * s[2] * s[2] * lambda;
    A[8] = mu_t + 2.0 * s[2] * s[2] * mu / beta2 + s[0] * s[0] * s[1] * s[1] * lambda;
    A[3] = A[1] = s[2] * gamma + 2.0 * s[0] * s[1] * mu / beta2;
    A[6] = A[2] = s[1] * gamma + 2.0 * s[0] * s[2] * mu / beta2;
    A[7] = A[5] = s[0] * gamma + 2.0 * s[1] * s[2] * mu / beta2;
    jacobi_eigen(3, A, ev, E);
    for (int m = 0; m < 3; ++m) {
        evals[6 + m] = ev[m];
        // Q_m = U diag(e_m) V^T
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                double v = 0.0;
                for (int k = 0; k < 3; ++k) v += U[r + 3 * k] * E[k + 3 * m] * Vt[k + 3 * c];
                evecs[(r + 3 * c) + 9 * (6 + m)] = v;
            }
    }
}

// ProjectedStableNeoHookeanMaterial::contract_multiple_stress_tensors_into, 3D
// (fenris/src/solid/materials.rs:413-441): eigenvalues clamped at zero, recomposed.
void snh_dpdf_projected(double mu, double lambda, const double* F, double* dp) {
    double ev[9], Q[81];
    snh_dpdf_eigen(mu, lambda, F, ev, Q);
    for (int k = 0; k < 81; ++k) dp[k] = 0.0;
    for (int m = 0; m < 9; ++m) {
        const double l = std::max(0.0, ev[m]);
        for (int c = 0; c < 9; ++c)
            for (int r = 0; r < 9; ++r) dp[r + 9 * c] += Q[r + 9 * m] * l * Q[c + 9 * m];
    }
}

// StableNeoHookeanMaterial::contract_stress_tensor_with (fenris/src/solid/materials.rs:324-347),
// reparametrized mu / lambda passed in.
bool snh_contract_with_r(double mu, double lambda, const double* F, const double* a, const double* b, double* out) {
    const double alpha = 1.0 + mu / lambda * 0.75;
    double B[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) B[r + 3 * c] = a[r] * b[c];
    double Ft[9], C[9];
    transpose3(F, Ft);
    matmul3(Ft, F, C);
    const double I_C = C[0] + C[4] + C[8];
    const double J = det3(F);
    const double beta = I_C + 1.0;
    const double beta2 = beta * beta;
    const double G = 1.0 - 1.0 / beta;
    const double H = (J - alpha) * J;
    double Finv[9], FinvT[9];
    if (!inv3(F, Finv)) return false;
    transpose3(Finv, FinvT);
    const double trB = B[0] + B[4] + B[8];
    double FB[9], FBFt[9], t1[9], t2[9], Bt[9], t3[9], t4[9];
    matmul3(F, B, FB);
    matmul3(FB, Ft, FBFt);
    matmul3(FinvT, B, t1);
    matmul3(t1, Finv, t2);
    transpose3(B, Bt);
    matmul3(FinvT, Bt, t3);
    matmul3(t3, Finv, t4);
    for (int i = 0; i < 9; ++i) {
        const double id = (i % 4 == 0) ? 1.0 : 0.0;
        const double Q = FBFt[i] * 2.0 / beta2 + id * G * trB;
        const double R = t2[i] * (2.0 * J - alpha) * J - t4[i] * H;
        out[i] = Q * mu + R * lambda;
    }
    return true;
}

// The default ElasticMaterialModel::contract_multiple_stress_tensors_into
// (fenris/src/solid/mod.rs:56-86), which InvertibleMaterial inherits: one closed-form
// contraction per node pair (i <= j) at the projected F, the transpose mirrored.
bool contract_pairwise(double mu, double lambda, const double* F, int n, const double* a, double* out) {
    const int dim = 3 * n;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double c3[9];
            if (!snh_contract_with_r(mu, lambda, F, a + 3 * i, a + 3 * j, c3)) return false;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) {
                    out[(3 * i + r) + dim * (3 * j + c)] += c3[r + 3 * c];
                    if (i != j) out[(3 * j + c) + dim * (3 * i + r)] += c3[r + 3 * c];
                }
        }
    return true;
}

// The stress of the elastic model kind (ORC_MAT_*), reparametrized mu / lambda.
bool material_stress_r(int kind, double thr, double mu, double lambda, const double* F, double* P) {
    if (kind == ORC_MAT_INVERTIBLE_SNH) {
        double Fp[9];
        project_deformation_gradient(F, thr, Fp);
        return snh_stress_r(mu, lambda, Fp, P);
    }
    return snh_stress_r(mu, lambda, F, P);
}

// The tangent dP/dF of the elastic model kind (for the invertible kind: at the projected F).
bool material_dpdf_r(int kind, double thr, double mu, double lambda, const double* F, double* dp) {
    if (kind == ORC_MAT_INVERTIBLE_SNH) {
        double Fp[9];
        project_deformation_gradient(F, thr, Fp);
        return snh_dpdf_r(mu, lambda, Fp, dp);
    }
    if (kind == ORC_MAT_PROJECTED_SNH) {
        double Finv[9];
        if (!inv3(F, Finv)) return false;
        snh_dpdf_projected(mu, lambda, F, dp);
        return true;
    }
    return snh_dpdf_r(mu, lambda, F, dp);
}

double norm2(const std::vector<double>& x) {
    double s = 0.0;
    for (double v : x) s += v * v;
    return std::sqrt(s);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
int orc_nodes_per_element(int t) { return nodes_per(t); }
int orc_geometry_nodes(int t) { return geom_nodes(t); }

// fenris/src/solid/materials.rs:27-38 (includes the errata's factor-4 lambda bug,
// errata/tex/errata.tex:24-38)
void orc_lame_from_young_poisson(double young, double poisson, double* mu, double* lambda) {
    const double m = 0.5 * young / (1.0 + poisson);
    *mu = m;
    *lambda = 0.5 * m * poisson / (1.0 - 2.0 * poisson);
}

void orc_snh_reparametrize(double mu, double lambda, double* mu_out, double* lambda_out) {
    snh_reparam(mu, lambda, mu_out, lambda_out);
}

// fenris/src/solid/materials.rs:269-297 (note: uses the un-reparametrized mu/lambda in the
// outer terms exactly as the reference does)
int orc_snh_energy(double mu0, double lambda0, int64_t nq, const double* F, double* psi) {
    const double d = 3.0;
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    for (int64_t q = 0; q < nq; ++q) {
        const double* f = F + 9 * q;
        const double a = mu / lambda * (d / (d + 1.0));
        const double alpha = 1.0 + a;
        double Ft[9], C[9];
        transpose3(f, Ft);
        matmul3(Ft, f, C);
        const double I_C = C[0] + C[4] + C[8];
        const double J = det3(f);
        const double Jma = J - alpha;
        psi[q] = 0.5 * mu0 * (I_C - d) + 0.5 * lambda0 * ((Jma * Jma) - (a * a)) -
                 0.5 * mu0 * std::log((I_C + 1.0) / (d + 1.0));
    }
    return ORC_OK;
}

int orc_snh_stress(double mu0, double lambda0, int64_t nq, const double* F, double* P) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    for (int64_t q = 0; q < nq; ++q)
        if (!snh_stress_r(mu, lambda, F + 9 * q, P + 9 * q)) return fail(ORC_ERR_SINGULAR, "singular F");
    return ORC_OK;
}

int orc_snh_dpdf(double mu0, double lambda0, int64_t nq, const double* F, double* dpdf) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    for (int64_t q = 0; q < nq; ++q)
        if (!snh_dpdf_r(mu, lambda, F + 9 * q, dpdf + 81 * q)) return fail(ORC_ERR_SINGULAR, "singular F");
    return ORC_OK;
}

// fenris/src/solid/materials.rs:324-347
int orc_snh_contract_with(double mu0, double lambda0, const double* F, const double* a, const double* b,
                          double* out) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    if (!snh_contract_with_r(mu, lambda, F, a, b, out)) return fail(ORC_ERR_SINGULAR, "singular F");
    return ORC_OK;
}

int orc_snh_contract_multiple(double mu0, double lambda0, const double* F, int n, const double* a, double* out) {
    double mu, lambda, dp[81];
    snh_reparam(mu0, lambda0, &mu, &lambda);
    if (!snh_dpdf_r(mu, lambda, F, dp)) return fail(ORC_ERR_SINGULAR, "singular F");
    contract_stiffness(dp, n, a, out);
    return ORC_OK;
}

int orc_svd3(int64_t nq, const double* F, double* U, double* S, double* V) {
    for (int64_t q = 0; q < nq; ++q) svd3(F + 9 * q, U + 9 * q, S + 3 * q, V + 9 * q);
    return ORC_OK;
}

int orc_rotation_svd(int64_t nq, const double* F, double* U, double* S, double* Vt) {
    for (int64_t q = 0; q < nq; ++q) rotation_svd3(F + 9 * q, U + 9 * q, S + 3 * q, Vt + 9 * q);
    return ORC_OK;
}

int orc_project_deformation_gradient(double threshold, int64_t nq, const double* F, double* Fout) {
    for (int64_t q = 0; q < nq; ++q) project_deformation_gradient(F + 9 * q, threshold, Fout + 9 * q);
    return ORC_OK;
}

int orc_snh_dpdf_eigen(double mu0, double lambda0, const double* F, double* evals, double* evecs) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    snh_dpdf_eigen(mu, lambda, F, evals, evecs);
    return ORC_OK;
}

int orc_material_stress(int kind, double threshold, double mu0, double lambda0, int64_t nq, const double* F,
                        double* P) {
    if (kind < ORC_MAT_SNH || kind > ORC_MAT_PROJECTED_SNH) return fail(ORC_ERR_INVALID, "unknown material kind");
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    for (int64_t q = 0; q < nq; ++q)
        if (!material_stress_r(kind, threshold, mu, lambda, F + 9 * q, P + 9 * q))
            return fail(ORC_ERR_SINGULAR, "singular F");
    return ORC_OK;
}

int orc_material_dpdf(int kind, double threshold, double mu0, double lambda0, int64_t nq, const double* F,
                      double* dpdf) {
    if (kind < ORC_MAT_SNH || kind > ORC_MAT_PROJECTED_SNH) return fail(ORC_ERR_INVALID, "unknown material kind");
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    for (int64_t q = 0; q < nq; ++q)
        if (!material_dpdf_r(kind, threshold, mu, lambda, F + 9 * q, dpdf + 81 * q))
            return fail(ORC_ERR_SINGULAR, "singular F");
    return ORC_OK;
}

int orc_material_contract_multiple(int kind, double threshold, double mu0, double lambda0, const double* F, int n,
                                   const double* a, double* out) {
    if (kind < ORC_MAT_SNH || kind > ORC_MAT_PROJECTED_SNH) return fail(ORC_ERR_INVALID, "unknown material kind");
    double mu, lambda, dp[81];
    snh_reparam(mu0, lambda0, &mu, &lambda);
    if (kind == ORC_MAT_INVERTIBLE_SNH) {
        double Fp[9];
        project_deformation_gradient(F, threshold, Fp);
        if (!contract_pairwise(mu, lambda, Fp, n, a, out)) return fail(ORC_ERR_SINGULAR, "singular F");
        return ORC_OK;
    }
    if (!material_dpdf_r(kind, threshold, mu, lambda, F, dp)) return fail(ORC_ERR_SINGULAR, "singular F");
    contract_stiffness(dp, n, a, out);
    return ORC_OK;
}

int orc_element_basis(int t, int64_t npts, const double* xi, double* phi) {
    const int n = nodes_per(t);
    if (n < 0) return fail(ORC_ERR_INVALID, "bad element type");
    for (int64_t p = 0; p < npts; ++p) basis(t, xi + 3 * p, phi + n * p);
    return ORC_OK;
}

int orc_element_gradients(int t, int64_t npts, const double* xi, double* grad) {
    const int n = nodes_per(t);
    if (n < 0) return fail(ORC_ERR_INVALID, "bad element type");
    for (int64_t p = 0; p < npts; ++p) gradients(t, xi + 3 * p, grad + 3 * n * p);
    return ORC_OK;
}

int orc_element_reference_jacobian(int t, const double* nodes, const double* xi, double* J) {
    if (nodes_per(t) < 0) return fail(ORC_ERR_INVALID, "bad element type");
    reference_jacobian(t, nodes, xi, J);
    return ORC_OK;
}

}  // extern "C"

// ====================================================================== model
struct orc_model {
    int type = 0, n = 0, gn = 0;
    int64_t nv = 0, ne = 0, ninterior = 0;
    std::vector<double> vertices;
    std::vector<int32_t> conn;
    struct Quad {
        std::vector<double> iw, ixi;
        std::vector<int64_t> off;
        std::vector<double> fw, fxi;
    } mass, stiff, ell;
    double mass_reg = 0.0;
    int material = ORC_MAT_SNH;  // elastic model kind
    double threshold = 0.0;      // InvertibleMaterial threshold
    std::vector<std::vector<int32_t>> colors[2];  // per pass: color -> global element ids
    Csr pattern;                                  // values unused

    void quad_for(const Quad& q, int64_t e, const double** w, const double** xi, int64_t* nq) const {
        if (e < ninterior) {
            *w = q.iw.data();
            *xi = q.ixi.data();
            *nq = (int64_t)q.iw.size();
        } else {
            const int64_t k = e - ninterior;
            *w = q.fw.data() + q.off[k];
            *xi = q.fxi.data() + 3 * q.off[k];
            *nq = q.off[k + 1] - q.off[k];
        }
    }
    void geom(int64_t e, double* X) const {
        for (int k = 0; k < gn; ++k)
            for (int i = 0; i < 3; ++i) X[3 * k + i] = vertices[3 * (int64_t)conn[e * n + k] + i];
    }
};

namespace {
void copy_quad(const orc_quadrature& src, int64_t ninterface, orc_model::Quad& dst) {
    dst.iw.assign(src.interior_w, src.interior_w + src.interior_nq);
    dst.ixi.assign(src.interior_xi, src.interior_xi + 3 * src.interior_nq);
    if (ninterface > 0 && src.interface_offsets) {
        dst.off.assign(src.interface_offsets, src.interface_offsets + ninterface + 1);
        const int64_t tot = dst.off.back();
        dst.fw.assign(src.interface_w, src.interface_w + tot);
        dst.fxi.assign(src.interface_xi, src.interface_xi + 3 * tot);
    } else {
        dst.off.assign(ninterface + 1, 0);
    }
}

// assemble_generalized_element_elliptic_term (fenris/src/assembly.rs:674-711) with the
// elliptic operator of MaterialEllipticOperator (fenris/src/solid/assembly.rs:33-38).
bool element_force(const orc_model& m, int64_t e, double mu_r, double lambda_r, const double* u, double* fe) {
    const int n = m.n;
    double X[24];
    m.geom(e, X);
    std::vector<double> ue(3 * n), G(3 * n);
    for (int I = 0; I < n; ++I)
        for (int b = 0; b < 3; ++b) ue[3 * I + b] = u[3 * (int64_t)m.conn[e * n + I] + b];
    for (int i = 0; i < 3 * n; ++i) fe[i] = 0.0;
    const double *w, *xi;
    int64_t nq;
    m.quad_for(m.ell, e, &w, &xi, &nq);
    for (int64_t q = 0; q < nq; ++q) {
        gradients(m.type, xi + 3 * q, G.data());
        double J[9], Jinv[9];
        reference_jacobian(m.type, X, xi + 3 * q, J);
        const double detJ = det3(J);
        if (!inv3(J, Jinv)) return false;
        // u_grad = (u * (G^T * J_inv))^T
        double mm[9] = {0};  // (b, a)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                double s = 0.0;
                for (int I = 0; I < n; ++I) {
                    double t = 0.0;
                    for (int c = 0; c < 3; ++c) t += G[3 * I + c] * Jinv[c + 3 * a];
                    s += ue[3 * I + b] * t;
                }
                mm[b + 3 * a] = s;
            }
        double ugrad[9];
        transpose3(mm, ugrad);
        // F = u_grad^T + I; g = -P^T; g^T = -P
        double F[9], P[9];
        transpose3(ugrad, F);
        F[0] += 1.0;
        F[4] += 1.0;
        F[8] += 1.0;
        if (!material_stress_r(m.material, m.threshold, mu_r, lambda_r, F, P)) return false;
        // A = g^T * J_inv^T ; f_e += (A * G) * w * |detJ|
        double gT[9];
        for (int i = 0; i < 9; ++i) gT[i] = -P[i];
        double JinvT[9], A[9];
        transpose3(Jinv, JinvT);
        matmul3(gT, JinvT, A);
        for (int I = 0; I < n; ++I)
            for (int i = 0; i < 3; ++i) {
                double s = 0.0;
                for (int a = 0; a < 3; ++a) s += A[i + 3 * a] * G[3 * I + a];
                fe[3 * I + i] += s * w[q] * std::fabs(detJ);
            }
    }
    return true;
}

// assemble_generalized_element_stiffness (fenris/src/assembly.rs:747-797)
bool element_stiffness(const orc_model& m, int64_t e, double mu_r, double lambda_r, const double* u, double* Ke) {
    const int n = m.n, dim = 3 * n;
    double X[24];
    m.geom(e, X);
    std::vector<double> ue(3 * n), Gref(3 * n), G(3 * n);
    for (int I = 0; I < n; ++I)
        for (int b = 0; b < 3; ++b) ue[3 * I + b] = u[3 * (int64_t)m.conn[e * n + I] + b];
    for (int i = 0; i < dim * dim; ++i) Ke[i] = 0.0;
    const double *w, *xi;
    int64_t nq;
    m.quad_for(m.stiff, e, &w, &xi, &nq);
    for (int64_t q = 0; q < nq; ++q) {
        double J[9], Jinv[9], JinvT[9];
        reference_jacobian(m.type, X, xi + 3 * q, J);
        const double detJ = det3(J);
        if (!inv3(J, Jinv)) return false;
        transpose3(Jinv, JinvT);
        gradients(m.type, xi + 3 * q, Gref.data());
        for (int I = 0; I < n; ++I)
            for (int a = 0; a < 3; ++a) {
                double s = 0.0;
                for (int c = 0; c < 3; ++c) s += JinvT[a + 3 * c] * Gref[3 * I + c];
                G[3 * I + a] = s;
            }
        double ugrad[9];  // G * u^T
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int I = 0; I < n; ++I) s += G[3 * I + a] * ue[3 * I + b];
                ugrad[a + 3 * b] = s;
            }
        const double scale = w[q] * std::fabs(detJ);
        const double sq = std::sqrt(scale);
        for (int i = 0; i < 3 * n; ++i) G[i] *= sq;
        double F[9], dp[81];
        transpose3(ugrad, F);
        F[0] += 1.0;
        F[4] += 1.0;
        F[8] += 1.0;
        if (m.material == ORC_MAT_INVERTIBLE_SNH) {
            double Fp[9];
            project_deformation_gradient(F, m.threshold, Fp);
            if (!contract_pairwise(mu_r, lambda_r, Fp, n, G.data(), Ke)) return false;
            continue;
        }
        if (!material_dpdf_r(m.material, m.threshold, mu_r, lambda_r, F, dp)) return false;
        contract_stiffness(dp, n, G.data(), Ke);
    }
    return true;
}

// assemble_generalized_element_mass (fenris/src/assembly.rs:624-671), scalar part
void element_mass_scalar(const orc_model& m, int64_t e, double density, const orc_model::Quad& quad, bool interior_rule,
                         double* Ms) {
    const int n = m.n;
    double X[24];
    m.geom(e, X);
    std::vector<double> phi(n);
    for (int i = 0; i < n * n; ++i) Ms[i] = 0.0;
    const double *w, *xi;
    int64_t nq;
    if (interior_rule) {
        w = quad.iw.data();
        xi = quad.ixi.data();
        nq = (int64_t)quad.iw.size();
    } else {
        m.quad_for(quad, e, &w, &xi, &nq);
    }
    for (int64_t q = 0; q < nq; ++q) {
        double J[9];
        reference_jacobian(m.type, X, xi + 3 * q, J);
        const double detJ = det3(J);
        basis(m.type, xi + 3 * q, phi.data());
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Ms[i + n * j] += density * std::fabs(detJ) * w[q] * phi[i] * phi[j];
    }
}
}  // namespace

extern "C" {

int orc_model_create(const orc_model_desc* d, orc_model** out) {
    const int n = nodes_per(d->element_type);
    if (n < 0) return fail(ORC_ERR_INVALID, "bad element type");
    if (d->num_interior < 0 || d->num_interior > d->num_elements) return fail(ORC_ERR_INVALID, "bad num_interior");
    auto* m = new orc_model();
    m->type = d->element_type;
    m->n = n;
    m->gn = geom_nodes(d->element_type);
    m->nv = d->num_vertices;
    m->ne = d->num_elements;
    m->ninterior = d->num_interior;
    m->vertices.assign(d->vertices, d->vertices + 3 * d->num_vertices);
    m->conn.assign(d->connectivity, d->connectivity + (int64_t)n * d->num_elements);
    for (int32_t c : m->conn)
        if (c < 0 || c >= m->nv) {
            delete m;
            return fail(ORC_ERR_INVALID, "connectivity out of bounds");
        }
    const int64_t nif = m->ne - m->ninterior;
    copy_quad(d->mass_quadrature, nif, m->mass);
    copy_quad(d->stiffness_quadrature, nif, m->stiff);
    copy_quad(d->elliptic_quadrature, nif, m->ell);
    m->mass_reg = d->mass_regularization_factor;
    // colors: NodalModel (fenris/src/model.rs:90-102) and EmbeddedModel interior /
    // interface passes (fenris/src/embedding/mod.rs:262-270)
    const int64_t bounds[3] = {0, m->ninterior, m->ne};
    for (int p = 0; p < 2; ++p) {
        const int64_t cnt = bounds[p + 1] - bounds[p];
        std::vector<int64_t> off(cnt + 1);
        for (int64_t i = 0; i <= cnt; ++i) off[i] = i * n;
        auto cols = greedy_coloring(cnt, off.data(), m->conn.data() + bounds[p] * n);
        for (auto& c : cols)
            for (auto& l : c) l += (int32_t)bounds[p];
        m->colors[p] = cols;
    }
    std::vector<int64_t> off(m->ne + 1);
    for (int64_t i = 0; i <= m->ne; ++i) off[i] = i * n;
    std::vector<int32_t> none;
    build_pattern(m->nv, m->ne, off.data(), m->conn.data(), 3, none, m->pattern);
    if (m->pattern.col.size() >= (size_t)INT32_MAX) {
        delete m;
        return fail(ORC_ERR_INVALID, "nnz exceeds int32");
    }
    *out = m;
    return ORC_OK;
}

void orc_model_destroy(orc_model* m) { delete m; }

int orc_model_set_material(orc_model* m, int kind, double threshold) {
    if (kind != ORC_MAT_SNH && kind != ORC_MAT_INVERTIBLE_SNH && kind != ORC_MAT_PROJECTED_SNH)
        return fail(ORC_ERR_INVALID, "unknown material kind");
    m->material = kind;
    m->threshold = kind == ORC_MAT_INVERTIBLE_SNH ? threshold : 0.0;
    return ORC_OK;
}
int64_t orc_model_ndof(const orc_model* m) { return 3 * m->nv; }
int orc_model_num_colors(const orc_model* m, int pass) { return (int)m->colors[pass].size(); }

int orc_model_colors(const orc_model* m, int pass, int32_t* offsets, int32_t* labels) {
    int32_t k = 0;
    offsets[0] = 0;
    for (size_t c = 0; c < m->colors[pass].size(); ++c) {
        for (int32_t l : m->colors[pass][c]) labels[k++] = l;
        offsets[c + 1] = k;
    }
    return ORC_OK;
}

int64_t orc_model_nnz(const orc_model* m) { return (int64_t)m->pattern.col.size(); }

int orc_model_pattern(const orc_model* m, int32_t* rowptr, int32_t* col) {
    std::copy(m->pattern.rowptr.begin(), m->pattern.rowptr.end(), rowptr);
    std::copy(m->pattern.col.begin(), m->pattern.col.end(), col);
    return ORC_OK;
}

// ElasticityModel::assemble_mass -> COO -> to_csr(Add) (fenris/src/assembly.rs:1355-1380,
// fenris/src/embedding/mod.rs:404-451). Duplicates are summed in element order.
int orc_model_assemble_mass(const orc_model* m, double density, double* values) {
    Csr csr = m->pattern;
    const int n = m->n;
    std::vector<double> Ms(n * n);
    auto scatter = [&](int64_t e) -> bool {
        for (int il = 0; il < n; ++il)
            for (int jl = 0; jl < n; ++jl)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        const int64_t r = 3 * (int64_t)m->conn[e * n + il] + i;
                        const int64_t c = 3 * (int64_t)m->conn[e * n + jl] + j;
                        const int64_t k = csr.find(r, c);
                        if (k < 0) return false;
                        csr.val[k] += (i == j) ? Ms[il + n * jl] : 0.0;
                    }
        return true;
    };
    for (int64_t e = 0; e < m->ne; ++e) {
        element_mass_scalar(*m, e, density, m->mass, false, Ms.data());
        if (!scatter(e)) return fail(ORC_ERR_PATTERN, "missing CSR entry");
    }
    if (m->mass_reg > 0.0) {
        for (int64_t e = 0; e < m->ne; ++e) {
            element_mass_scalar(*m, e, m->mass_reg * density, m->mass, true, Ms.data());
            if (!scatter(e)) return fail(ORC_ERR_PATTERN, "missing CSR entry");
        }
    }
    std::copy(csr.val.begin(), csr.val.end(), values);
    return ORC_OK;
}

// assemble_generalized_elliptic_term_into_par (fenris/src/assembly.rs:915-966) through
// assemble_pseudo_forces_into_par: color by color, f += f_e. Interior then interface pass
// for embedded models (fenris/src/embedding/mod.rs:516-554).
int orc_model_forces_add(const orc_model* m, double mu0, double lambda0, const double* u, double* f) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    const int n = m->n;
    int status = ORC_OK;
    for (int p = 0; p < 2; ++p)
        for (const auto& color : m->colors[p]) {
            const int64_t cnt = (int64_t)color.size();
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < cnt; ++k) {
                const int64_t e = color[k];
                std::vector<double> fe(3 * n);
                if (!element_force(*m, e, mu, lambda, u, fe.data())) {
#pragma omp critical
                    status = ORC_ERR_SINGULAR;
                    continue;
                }
                for (int I = 0; I < n; ++I)
                    for (int i = 0; i < 3; ++i) f[3 * (int64_t)m->conn[e * n + I] + i] += fe[3 * I + i];
            }
            if (status != ORC_OK) return fail(status, "singular deformation gradient or Jacobian");
        }
    return ORC_OK;
}

// CsrParAssembler::assemble_into_csr (fenris/src/assembly.rs:253-313): per color, element
// matrices added row by row with add_element_row_to_csr_row. values += K.
}  // extern "C"

namespace {
// DefaultSemidefiniteProjection (fenris/src/assembly.rs:552-564) of a symmetric dim x dim matrix
// (column-major, in place): eigen-decomposition, eigenvalues clamped at zero, recomposed.
void psd_project(int dim, double* K) {
    std::vector<double> a(K, K + (size_t)dim * dim), ev, V;
    jacobi_eigen(dim, a, ev, V);
    for (int c = 0; c < dim; ++c)
        for (int r = 0; r < dim; ++r) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += V[r + (size_t)dim * k] * std::max(0.0, ev[k]) * V[c + (size_t)dim * k];
            K[r + (size_t)dim * c] = s;
        }
}

int stiffness_add_impl(const orc_model* m, double mu0, double lambda0, const double* u, double* values, bool project) {
    double mu, lambda;
    snh_reparam(mu0, lambda0, &mu, &lambda);
    const int n = m->n, dim = 3 * n;
    Csr csr;
    csr.n = m->pattern.n;
    csr.rowptr = m->pattern.rowptr;
    csr.col = m->pattern.col;
    csr.val.assign(values, values + csr.col.size());
    int status = ORC_OK;
    for (int p = 0; p < 2; ++p)
        for (const auto& color : m->colors[p]) {
            const int64_t cnt = (int64_t)color.size();
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t k = 0; k < cnt; ++k) {
                const int64_t e = color[k];
                std::vector<double> Ke((size_t)dim * dim);
                if (!element_stiffness(*m, e, mu, lambda, u, Ke.data())) {
#pragma omp critical
                    status = ORC_ERR_SINGULAR;
                    continue;
                }
                if (project) psd_project(dim, Ke.data());
                const int32_t* g = m->conn.data() + e * n;
                std::vector<int> perm(n);
                std::iota(perm.begin(), perm.end(), 0);
                std::sort(perm.begin(), perm.end(), [&](int a, int b) { return g[a] < g[b]; });
                for (int ln = 0; ln < n; ++ln)
                    for (int i = 0; i < 3; ++i)
                        if (!add_element_row(csr, 3 * (int64_t)g[ln] + i, g, perm, 3, Ke.data(), dim, 3 * ln + i)) {
#pragma omp critical
                            status = ORC_ERR_PATTERN;
                        }
            }
            if (status != ORC_OK) return fail(status, "stiffness assembly failed");
        }
    std::copy(csr.val.begin(), csr.val.end(), values);
    return ORC_OK;
}
}  // namespace

extern "C" {

int orc_model_stiffness_add(const orc_model* m, double mu0, double lambda0, const double* u, double* values) {
    return stiffness_add_impl(m, mu0, lambda0, u, values, false);
}

// assemble_transformed_stiffness_into_par with DefaultSemidefiniteProjection (fenris/src/solid/
// impl_model.rs:146-164, fenris/src/assembly.rs:552-564): values += sum of projected K_e
int orc_model_stiffness_add_projected(const orc_model* m, double mu0, double lambda0, const double* u,
                                      double* values) {
    return stiffness_add_impl(m, mu0, lambda0, u, values, true);
}

// cuthill_mckee (fenris/src/reorder.rs:169-235) on a square CSR pattern: per connected component,
// start from the first unvisited vertex of minimum degree (lane length), breadth-first, the
// adjacent vertices of each popped vertex visited by increasing degree. The reference sorts them
// with sort_unstable_by_key; this restatement breaks ties by CSR position (increasing index).
int orc_cuthill_mckee(int64_t n, const int32_t* rowptr, const int32_t* col, int64_t* perm) {
    std::vector<char> visited(n, 0);
    std::vector<int64_t> out;
    out.reserve(n);
    auto degree = [&](int64_t v) { return (int64_t)(rowptr[v + 1] - rowptr[v]); };
    while ((int64_t)out.size() < n) {
        int64_t start = -1;
        for (int64_t v = 0; v < n; ++v)
            if (!visited[v] && (start < 0 || degree(v) < degree(start))) start = v;
        std::vector<int64_t> queue{start};
        visited[start] = 1;
        for (size_t h = 0; h < queue.size(); ++h) {
            const int64_t v = queue[h];
            out.push_back(v);
            std::vector<int64_t> adj(col + rowptr[v], col + rowptr[v + 1]);
            std::stable_sort(adj.begin(), adj.end(), [&](int64_t a, int64_t b) { return degree(a) < degree(b); });
            for (int64_t a : adj)
                if (!visited[a]) {
                    visited[a] = 1;
                    queue.push_back(a);
                }
        }
    }
    std::copy(out.begin(), out.end(), perm);
    return ORC_OK;
}

// reorder_mesh_par (fenris/src/reorder.rs:52-88): reverse Cuthill-McKee of the vertex pattern
// (CsrParAssembler::assemble_pattern), elements stably sorted by their smallest new vertex index.
int orc_mesh_reorder(int n, int64_t nv, int64_t ne, const int32_t* conn, int64_t* vperm, int64_t* eperm) {
    std::vector<int64_t> off(ne + 1);
    for (int64_t e = 0; e <= ne; ++e) off[e] = e * n;
    Csr pat;
    std::vector<int32_t> no_diag;
    build_pattern(nv, ne, off.data(), conn, 1, no_diag, pat);
    orc_cuthill_mckee(nv, pat.rowptr.data(), pat.col.data(), vperm);
    std::reverse(vperm, vperm + nv);
    std::vector<int64_t> inv(nv);
    for (int64_t i = 0; i < nv; ++i) inv[vperm[i]] = i;
    std::vector<int64_t> idx(ne);
    std::iota(idx.begin(), idx.end(), 0);
    auto key = [&](int64_t e) {
        int64_t k = inv[conn[e * n]];
        for (int j = 1; j < n; ++j) k = std::min(k, inv[conn[e * n + j]]);
        return k;
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return key(a) < key(b); });
    std::copy(idx.begin(), idx.end(), eperm);
    return ORC_OK;
}

int orc_psd_project(int dim, const double* a, double* out) {
    std::copy(a, a + (size_t)dim * dim, out);
    psd_project(dim, out);
    return ORC_OK;
}

int orc_greedy_coloring(int64_t nsets, const int64_t* offsets, const int32_t* nodes, int32_t* color_of_set,
                        int32_t* num_colors) {
    auto cols = greedy_coloring(nsets, offsets, nodes);
    for (size_t c = 0; c < cols.size(); ++c)
        for (int32_t l : cols[c]) color_of_set[l] = (int32_t)c;
    *num_colors = (int32_t)cols.size();
    return ORC_OK;
}

int orc_coo_pattern(int64_t n, int64_t nsets, const int64_t* offsets, const int32_t* nodes, int sdim,
                    int64_t* nnz_out, int32_t* rowptr, int32_t* col) {
    Csr csr;
    std::vector<int32_t> none;
    build_pattern(n, nsets, offsets, nodes, sdim, none, csr);
    *nnz_out = (int64_t)csr.col.size();
    if (rowptr) std::copy(csr.rowptr.begin(), csr.rowptr.end(), rowptr);
    if (col) std::copy(csr.col.begin(), csr.col.end(), col);
    return ORC_OK;
}

// spmv_csr (fenris/src/sparse.rs:1168-1199): y <- beta*y + alpha*A*x
void orc_spmv(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals, double beta, double* y,
              double alpha, const double* x) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double yi = beta * y[i];
        double s = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += vals[k] * x[col[k]];
        yi += alpha * s;
        y[i] = yi;
    }
}

// apply_homogeneous_dirichlet_bc_csr (fenris/src/assembly.rs:799-869)
void orc_dirichlet_csr(int64_t n, const int32_t* rowptr, const int32_t* col, double* vals, int64_t nnodes,
                       const int32_t* nodes, int sdim) {
    const int d = sdim;
    double scale = 1.0;
    for (int64_t i = 0; i < n; ++i) {
        double di = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
            if (col[k] == i) {
                di = vals[k];
                break;
            }
        if (di != 0.0) {
            scale = std::fabs(di);
            break;
        }
    }
    std::vector<char> member(n, 0), visit(n, 0);
    for (int64_t a = 0; a < nnodes; ++a)
        for (int i = 0; i < d; ++i) {
            const int64_t r = d * (int64_t)nodes[a] + i;
            member[r] = 1;
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                if (col[k] == r)
                    vals[k] = scale;
                else {
                    vals[k] = 0.0;
                    visit[col[k]] = 1;
                }
            }
        }
    for (int64_t r = 0; r < n; ++r)
        if (visit[r] && !member[r])
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
                if (member[col[k]]) vals[k] = 0.0;
}

// apply_homogeneous_dirichlet_bc_rhs (fenris/src/assembly.rs:897-913)
void orc_dirichlet_rhs(double* rhs, int64_t nnodes, const int32_t* nodes) {
    for (int64_t a = 0; a < nnodes; ++a)
        for (int i = 0; i < 3; ++i) rhs[3 * (int64_t)nodes[a] + i] = 0.0;
}

int orc_symmetric_eigen(int m, const double* a, double* evals, double* evecs) {
    std::vector<double> A(a, a + (size_t)m * m), ev, v;
    if (m > ORC_JACOBI_MAX_DIM)
        tridiagonal_qr_eigen(m, A, ev, v);
    else
        jacobi_eigen(m, A, ev, v);
    std::copy(ev.begin(), ev.end(), evals);
    std::copy(v.begin(), v.end(), evecs);
    return ORC_OK;
}

int orc_invert_block(int m, const double* block, double* out) {
    invert_block(m, block, out);
    return ORC_OK;
}

}  // extern "C"

// ====================================================================== Schwarz
struct orc_schwarz {
    int64_t num_nodes = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> nodes, untreated;
    std::vector<std::vector<int32_t>> colors;
    Csr P;
};

extern "C" {

// build_preconditioner_pattern + sequential_greedy_coloring
// (simulation_toolbox/src/fem/schwarz_precond.rs:134-169, :45-63)
int orc_schwarz_create(int64_t num_nodes, int64_t nsub, const int64_t* offsets, const int32_t* nodes,
                       int64_t nuntreated, const int32_t* untreated, orc_schwarz** out) {
    auto* s = new orc_schwarz();
    s->num_nodes = num_nodes;
    s->off.assign(offsets, offsets + nsub + 1);
    s->nodes.assign(nodes, nodes + offsets[nsub]);
    s->untreated.assign(untreated, untreated + nuntreated);
    s->colors = greedy_coloring(nsub, s->off.data(), s->nodes.data());
    build_pattern(num_nodes, nsub, s->off.data(), s->nodes.data(), 3, s->untreated, s->P);
    *out = s;
    return ORC_OK;
}

void orc_schwarz_destroy(orc_schwarz* s) { delete s; }
int64_t orc_schwarz_nnz(const orc_schwarz* s) { return (int64_t)s->P.col.size(); }

int orc_schwarz_pattern(const orc_schwarz* s, int32_t* rowptr, int32_t* col) {
    std::copy(s->P.rowptr.begin(), s->P.rowptr.end(), rowptr);
    std::copy(s->P.col.begin(), s->P.col.end(), col);
    return ORC_OK;
}

// build_schwarz_preconditioner_into (simulation_toolbox/src/fem/schwarz_precond.rs:171-216)
// with SchwarzElementAssembler (:250-327): gather B_s from J, invert, add per color;
// untreated nodes get 1/J_ii.
int orc_schwarz_update(orc_schwarz* s, const int32_t* rowptr, const int32_t* col, const double* J) {
    Csr Jm;
    Jm.n = 3 * s->num_nodes;
    Jm.rowptr.assign(rowptr, rowptr + Jm.n + 1);
    Jm.col.assign(col, col + rowptr[Jm.n]);
    std::fill(s->P.val.begin(), s->P.val.end(), 0.0);
    int status = ORC_OK;
    // the block inverses are independent of each other: all of them first, in parallel; then the
    // scatter into P in color order (the additions into P in the reference's order)
    const int64_t nsub = (int64_t)s->off.size() - 1;
    std::vector<std::vector<double>> inv((size_t)nsub);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t sub = 0; sub < nsub; ++sub) {
        const int32_t* g = s->nodes.data() + s->off[sub];
        const int nn = (int)(s->off[sub + 1] - s->off[sub]);
        const int dim = 3 * nn;
        // B = J restricted to the subdomain's DOFs. The reference's gather (schwarz_precond.rs:
        // 304-320) requires every node pair of a subdomain to be coupled in J and panics otherwise;
        // an uncoupled pair is a zero of J, which is what B receives here.
        std::vector<double> B((size_t)dim * dim, 0.0);
        for (int il = 0; il < nn; ++il)
            for (int i = 0; i < 3; ++i) {
                const int64_t grow = 3 * (int64_t)g[il] + i;
                for (int jl = 0; jl < nn; ++jl)
                    for (int j = 0; j < 3; ++j) {
                        const int64_t kk = Jm.find(grow, 3 * (int64_t)g[jl] + j);
                        if (kk >= 0) B[(3 * il + i) + (size_t)dim * (3 * jl + j)] = J[kk];
                    }
            }
        inv[sub].resize((size_t)dim * dim);
        invert_block(dim, B.data(), inv[sub].data());
    }
    for (const auto& color : s->colors) {
        const int64_t cnt = (int64_t)color.size();
#pragma omp parallel for schedule(dynamic, 2)
        for (int64_t k = 0; k < cnt; ++k) {
            const int64_t sub = color[k];
            const int32_t* g = s->nodes.data() + s->off[sub];
            const int nn = (int)(s->off[sub + 1] - s->off[sub]);
            const int dim = 3 * nn;
            const std::vector<double>& Binv = inv[sub];
            std::vector<int> perm(nn);
            std::iota(perm.begin(), perm.end(), 0);
            std::sort(perm.begin(), perm.end(), [&](int a, int b) { return g[a] < g[b]; });
            for (int ln = 0; ln < nn; ++ln)
                for (int i = 0; i < 3; ++i)
                    if (!add_element_row(s->P, 3 * (int64_t)g[ln] + i, g, perm, 3, Binv.data(), dim, 3 * ln + i)) {
#pragma omp critical
                        status = ORC_ERR_PATTERN;
                    }
        }
        if (status != ORC_OK) return fail(status, "Schwarz block gather/scatter failed");
    }
    for (int32_t node : s->untreated)
        for (int i = 0; i < 3; ++i) {
            const int64_t r = 3 * (int64_t)node + i;
            if (s->P.rowptr[r + 1] - s->P.rowptr[r] != 1) return fail(ORC_ERR_PATTERN, "untreated row not diagonal");
            const int64_t kk = Jm.find(r, r);
            const double entry = kk < 0 ? 0.0 : J[kk];
            if (entry == 0.0) return fail(ORC_ERR_SINGULAR, "Schwarz: diagonal entry is zero");
            s->P.val[s->P.rowptr[r]] = 1.0 / entry;
        }
    return ORC_OK;
}

int orc_schwarz_values(const orc_schwarz* s, double* values) {
    std::copy(s->P.val.begin(), s->P.val.end(), values);
    return ORC_OK;
}

}  // extern "C"

// ====================================================================== CG
namespace {
struct Op {
    virtual void apply(const double* x, double* y) const = 0;
    virtual ~Op() {}
};
struct CsrOp : Op {
    int64_t n;
    const int32_t *rp, *c;
    const double* v;
    CsrOp(int64_t n_, const int32_t* rp_, const int32_t* c_, const double* v_) : n(n_), rp(rp_), c(c_), v(v_) {}
    void apply(const double* x, double* y) const override { orc_spmv(n, rp, c, v, 0.0, y, 1.0, x); }
};
struct DiagOp : Op {
    std::vector<double> d;
    void apply(const double* x, double* y) const override {
        for (size_t i = 0; i < d.size(); ++i) y[i] = d[i] * x[i];
    }
};
struct IdOp : Op {
    int64_t n;
    void apply(const double* x, double* y) const override { std::copy(x, x + n, y); }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// ConjugateGradient::solve_with_guess_ (fenris/src/cg.rs:377-482)
int pcg(int64_t n, const Op& A, const Op& P, const double* b_, double* x, double tol, int64_t max_iter,
        int64_t* iters) {
    std::vector<double> b(b_, b_ + n), r(n), z(n), p(n), Ap(n), xv(x, x + n);
    *iters = 0;
    A.apply(xv.data(), r.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    P.apply(r.data(), z.data());
    p = z;
    double zTr = dot(z, r);
    const double b_norm = norm2(b);
    if (b_norm == 0.0) {
        std::fill(x, x + n, 0.0);
        return ORC_OK;
    }
    int status = ORC_OK;
    while (true) {
        if (norm2(r) <= tol * b_norm) break;
        if (max_iter >= 0 && *iters >= max_iter) {
            status = ORC_ERR_CG_MAX_ITER;
            break;
        }
        A.apply(p.data(), Ap.data());
        const double pAp = dot(p, Ap);
        if (pAp <= 0.0) {
            status = ORC_ERR_CG_INDEFINITE_OPERATOR;
            break;
        }
        if (zTr <= 0.0) {
            status = ORC_ERR_CG_INDEFINITE_PRECONDITIONER;
            break;
        }
        const double alpha = zTr / pAp;
        for (int64_t i = 0; i < n; ++i) xv[i] = xv[i] + alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] - alpha * Ap[i];
        *iters += 1;
        P.apply(r.data(), z.data());
        const double zTr_next = dot(z, r);
        const double beta = zTr_next / zTr;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        zTr = zTr_next;
    }
    std::copy(xv.begin(), xv.end(), x);
    if (status != ORC_OK) g_err = "CG stopped early";
    return status;
}

// BacktrackingLineSearch::step (hamilton2/src/newton.rs:171-246)
template <class Eval>
int line_search(int64_t n, Eval& eval, std::vector<double>& f, std::vector<double>& x, const std::vector<double>& p,
                int64_t* evals) {
    const double c = 1e-4, alpha_min = 1e-6;
    double fn = 0.0;
    for (double v : f) fn += v * v;
    const double g_initial = 0.5 * fn;
    double alpha_prev = 0.0, alpha = 1.0;
    const double initial[3] = {1.0, 0.75, 0.5};
    int idx = 0;
    double tail = 0.25;
    while (true) {
        const double delta = alpha - alpha_prev;
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + delta * p[i];
        eval(x.data(), f.data());
        ++*evals;
        double s = 0.0;
        for (double v : f) s += v * v;
        const double g = 0.5 * s;
        if (g <= (1.0 - c * alpha) * g_initial) break;
        if (alpha < alpha_min) return fail(ORC_ERR_LINE_SEARCH, "line search: alpha below minimum");
        alpha_prev = alpha;
        ++idx;
        if (idx < 3)
            alpha = initial[idx];
        else {
            alpha = tail;
            tail *= 0.25;
        }
    }
    return ORC_OK;
}

// newton_line_search (hamilton2/src/newton.rs:76-131)
template <class Eval, class Solve>
int newton(int64_t n, std::vector<double>& x, Eval& eval, Solve& solve, double tol, int64_t max_iter, bool ls,
           int64_t* iters, int64_t* evals, double* final_norm) {
    std::vector<double> f(n), mdx(n);
    eval(x.data(), f.data());
    ++*evals;
    int64_t it = 0;
    while (true) {
        const double fnorm = norm2(f);
        *final_norm = fnorm;
        if (!(fnorm > tol)) break;
        if (max_iter >= 0 && it == max_iter) {
            *iters = it;
            return fail(ORC_ERR_NEWTON_MAX_ITER, "Newton: maximum iterations reached");
        }
        int st = solve(x.data(), f.data(), mdx.data());
        if (st != ORC_OK) {
            *iters = it;
            return ORC_ERR_JACOBIAN;
        }
        for (auto& v : mdx) v *= -1.0;
        if (ls) {
            st = line_search(n, eval, f, x, mdx, evals);
            if (st != ORC_OK) {
                *iters = it;
                return st;
            }
        } else {
            for (int64_t i = 0; i < n; ++i) x[i] = x[i] + mdx[i];
            eval(x.data(), f.data());
            ++*evals;
        }
        ++it;
    }
    *iters = it;
    return ORC_OK;
}

// backward_euler_step (hamilton2/src/integrators/euler.rs:155-199) for any dynamic system with
//   eval_f(t, u, v, f), apply_mass(y, x) (y += M x) and
//   solve_jac(t, u, v, alpha, beta, rhs, sol) (solves (M - alpha df/du - beta df/dv) sol = rhs),
// through BackwardEulerImplicitFunction (euler.rs:76-141) and newton_line_search.
template <class EvalF, class ApplyM, class SolveJ>
int be_step_generic(int64_t n, double t0, double dt, double* u, double* v, EvalF& eval_f, ApplyM& apply_mass,
                    SolveJ& solve_jac, double settings_tolerance, int64_t max_newton, int64_t* newton_iters,
                    int64_t* evals, double* fnorm_out, double* tol_out) {
    const double t_next = t0 + dt;
    std::vector<double> v_n(v, v + n), u_p(n);
    for (int64_t i = 0; i < n; ++i) u_p[i] = u[i] + v[i] * dt;
    std::vector<double> uu(n), vv(n);
    auto eval = [&](const double* w, double* F) {
        for (int64_t i = 0; i < n; ++i) {
            uu[i] = u_p[i] + w[i] * dt;
            vv[i] = v_n[i] + w[i];
        }
        if (eval_f(t_next, uu.data(), vv.data(), F) != ORC_OK) {
            for (int64_t i = 0; i < n; ++i) F[i] = NAN;
            return;
        }
        const double sc = dt * -1.0;
        for (int64_t i = 0; i < n; ++i) F[i] *= sc;
        apply_mass(F, w);
    };
    auto solve = [&](const double* w, const double* rhs, double* sol) -> int {
        for (int64_t i = 0; i < n; ++i) {
            uu[i] = u_p[i] + w[i] * dt;
            vv[i] = v_n[i] + w[i];
        }
        return solve_jac(t_next, uu.data(), vv.data(), -dt * dt, -dt, rhs, sol);
    };
    std::vector<double> w(n, 0.0);
    const double tol = dt * settings_tolerance;
    *tol_out = tol;
    int st = newton(n, w, eval, solve, tol, max_newton, true, newton_iters, evals, fnorm_out);
    if (st != ORC_OK) return st;
    for (int64_t i = 0; i < n; ++i) v[i] += w[i];
    for (int64_t i = 0; i < n; ++i) u[i] += v[i] * dt;
    return ORC_OK;
}

// integrate_model_with_substeps (simulation_toolbox/src/fem/newton_cg.rs:146-228): the step is
// retried from the saved initial state with 5 times as many substeps of dt / substeps after any
// failure, and abandoned once the count would reach 1000. step(t, dt, u, v) advances the state by
// one (sub)step and returns a status. On success u and v hold the result and *substeps the count.
template <class Step>
int substep_loop(int64_t n, double t, double dt, double* u, double* v, Step& step, int* substeps, int* last_error) {
    const int max_substeps = 1000, factor = 5;
    int num = 1;
    std::vector<double> us(u, u + n), vs(v, v + n);
    while (true) {
        const double sdt = dt / (double)num;
        int failed = ORC_OK;
        for (int k = 0; k < num && failed == ORC_OK; ++k) failed = step(t + k * sdt, sdt, us.data(), vs.data());
        if (failed == ORC_OK) break;
        *last_error = failed;
        num *= factor;
        if (num >= max_substeps) return fail(ORC_ERR_SUBSTEPPING, "Substepping failed; last error " + std::to_string(failed));
        std::copy(u, u + n, us.begin());
        std::copy(v, v + n, vs.begin());
    }
    *substeps = num;
    std::copy(us.begin(), us.end(), u);
    std::copy(vs.begin(), vs.end(), v);
    return ORC_OK;
}

}  // namespace

extern "C" {

int orc_integrate_with_substeps(int64_t n, double* u, double* v, double t0, double dt, void* ctx,
                                orc_dyn_eval_fn eval_f, orc_dyn_mass_fn apply_mass, orc_dyn_solve_fn solve_jac,
                                double tolerance, int64_t max_newton_iter, int64_t* substeps,
                                int64_t* newton_iterations) {
    auto ef = [&](double t, const double* uu, const double* vv, double* f) { return eval_f(ctx, t, uu, vv, f); };
    auto mf = [&](double* y, const double* x) { apply_mass(ctx, y, x); };
    auto sf = [&](double t, const double* uu, const double* vv, double a, double b, const double* rhs, double* sol) {
        return solve_jac(ctx, t, uu, vv, a, b, rhs, sol);
    };
    int64_t total = 0;
    auto step = [&](double t, double sdt, double* us, double* vs) -> int {
        int64_t it = 0, evals = 0;
        double fn = 0.0, tol = 0.0;
        const int st = be_step_generic(n, t, sdt, us, vs, ef, mf, sf, tolerance, max_newton_iter, &it, &evals, &fn, &tol);
        total += it;
        return st;
    };
    int num = 0, last = ORC_OK;
    const int st = substep_loop(n, t0, dt, u, v, step, &num, &last);
    if (substeps) *substeps = num;
    if (newton_iterations) *newton_iterations = total;
    return st;
}

int orc_pcg_csr(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals, const int32_t* prowptr,
                const int32_t* pcol, const double* pvals, const double* b, double* x, double tol, int64_t max_iter,
                int64_t* iterations) {
    CsrOp A(n, rowptr, col, vals);
    if (prowptr) {
        CsrOp P(n, prowptr, pcol, pvals);
        return pcg(n, A, P, b, x, tol, max_iter, iterations);
    }
    IdOp I;
    I.n = n;
    return pcg(n, A, I, b, x, tol, max_iter, iterations);
}

int orc_backward_euler_step(int64_t n, double* u, double* v, double t0, double dt, void* ctx, orc_dyn_eval_fn eval_f,
                            orc_dyn_mass_fn apply_mass, orc_dyn_solve_fn solve_jac, double tolerance,
                            int64_t max_newton_iter, int64_t* iterations) {
    auto ef = [&](double t, const double* uu, const double* vv, double* f) { return eval_f(ctx, t, uu, vv, f); };
    auto mf = [&](double* y, const double* x) { apply_mass(ctx, y, x); };
    auto sf = [&](double t, const double* uu, const double* vv, double a, double b, const double* rhs, double* sol) {
        return solve_jac(ctx, t, uu, vv, a, b, rhs, sol);
    };
    int64_t evals = 0;
    double fn = 0.0, tol = 0.0;
    return be_step_generic(n, t0, dt, u, v, ef, mf, sf, tolerance, max_newton_iter, iterations, &evals, &fn, &tol);
}

int orc_newton(int64_t n, double* x, void* ctx, orc_eval_fn eval, orc_solve_fn solve, double tolerance,
               int64_t max_iter, int ls, int64_t* iterations) {
    std::vector<double> xv(x, x + n);
    auto ev = [&](const double* xx, double* f) { eval(ctx, xx, f); };
    auto so = [&](const double* xx, const double* rhs, double* sol) { return solve(ctx, xx, rhs, sol); };
    int64_t evals = 0;
    double fn = 0.0;
    const int st = newton(n, xv, ev, so, tolerance, max_iter, ls != 0, iterations, &evals, &fn);
    std::copy(xv.begin(), xv.end(), x);
    return st;
}

}  // extern "C"

// ====================================================================== FE dynamic system
struct orc_sim {
    const orc_model* model = nullptr;
    orc_sim_desc d{};
    std::vector<int32_t> dir;
    Csr M, K, J;
    std::vector<double> gforce;
    Csr D;                 // damping matrix of the current (sub)step
    bool damping = false;
    double rep_force = 0.0;
    orc_schwarz* schwarz = nullptr;
    std::vector<double> diag_inv;
    // state for set_state
    std::vector<double> su, sv;
    orc_step_stats* stats = nullptr;
    // prescribed Dirichlet values, three per Dirichlet node (zeros: the Homogeneous condition)
    std::vector<double> bc_u, bc_v;
};

namespace {
// apply_dirichlet_bcs_unknowns (simulation_toolbox/src/fem/system_assembly.rs:78-104): the
// condition's displacement and velocity values on the Dirichlet unknowns
void apply_bc_unknowns(const orc_sim* s, double* u, double* v) {
    for (size_t k = 0; k < s->dir.size(); ++k) {
        const int64_t node = s->dir[k];
        for (int i = 0; i < 3; ++i) {
            u[3 * node + i] = s->bc_u[3 * k + i];
            v[3 * node + i] = s->bc_v[3 * k + i];
        }
    }
}

// NewtonCgDynamicSystem::eval_f (simulation_toolbox/src/fem/newton_cg.rs:441-478)
int eval_f(orc_sim* s, const double* u_in, const double* v_in, double* f) {
    const int64_t n = s->M.n;
    std::vector<double> u(u_in, u_in + n), v(v_in, v_in + n);
    for (int64_t i = 0; i < n; ++i) f[i] = 1.0 * s->gforce[i];
    apply_bc_unknowns(s, u.data(), v.data());
    const double t0 = now_s();
    int st = orc_model_forces_add(s->model, s->d.mu, s->d.lambda, u.data(), f);
    if (s->stats) s->stats->t_forces += now_s() - t0;
    if (st != ORC_OK) return st;
    // f += -D v
    if (s->damping) orc_spmv(n, s->D.rowptr.data(), s->D.col.data(), s->D.val.data(), 1.0, f, -1.0, v.data());
    orc_dirichlet_rhs(f, (int64_t)s->dir.size(), s->dir.data());
    return ORC_OK;
}

// compute_damping_matrix_into (simulation_toolbox/src/fem/system_assembly.rs:145-194), called by
// integrate_model (newton_cg.rs:256-265) with the (sub)step's initial displacement:
// D = 0 + a_K K+(u) + a_M M, K+ the stiffness without boundary conditions, element matrices
// projected onto the positive semi-definite cone unless the material is positive semi-definite.
int compute_damping(orc_sim* s, const double* u) {
    const int64_t nnz = (int64_t)s->M.col.size();
    const double aK = s->d.stiffness_damping_coefficient, aM = s->d.mass_damping_coefficient;
    if (aK > 0.0) {
        std::fill(s->K.val.begin(), s->K.val.end(), 0.0);
        const int st = s->model->material == ORC_MAT_PROJECTED_SNH
                           ? orc_model_stiffness_add(s->model, s->d.mu, s->d.lambda, u, s->K.val.data())
                           : orc_model_stiffness_add_projected(s->model, s->d.mu, s->d.lambda, u, s->K.val.data());
        if (st != ORC_OK) return st;
    }
    for (int64_t k = 0; k < nnz; ++k) {
        double v = 0.0;
        if (aK > 0.0) v += aK * s->K.val[k];
        if (aM > 0.0) v += aM * s->M.val[k];
        s->D.val[k] = v;
    }
    return ORC_OK;
}

// NewtonCgDynamicSystem::solve_jacobian_combination (newton_cg.rs:680-802) with
// (alpha, beta) = (-dt^2, -dt) from BackwardEulerImplicitFunction (hamilton2/src/integrators/euler.rs:122-140)
int solve_jacobian(orc_sim* s, double alpha, double beta, const double* rhs, double* sol) {
    const int64_t n = s->M.n;
    const int64_t nnz = (int64_t)s->M.col.size();
    double t0 = now_s();
    std::fill(s->K.val.begin(), s->K.val.end(), 0.0);
    int st = orc_model_stiffness_add(s->model, s->d.mu, s->d.lambda, s->su.data(), s->K.val.data());
    if (s->stats) s->stats->t_stiffness += now_s() - t0;
    if (st != ORC_OK) return st;
    t0 = now_s();
    // compute_jacobian_combination_into (system_assembly.rs:197-235): J = -alpha K - beta D + M
    const double kc = -alpha, dc = -beta;
    for (int64_t k = 0; k < nnz; ++k) {
        double v = 0.0;
        v += kc * s->K.val[k];
        if (s->damping) v += dc * s->D.val[k];
        v += 1.0 * s->M.val[k];
        s->J.val[k] = v;
    }
    orc_dirichlet_csr(n, s->J.rowptr.data(), s->J.col.data(), s->J.val.data(), (int64_t)s->dir.size(), s->dir.data(), 3);
    if (s->stats) s->stats->t_jacobian += now_s() - t0;
    t0 = now_s();
    std::unique_ptr<Op> P;
    if (s->schwarz) {
        st = orc_schwarz_update(s->schwarz, s->J.rowptr.data(), s->J.col.data(), s->J.val.data());
        if (st != ORC_OK) return st;
        P.reset(new CsrOp(n, s->schwarz->P.rowptr.data(), s->schwarz->P.col.data(), s->schwarz->P.val.data()));
    } else {
        // DiagonalPreconditioner::update (newton_cg.rs:558-575)
        auto* D = new DiagOp();
        D->d.resize(n);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t kk = s->J.find(i, i);
            const double a = kk < 0 ? 0.0 : s->J.val[kk];
            if (a == 0.0) {
                delete D;
                return fail(ORC_ERR_SINGULAR, "Diagonal matrix is not invertible.");
            }
            D->d[i] = 1.0 / a;
        }
        P.reset(D);
    }
    if (s->stats) s->stats->t_precond += now_s() - t0;
    t0 = now_s();
    CsrOp A(n, s->J.rowptr.data(), s->J.col.data(), s->J.val.data());
    std::fill(sol, sol + n, 0.0);
    int64_t it = 0;
    st = pcg(n, A, *P, rhs, sol, s->d.cg_tolerance, s->d.cg_max_iter, &it);
    if (s->stats) {
        s->stats->t_cg += now_s() - t0;
        s->stats->cg_iterations += it;
    }
    if (st == ORC_ERR_CG_MAX_ITER || st == ORC_ERR_CG_INDEFINITE_OPERATOR) st = ORC_OK;
    return st;
}

// the FE system's Backward Euler step (NewtonCgDynamicSystem, newton_cg.rs:413-820)
int be_step(orc_sim* s, double t0, double dt, double* u, double* v, int64_t* newton_iters) {
    const int64_t n = s->M.n;
    auto eval_fn = [&](double t, const double* uu, const double* vv, double* F) { return eval_f(s, uu, vv, F); };
    auto mass_fn = [&](double* y, const double* x) {
        orc_spmv(n, s->M.rowptr.data(), s->M.col.data(), s->M.val.data(), 1.0, y, 1.0, x);
    };
    auto solve_fn = [&](double t, const double* uu, const double* vv, double alpha, double beta, const double* rhs,
                        double* sol) -> int {
        // set_state (newton_cg.rs:662-672)
        s->su.assign(uu, uu + n);
        s->sv.assign(vv, vv + n);
        apply_bc_unknowns(s, s->su.data(), s->sv.data());
        return solve_jacobian(s, alpha, beta, rhs, sol);
    };
    int64_t evals = 0;
    double fnorm = 0.0, tol = 0.0;
    const int st = be_step_generic(n, t0, dt, u, v, eval_fn, mass_fn, solve_fn, s->d.newton_tolerance_factor * s->rep_force,
                                   s->d.max_newton_iter, newton_iters, &evals, &fnorm, &tol);
    if (s->stats) {
        s->stats->line_search_evals += evals;
        s->stats->final_residual = fnorm;
        s->stats->tolerance = tol;
    }
    return st;
}
}  // namespace

extern "C" {

// FiniteElementElasticModel3d::ensure_model_matrix_storage_initialized (simulation_toolbox/src/fem/
// fe_model.rs:490-515) + compute_representative_force (:423-431) + create_preconditioner
// (newton_cg.rs:108-144)
int orc_sim_create(const orc_model* model, const orc_sim_desc* d, orc_sim** out) {
    auto* s = new orc_sim();
    s->model = model;
    s->d = *d;
    s->dir.assign(d->dirichlet_nodes, d->dirichlet_nodes + d->num_dirichlet);
    s->bc_u.assign(3 * s->dir.size(), 0.0);
    s->bc_v.assign(3 * s->dir.size(), 0.0);
    s->M = model->pattern;
    int st = orc_model_assemble_mass(model, d->density, s->M.val.data());
    if (st != ORC_OK) {
        delete s;
        return st;
    }
    const int64_t n = s->M.n;
    orc_dirichlet_csr(n, s->M.rowptr.data(), s->M.col.data(), s->M.val.data(), (int64_t)s->dir.size(), s->dir.data(), 3);
    const double a = 10.0;
    std::vector<double> arep(n, std::sqrt((a * a) / 3.0)), tmp(n, 0.0);
    orc_spmv(n, s->M.rowptr.data(), s->M.col.data(), s->M.val.data(), 0.0, tmp.data(), 1.0, arep.data());
    s->rep_force = norm2(tmp);
    s->K = model->pattern;
    s->J = model->pattern;
    s->D = model->pattern;
    std::fill(s->D.val.begin(), s->D.val.end(), 0.0);
    if (!(d->mass_damping_coefficient >= 0.0) || !(d->stiffness_damping_coefficient >= 0.0)) {
        delete s;
        return fail(ORC_ERR_INVALID, "damping coefficients must be >= 0");
    }
    s->damping = d->mass_damping_coefficient > 0.0 || d->stiffness_damping_coefficient > 0.0;
    // compute_gravity_density (system_assembly.rs:30-75) then M * g
    std::vector<double> gd(n, 0.0);
    if (d->gravity_enabled)
        for (int64_t i = 0; i < n; ++i)
            if ((i % 3) + 1 == 3) gd[i] = d->gravity;
    s->gforce.assign(n, 0.0);
    orc_spmv(n, s->M.rowptr.data(), s->M.col.data(), s->M.val.data(), 0.0, s->gforce.data(), 1.0, gd.data());
    if (d->num_subdomains > 0) {
        st = orc_schwarz_create(model->nv, d->num_subdomains, d->subdomain_offsets, d->subdomain_nodes,
                                d->num_untreated, d->untreated_nodes, &s->schwarz);
        if (st != ORC_OK) {
            delete s;
            return st;
        }
    }
    *out = s;
    return ORC_OK;
}

void orc_sim_destroy(orc_sim* s) {
    if (!s) return;
    if (s->schwarz) orc_schwarz_destroy(s->schwarz);
    delete s;
}

double orc_sim_representative_force(const orc_sim* s) { return s->rep_force; }

// DirichletBoundaryConditions::apply_displacement_bcs / apply_velocity_bcs values of the sim's
// Dirichlet nodes (simulation_toolbox/src/fem/bcs.rs:13-27), three per node; NULL: zeros
int orc_sim_set_dirichlet_values(orc_sim* s, const double* u_values, const double* v_values) {
    const size_t m = 3 * s->dir.size();
    for (size_t k = 0; k < m; ++k) {
        s->bc_u[k] = u_values ? u_values[k] : 0.0;
        s->bc_v[k] = v_values ? v_values[k] : 0.0;
    }
    return ORC_OK;
}

int orc_sim_mass_values(const orc_sim* s, double* values) {
    std::copy(s->M.val.begin(), s->M.val.end(), values);
    return ORC_OK;
}

int orc_sim_damping_values(const orc_sim* s, double* values) {
    std::copy(s->D.val.begin(), s->D.val.end(), values);
    return ORC_OK;
}

int orc_sim_jacobian_values(const orc_sim* s, double* values) {
    std::copy(s->J.val.begin(), s->J.val.end(), values);
    return ORC_OK;
}

// BackwardEulerImplicitFunction::eval_into (hamilton2/src/integrators/euler.rs:97-114)
int orc_sim_residual(orc_sim* s, double t, double dt, const double* u, const double* v, const double* w, double* F) {
    (void)t;
    const int64_t n = s->M.n;
    std::vector<double> uu(n), vv(n);
    for (int64_t i = 0; i < n; ++i) {
        uu[i] = (u[i] + v[i] * dt) + w[i] * dt;
        vv[i] = v[i] + w[i];
    }
    int st = eval_f(s, uu.data(), vv.data(), F);
    if (st != ORC_OK) return st;
    for (int64_t i = 0; i < n; ++i) F[i] *= dt * -1.0;
    orc_spmv(n, s->M.rowptr.data(), s->M.col.data(), s->M.val.data(), 1.0, F, 1.0, w);
    return ORC_OK;
}

// integrate_model_with_substeps (simulation_toolbox/src/fem/newton_cg.rs:146-228) and
// integrate_model (:230-295)
int orc_sim_step(orc_sim* s, double t, double dt, double* u, double* v, orc_step_stats* stats) {
#ifdef _OPENMP
    if (s->d.num_threads > 0) omp_set_num_threads(s->d.num_threads);
#endif
    orc_step_stats local{};
    s->stats = stats ? stats : &local;
    *s->stats = orc_step_stats{};
    const int64_t n = s->M.n;
    apply_bc_unknowns(s, u, v);
    auto step = [&](double ts, double sdt, double* us, double* vs) -> int {
        int64_t it = 0;
        if (s->damping) {
            const int dst = compute_damping(s, us);
            if (dst != ORC_OK) return dst;
        }
        const int st = be_step(s, ts, sdt, us, vs, &it);
        s->stats->newton_iterations += it;
        if (st != ORC_OK) return st;
        apply_bc_unknowns(s, us, vs);
        return ORC_OK;
    };
    int num_substeps = 0, last = ORC_OK;
    std::vector<double> u0(u, u + n), v0(v, v + n);
    const int sst = substep_loop(n, t, dt, u0.data(), v0.data(), step, &num_substeps, &last);
    if (sst != ORC_OK) {
        s->stats = nullptr;
        return sst;
    }
    std::vector<double> us(u0), vs(v0);
    s->stats->substeps = num_substeps;
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(us[i]) || !std::isfinite(vs[i])) {
            s->stats = nullptr;
            return fail(ORC_ERR_NOT_FINITE, "non-finite state after step");
        }
    std::copy(us.begin(), us.end(), u);
    std::copy(vs.begin(), vs.end(), v);
    s->stats = nullptr;
    return ORC_OK;
}

}  // extern "C"

// ====================================================================== learned correction network
namespace {
constexpr int MI = 64, MH = 128, MO = 64;
constexpr int P_W1 = 0, P_B1 = P_W1 + MH * MI, P_W2 = P_B1 + MH, P_B2 = P_W2 + MH * MH, P_W3 = P_B2 + MH,
              P_B3 = P_W3 + MO * MH, P_N = P_B3 + MO;

float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u = ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
double gelu_d(double x) { return 0.5 * x * (1.0 + std::erf(x / std::sqrt(2.0))); }
double gelu_grad_d(double x) {
    return 0.5 * (1.0 + std::erf(x / std::sqrt(2.0))) + x * std::exp(-0.5 * x * x) / std::sqrt(2.0 * M_PI);
}

struct MlpTape {
    std::vector<double> x, z1, h1, z2, h2, y;
};

void mlp_fwd_one(const float* p, const float* xin, bool em, MlpTape& t) {
    auto w = [&](int i) { return em ? (double)bf16r(p[i]) : (double)p[i]; };
    t.x.resize(MI);
    t.z1.resize(MH);
    t.h1.resize(MH);
    t.z2.resize(MH);
    t.h2.resize(MH);
    t.y.resize(MO);
    for (int i = 0; i < MI; ++i) t.x[i] = em ? (double)bf16r(xin[i]) : (double)xin[i];
    for (int o = 0; o < MH; ++o) {
        double a = 0.0;
        for (int i = 0; i < MI; ++i) a += w(P_W1 + o * MI + i) * t.x[i];
        a += p[P_B1 + o];
        t.z1[o] = a;
        t.h1[o] = em ? (double)bf16r((float)gelu_d(a)) : gelu_d(a);
    }
    for (int o = 0; o < MH; ++o) {
        double a = 0.0;
        for (int i = 0; i < MH; ++i) a += w(P_W2 + o * MH + i) * t.h1[i];
        a += p[P_B2 + o];
        t.z2[o] = a;
        t.h2[o] = em ? (double)bf16r((float)gelu_d(a)) : gelu_d(a);
    }
    for (int o = 0; o < MO; ++o) {
        double a = 0.0;
        for (int i = 0; i < MH; ++i) a += w(P_W3 + o * MH + i) * t.h2[i];
        t.y[o] = a + p[P_B3 + o];
    }
}
}  // namespace

extern "C" {

int orc_mlp_forward(const float* params, int64_t n, const float* x, int emulate_bf16, double* y) {
    if (!params || (n > 0 && (!x || !y)) || n < 0) return fail(ORC_ERR_INVALID, "bad network arguments");
    MlpTape t;
    for (int64_t s = 0; s < n; ++s) {
        mlp_fwd_one(params, x + s * MI, emulate_bf16 != 0, t);
        std::copy(t.y.begin(), t.y.end(), y + s * MO);
    }
    return ORC_OK;
}

int orc_mlp_grad(const float* params, int64_t n, const float* x, const float* dy, int emulate_bf16, double* grad) {
    if (!params || n < 1 || !x || !dy || !grad) return fail(ORC_ERR_INVALID, "bad network arguments");
    const bool em = emulate_bf16 != 0;
    auto w = [&](int i) { return em ? (double)bf16r(params[i]) : (double)params[i]; };
    std::fill(grad, grad + P_N, 0.0);
    MlpTape t;
    std::vector<double> d3(MO), d2(MH), d1(MH);
    for (int64_t s = 0; s < n; ++s) {
        mlp_fwd_one(params, x + s * MI, em, t);
        for (int o = 0; o < MO; ++o) d3[o] = dy[s * MO + o];
        for (int o = 0; o < MO; ++o) {
            for (int i = 0; i < MH; ++i) grad[P_W3 + o * MH + i] += d3[o] * t.h2[i];
            grad[P_B3 + o] += d3[o];
        }
        for (int i = 0; i < MH; ++i) {
            double a = 0.0;
            for (int o = 0; o < MO; ++o) a += w(P_W3 + o * MH + i) * d3[o];
            d2[i] = a * gelu_grad_d(t.z2[i]);
        }
        for (int o = 0; o < MH; ++o) {
            for (int i = 0; i < MH; ++i) grad[P_W2 + o * MH + i] += d2[o] * t.h1[i];
            grad[P_B2 + o] += d2[o];
        }
        for (int i = 0; i < MH; ++i) {
            double a = 0.0;
            for (int o = 0; o < MH; ++o) a += w(P_W2 + o * MH + i) * d2[o];
            d1[i] = a * gelu_grad_d(t.z1[i]);
        }
        for (int o = 0; o < MH; ++o) {
            for (int i = 0; i < MI; ++i) grad[P_W1 + o * MI + i] += d1[o] * t.x[i];
            grad[P_B1 + o] += d1[o];
        }
    }
    return ORC_OK;
}

}  // extern "C"
