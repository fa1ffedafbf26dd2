// Device functions: Lagrange element bases (fenris/src/element.rs) and the Stable
// Neo-Hookean material (fenris/src/solid/materials.rs). 3x3 matrices are column-major.
#pragma once

#include <hip/hip_runtime.h>

namespace hoefem {
namespace dev {

template <int ET>
struct ElemTraits;
template <> struct ElemTraits<0> { static constexpr int N = 4, GN = 4; static constexpr bool TET = true; };
template <> struct ElemTraits<1> { static constexpr int N = 10, GN = 4; static constexpr bool TET = true; };
template <> struct ElemTraits<2> { static constexpr int N = 20, GN = 4; static constexpr bool TET = true; };
template <> struct ElemTraits<3> { static constexpr int N = 8, GN = 8; static constexpr bool TET = false; };
template <> struct ElemTraits<4> { static constexpr int N = 20, GN = 8; static constexpr bool TET = false; };
template <> struct ElemTraits<5> { static constexpr int N = 27, GN = 8; static constexpr bool TET = false; };

// Hex27 reference nodes (the first 20 are the Hex20 nodes, the first 8 the Hex8 nodes),
// fenris/src/element.rs:1250-1298.
__constant__ static const signed char HEXN[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {-1, 0, -1}, {-1, -1, 0}, {1, 0, -1}, {1, -1, 0}, {0, 1, -1},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1},  {-1, 0, 1},  {1, 0, 1},  {0, 1, 1},  {0, 0, -1},
    {0, -1, 0},   {-1, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},  {0, 0, 0}};
__constant__ static const double TETG[4][3] = {{-0.5, -0.5, -0.5}, {0.5, 0.0, 0.0}, {0.0, 0.5, 0.0}, {0.0, 0.0, 0.5}};
__constant__ static const signed char TET10E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
__constant__ static const signed char TET20E[12][2] = {{0, 1}, {1, 0}, {0, 2}, {2, 0}, {0, 3}, {3, 0},
                                                       {1, 2}, {2, 1}, {1, 3}, {3, 1}, {2, 3}, {3, 2}};
__constant__ static const signed char TET20F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

__device__ inline void tet4_psi(const double* xi, double* p) {
    p[0] = -0.5 * xi[0] - 0.5 * xi[1] - 0.5 * xi[2] - 0.5;
    p[1] = 0.5 * xi[0] + 0.5;
    p[2] = 0.5 * xi[1] + 0.5;
    p[3] = 0.5 * xi[2] + 0.5;
}
__device__ inline double lin1(double a, double x) { return (1.0 + a * x) / 2.0; }
__device__ inline double lin1g(double a) { return a / 2.0; }
__device__ inline double quad1(double a, double x) {
    const double a2 = a * a;
    const double x2 = x * x;
    return (3.0 / 2.0 * a2 - 1.0) * x2 + 0.5 * a * x + 1.0 - a2;
}
__device__ inline double quad1g(double a, double x) {
    const double a2 = a * a;
    return 2.0 * (3.0 / 2.0 * a2 - 1.0) * x + 0.5 * a;
}

// gradient of basis function I at xi (psi: Tet4 barycentric values, used by the tet family)
template <int ET>
__device__ inline void grad_node(int I, const double* xi, const double* p, double* g) {
    if constexpr (ET == 0) {
        g[0] = TETG[I][0]; g[1] = TETG[I][1]; g[2] = TETG[I][2];
    } else if constexpr (ET == 1) {
        if (I < 4) {
            const double s = 4.0 * p[I] - 1.0;
            for (int k = 0; k < 3; ++k) g[k] = TETG[I][k] * s;
        } else {
            const int i = TET10E[I - 4][0], j = TET10E[I - 4][1];
            for (int k = 0; k < 3; ++k) g[k] = TETG[i][k] * (4.0 * p[j]) + TETG[j][k] * (4.0 * p[i]);
        }
    } else if constexpr (ET == 2) {
        if (I < 4) {
            const double s = 0.5 * (27.0 * p[I] * p[I] - 18.0 * p[I] + 2.0);
            for (int k = 0; k < 3; ++k) g[k] = TETG[I][k] * s;
        } else if (I < 16) {
            const int a = TET20E[I - 4][0], b = TET20E[I - 4][1];
            const double pa = p[a], pb = p[b];
            for (int k = 0; k < 3; ++k)
                g[k] = (TETG[a][k] * (pb * (6.0 * pa - 1.0)) + TETG[b][k] * (pa * (3.0 * pa - 1.0))) * (9.0 / 2.0);
        } else {
            const int a = TET20F[I - 16][0], b = TET20F[I - 16][1], c = TET20F[I - 16][2];
            for (int k = 0; k < 3; ++k)
                g[k] = (TETG[a][k] * p[b] * p[c] + TETG[b][k] * p[a] * p[c] + TETG[c][k] * p[a] * p[b]) * 27.0;
        }
    } else if constexpr (ET == 3) {
        const double a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        g[0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
        g[1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
        g[2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
    } else if constexpr (ET == 5) {
        const double a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        g[0] = quad1g(a, xi[0]) * quad1(b, xi[1]) * quad1(c, xi[2]);
        g[1] = quad1(a, xi[0]) * quad1g(b, xi[1]) * quad1(c, xi[2]);
        g[2] = quad1(a, xi[0]) * quad1(b, xi[1]) * quad1g(c, xi[2]);
    } else {  // Hex20, fenris/src/element.rs:1556-1625
        const double a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        if (I < 8) {
            const double f = a * xi[0] + b * xi[1] + c * xi[2] - 2.0;
            const double gg = (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]);
            const double s = 1.0 / 8.0;
            g[0] = s * (a * gg + f * a * (1.0 + b * xi[1]) * (1.0 + c * xi[2]));
            g[1] = s * (b * gg + f * b * (1.0 + a * xi[0]) * (1.0 + c * xi[2]));
            g[2] = s * (c * gg + f * c * (1.0 + a * xi[0]) * (1.0 + b * xi[1]));
        } else {
            const double a2 = a * a, b2 = b * b, c2 = c * c;
            const double h = (1.0 - (1.0 - a2) * xi[0] * xi[0]) * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
                             (1.0 - (1.0 - c2) * xi[2] * xi[2]);
            const double gg = (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]);
            const double s = 1.0 / 4.0;
            const double dh0 = -2.0 * (1.0 - a2) * xi[0] * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
                               (1.0 - (1.0 - c2) * xi[2] * xi[2]);
            const double dh1 = -2.0 * (1.0 - b2) * xi[1] * (1.0 - (1.0 - a2) * xi[0] * xi[0]) *
                               (1.0 - (1.0 - c2) * xi[2] * xi[2]);
            const double dh2 = -2.0 * (1.0 - c2) * xi[2] * (1.0 - (1.0 - a2) * xi[0] * xi[0]) *
                               (1.0 - (1.0 - b2) * xi[1] * xi[1]);
            g[0] = s * (dh0 * gg + h * a * (1.0 + b * xi[1]) * (1.0 + c * xi[2]));
            g[1] = s * (dh1 * gg + h * b * (1.0 + a * xi[0]) * (1.0 + c * xi[2]));
            g[2] = s * (dh2 * gg + h * c * (1.0 + a * xi[0]) * (1.0 + b * xi[1]));
        }
    }
}

// value of basis function I at xi
template <int ET>
__device__ inline double basis_node(int I, const double* xi, const double* p) {
    if constexpr (ET == 0) {
        return p[I];
    } else if constexpr (ET == 1) {
        if (I < 4) return p[I] * (2.0 * p[I] - 1.0);
        return 4.0 * p[TET10E[I - 4][0]] * p[TET10E[I - 4][1]];
    } else if constexpr (ET == 2) {
        if (I < 4) return 0.5 * p[I] * (3.0 * p[I] - 1.0) * (3.0 * p[I] - 2.0);
        if (I < 16) {
            const int c = TET20E[I - 4][0], o = TET20E[I - 4][1];
            return (9.0 / 2.0) * p[c] * p[o] * (3.0 * p[c] - 1.0);
        }
        return 27.0 * p[TET20F[I - 16][0]] * p[TET20F[I - 16][1]] * p[TET20F[I - 16][2]];
    } else if constexpr (ET == 3) {
        return lin1(HEXN[I][0], xi[0]) * lin1(HEXN[I][1], xi[1]) * lin1(HEXN[I][2], xi[2]);
    } else if constexpr (ET == 5) {
        return quad1(HEXN[I][0], xi[0]) * quad1(HEXN[I][1], xi[1]) * quad1(HEXN[I][2], xi[2]);
    } else {
        const double a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        if (I < 8)
            return (1.0 / 8.0) * (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]) *
                   (a * xi[0] + b * xi[1] + c * xi[2] - 2.0);
        const double a2 = a * a, b2 = b * b, c2 = c * c;
        return (1.0 / 4.0) * (1.0 - (1.0 - a2) * xi[0] * xi[0]) * (1.0 - (1.0 - b2) * xi[1] * xi[1]) *
               (1.0 - (1.0 - c2) * xi[2] * xi[2]) * (1.0 + a * xi[0]) * (1.0 + b * xi[1]) * (1.0 + c * xi[2]);
    }
}

// reference Jacobian J = X * G_geom^T (Tet4 / Hex8 geometry). X: [GN][3]
template <int ET>
__device__ inline void ref_jacobian(const double* X, const double* xi, double* J) {
    constexpr int GN = ElemTraits<ET>::GN;
    double G[GN][3];
    if constexpr (ElemTraits<ET>::TET) {
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) G[k][c] = TETG[k][c];
    } else {
        for (int k = 0; k < 8; ++k) {
            const double a = HEXN[k][0], b = HEXN[k][1], c = HEXN[k][2];
            G[k][0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
            G[k][1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
            G[k][2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < GN; ++k) s += X[3 * k + i] * G[k][c];
            J[i + 3 * c] = s;
        }
}

__device__ inline double det3(const double* m) {
    const double m11 = m[0], m12 = m[3], m13 = m[6];
    const double m21 = m[1], m22 = m[4], m23 = m[7];
    const double m31 = m[2], m32 = m[5], m33 = m[8];
    const double minor_1 = m22 * m33 - m32 * m23;
    const double minor_2 = m21 * m33 - m31 * m23;
    const double minor_3 = m21 * m32 - m31 * m22;
    return m11 * minor_1 - m12 * minor_2 + m13 * minor_3;
}

__device__ inline bool inv3(const double* m, double* out) {
    const double m11 = m[0], m12 = m[3], m13 = m[6];
    const double m21 = m[1], m22 = m[4], m23 = m[7];
    const double m31 = m[2], m32 = m[5], m33 = m[8];
    const double a = m22 * m33 - m32 * m23;
    const double b = m21 * m33 - m31 * m23;
    const double c = m21 * m32 - m31 * m22;
    const double det = m11 * a - m12 * b + m13 * c;
    if (det == 0.0) return false;
    out[0] = a / det;
    out[3] = (m13 * m32 - m33 * m12) / det;
    out[6] = (m12 * m23 - m22 * m13) / det;
    out[1] = -b / det;
    out[4] = (m11 * m33 - m31 * m13) / det;
    out[7] = (m13 * m21 - m23 * m11) / det;
    out[2] = c / det;
    out[5] = (m12 * m31 - m32 * m11) / det;
    out[8] = (m11 * m22 - m21 * m12) / det;
    return true;
}

// reparametrize_lame_for_stable_neo_hookean (fenris/src/solid/materials.rs:243-253), d = 3
__host__ __device__ inline void snh_reparam(double mu, double lambda, double* mu_r, double* lambda_r) {
    const double d = 3.0;
    *mu_r = (d + 1.0) / d * mu;
    *lambda_r = lambda + (1.0 - 2.0 / (d * (d + 1.0))) * mu;
}

// compute_stress_tensor (fenris/src/solid/materials.rs:299-322); reparametrized parameters
__device__ inline bool snh_stress(double mu, double lambda, const double* F, double* P) {
    const double d = 3.0;
    const double alpha = 1.0 + mu / lambda * (d / (d + 1.0));
    double I_C = 0.0;
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const double J = det3(F);
    double Finv[9];
    if (!inv3(F, Finv)) return false;
    const double s1 = mu * (1.0 - 1.0 / (I_C + 1.0));
    const double s2 = lambda * (J - alpha);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P[r + 3 * c] = F[r + 3 * c] * s1 + (Finv[c + 3 * r] * J) * s2;
    return true;
}

// build_stable_neohookean_dp_df (fenris/src/solid/materials.rs:506-557) with
// add_volume_hessian_3d (:763-776). dp: column-major 9x9.
__device__ inline bool snh_dpdf(double mu, double lambda, const double* F, double* dp) {
    const double d = 3.0;
    const double alpha = 1.0 + mu / lambda * (d / (d + 1.0));
    double I_C = 0.0;
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const double J = det3(F);
    const double beta = I_C + 1.0;
    const double beta2 = beta * beta;
    double Finv[9];
    if (!inv3(F, Finv)) return false;
    double g[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) g[3 * c + r] = Finv[c + 3 * r] * J;
    const double diag = mu * (1.0 - 1.0 / beta);
    const double cf = mu * 2.0 / beta2;
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r) {
            double v = (r == c) ? diag : 0.0;
            v += cf * F[r] * F[c];
            v += lambda * g[r] * g[c];
            dp[r + 9 * c] = v;
        }
    const double scale = lambda * (J - alpha);
    // cross_product_matrix(F_col_k) * scale placed in the off-diagonal 3x3 blocks
    for (int k = 0; k < 3; ++k) {
        const double x0 = F[3 * k] * scale, x1 = F[3 * k + 1] * scale, x2 = F[3 * k + 2] * scale;
        // hat (column-major): [0, x2, -x1, -x2, 0, x0, x1, -x0, 0]
        const double h[9] = {0.0, x2, -x1, -x2, 0.0, x0, x1, -x0, 0.0};
        int r0a, c0a, r0b, c0b;
        if (k == 2) { r0a = 3; c0a = 0; r0b = 0; c0b = 3; }       // +f2 at (3,0), -f2 at (0,3)
        else if (k == 1) { r0a = 0; c0a = 6; r0b = 6; c0b = 0; }  // +f1 at (0,6), -f1 at (6,0)
        else { r0a = 6; c0a = 3; r0b = 3; c0b = 6; }              // +f0 at (6,3), -f0 at (3,6)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                dp[(r0a + r) + 9 * (c0a + c)] += h[r + 3 * c];
                dp[(r0b + r) + 9 * (c0b + c)] -= h[r + 3 * c];
            }
    }
    return true;
}

}  // namespace dev
}  // namespace hoefem
