// Device functions: Lagrange element bases (fenris/src/element.rs) and the Stable
// Neo-Hookean material (fenris/src/solid/materials.rs). 3x3 matrices are column-major.
// Templated on the arithmetic type T (double for the plain FP64 step, float for the
// FP32 element tangents of the mixed-precision step).
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
// Tet4 reference gradients times 2 (exact small integers; the 0.5 factor is applied in T)
__constant__ static const signed char TETG2[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
__constant__ static const signed char TET10E[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
__constant__ static const signed char TET20E[12][2] = {{0, 1}, {1, 0}, {0, 2}, {2, 0}, {0, 3}, {3, 0},
                                                       {1, 2}, {2, 1}, {1, 3}, {3, 1}, {2, 3}, {3, 2}};
__constant__ static const signed char TET20F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

template <class T>
__device__ inline T tetg(int i, int k) {
    return T(0.5) * T(TETG2[i][k]);
}

template <class T>
__device__ inline void tet4_psi(const T* xi, T* p) {
    p[0] = T(-0.5) * xi[0] - T(0.5) * xi[1] - T(0.5) * xi[2] - T(0.5);
    p[1] = T(0.5) * xi[0] + T(0.5);
    p[2] = T(0.5) * xi[1] + T(0.5);
    p[3] = T(0.5) * xi[2] + T(0.5);
}
template <class T>
__device__ inline T lin1(T a, T x) { return (T(1) + a * x) / T(2); }
template <class T>
__device__ inline T lin1g(T a) { return a / T(2); }
template <class T>
__device__ inline T quad1(T a, T x) {
    const T a2 = a * a;
    const T x2 = x * x;
    return (T(1.5) * a2 - T(1)) * x2 + T(0.5) * a * x + T(1) - a2;
}
template <class T>
__device__ inline T quad1g(T a, T x) {
    const T a2 = a * a;
    return T(2) * (T(1.5) * a2 - T(1)) * x + T(0.5) * a;
}

// gradient of basis function I at xi (p: Tet4 barycentric values, used by the tet family)
template <int ET, class T>
__device__ inline void grad_node(int I, const T* xi, const T* p, T* g) {
    if constexpr (ET == 0) {
        for (int k = 0; k < 3; ++k) g[k] = tetg<T>(I, k);
    } else if constexpr (ET == 1) {
        if (I < 4) {
            const T s = T(4) * p[I] - T(1);
            for (int k = 0; k < 3; ++k) g[k] = tetg<T>(I, k) * s;
        } else {
            const int i = TET10E[I - 4][0], j = TET10E[I - 4][1];
            for (int k = 0; k < 3; ++k) g[k] = tetg<T>(i, k) * (T(4) * p[j]) + tetg<T>(j, k) * (T(4) * p[i]);
        }
    } else if constexpr (ET == 2) {
        if (I < 4) {
            const T s = T(0.5) * (T(27) * p[I] * p[I] - T(18) * p[I] + T(2));
            for (int k = 0; k < 3; ++k) g[k] = tetg<T>(I, k) * s;
        } else if (I < 16) {
            const int a = TET20E[I - 4][0], b = TET20E[I - 4][1];
            const T pa = p[a], pb = p[b];
            for (int k = 0; k < 3; ++k)
                g[k] = (tetg<T>(a, k) * (pb * (T(6) * pa - T(1))) + tetg<T>(b, k) * (pa * (T(3) * pa - T(1)))) * T(4.5);
        } else {
            const int a = TET20F[I - 16][0], b = TET20F[I - 16][1], c = TET20F[I - 16][2];
            for (int k = 0; k < 3; ++k)
                g[k] = (tetg<T>(a, k) * p[b] * p[c] + tetg<T>(b, k) * p[a] * p[c] + tetg<T>(c, k) * p[a] * p[b]) * T(27);
        }
    } else if constexpr (ET == 3) {
        const T a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        g[0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
        g[1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
        g[2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
    } else if constexpr (ET == 5) {
        const T a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        g[0] = quad1g(a, xi[0]) * quad1(b, xi[1]) * quad1(c, xi[2]);
        g[1] = quad1(a, xi[0]) * quad1g(b, xi[1]) * quad1(c, xi[2]);
        g[2] = quad1(a, xi[0]) * quad1(b, xi[1]) * quad1g(c, xi[2]);
    } else {  // Hex20, fenris/src/element.rs:1556-1625
        const T a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        if (I < 8) {
            const T f = a * xi[0] + b * xi[1] + c * xi[2] - T(2);
            const T gg = (T(1) + a * xi[0]) * (T(1) + b * xi[1]) * (T(1) + c * xi[2]);
            const T s = T(0.125);
            g[0] = s * (a * gg + f * a * (T(1) + b * xi[1]) * (T(1) + c * xi[2]));
            g[1] = s * (b * gg + f * b * (T(1) + a * xi[0]) * (T(1) + c * xi[2]));
            g[2] = s * (c * gg + f * c * (T(1) + a * xi[0]) * (T(1) + b * xi[1]));
        } else {
            const T a2 = a * a, b2 = b * b, c2 = c * c;
            const T h = (T(1) - (T(1) - a2) * xi[0] * xi[0]) * (T(1) - (T(1) - b2) * xi[1] * xi[1]) *
                        (T(1) - (T(1) - c2) * xi[2] * xi[2]);
            const T gg = (T(1) + a * xi[0]) * (T(1) + b * xi[1]) * (T(1) + c * xi[2]);
            const T s = T(0.25);
            const T dh0 = T(-2) * (T(1) - a2) * xi[0] * (T(1) - (T(1) - b2) * xi[1] * xi[1]) *
                          (T(1) - (T(1) - c2) * xi[2] * xi[2]);
            const T dh1 = T(-2) * (T(1) - b2) * xi[1] * (T(1) - (T(1) - a2) * xi[0] * xi[0]) *
                          (T(1) - (T(1) - c2) * xi[2] * xi[2]);
            const T dh2 = T(-2) * (T(1) - c2) * xi[2] * (T(1) - (T(1) - a2) * xi[0] * xi[0]) *
                          (T(1) - (T(1) - b2) * xi[1] * xi[1]);
            g[0] = s * (dh0 * gg + h * a * (T(1) + b * xi[1]) * (T(1) + c * xi[2]));
            g[1] = s * (dh1 * gg + h * b * (T(1) + a * xi[0]) * (T(1) + c * xi[2]));
            g[2] = s * (dh2 * gg + h * c * (T(1) + a * xi[0]) * (T(1) + b * xi[1]));
        }
    }
}

// value of basis function I at xi
template <int ET, class T>
__device__ inline T basis_node(int I, const T* xi, const T* p) {
    if constexpr (ET == 0) {
        return p[I];
    } else if constexpr (ET == 1) {
        if (I < 4) return p[I] * (T(2) * p[I] - T(1));
        return T(4) * p[TET10E[I - 4][0]] * p[TET10E[I - 4][1]];
    } else if constexpr (ET == 2) {
        if (I < 4) return T(0.5) * p[I] * (T(3) * p[I] - T(1)) * (T(3) * p[I] - T(2));
        if (I < 16) {
            const int c = TET20E[I - 4][0], o = TET20E[I - 4][1];
            return T(4.5) * p[c] * p[o] * (T(3) * p[c] - T(1));
        }
        return T(27) * p[TET20F[I - 16][0]] * p[TET20F[I - 16][1]] * p[TET20F[I - 16][2]];
    } else if constexpr (ET == 3) {
        return lin1(T(HEXN[I][0]), xi[0]) * lin1(T(HEXN[I][1]), xi[1]) * lin1(T(HEXN[I][2]), xi[2]);
    } else if constexpr (ET == 5) {
        return quad1(T(HEXN[I][0]), xi[0]) * quad1(T(HEXN[I][1]), xi[1]) * quad1(T(HEXN[I][2]), xi[2]);
    } else {
        const T a = HEXN[I][0], b = HEXN[I][1], c = HEXN[I][2];
        if (I < 8)
            return T(0.125) * (T(1) + a * xi[0]) * (T(1) + b * xi[1]) * (T(1) + c * xi[2]) *
                   (a * xi[0] + b * xi[1] + c * xi[2] - T(2));
        const T a2 = a * a, b2 = b * b, c2 = c * c;
        return T(0.25) * (T(1) - (T(1) - a2) * xi[0] * xi[0]) * (T(1) - (T(1) - b2) * xi[1] * xi[1]) *
               (T(1) - (T(1) - c2) * xi[2] * xi[2]) * (T(1) + a * xi[0]) * (T(1) + b * xi[1]) * (T(1) + c * xi[2]);
    }
}

// reference Jacobian J = X * G_geom^T (Tet4 / Hex8 geometry). X: [GN][3]
template <int ET, class T>
__device__ inline void ref_jacobian(const T* X, const T* xi, T* J) {
    constexpr int GN = ElemTraits<ET>::GN;
    T G[GN][3];
    if constexpr (ElemTraits<ET>::TET) {
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) G[k][c] = tetg<T>(k, c);
    } else {
        for (int k = 0; k < 8; ++k) {
            const T a = HEXN[k][0], b = HEXN[k][1], c = HEXN[k][2];
            G[k][0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
            G[k][1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
            G[k][2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            T s = T(0);
            for (int k = 0; k < GN; ++k) s += X[3 * k + i] * G[k][c];
            J[i + 3 * c] = s;
        }
}

template <class T>
__device__ inline T det3(const T* m) {
    const T m11 = m[0], m12 = m[3], m13 = m[6];
    const T m21 = m[1], m22 = m[4], m23 = m[7];
    const T m31 = m[2], m32 = m[5], m33 = m[8];
    const T minor_1 = m22 * m33 - m32 * m23;
    const T minor_2 = m21 * m33 - m31 * m23;
    const T minor_3 = m21 * m32 - m31 * m22;
    return m11 * minor_1 - m12 * minor_2 + m13 * minor_3;
}

template <class T>
__device__ inline bool inv3(const T* m, T* out) {
    const T m11 = m[0], m12 = m[3], m13 = m[6];
    const T m21 = m[1], m22 = m[4], m23 = m[7];
    const T m31 = m[2], m32 = m[5], m33 = m[8];
    const T a = m22 * m33 - m32 * m23;
    const T b = m21 * m33 - m31 * m23;
    const T c = m21 * m32 - m31 * m22;
    const T det = m11 * a - m12 * b + m13 * c;
    if (det == T(0)) return false;
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
template <class T>
__device__ inline bool snh_stress(T mu, T lambda, const T* F, T* P) {
    const T alpha = T(1) + mu / lambda * T(0.75);
    T I_C = T(0);
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const T J = det3(F);
    T Finv[9];
    if (!inv3(F, Finv)) return false;
    const T s1 = mu * (T(1) - T(1) / (I_C + T(1)));
    const T s2 = lambda * (J - alpha);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P[r + 3 * c] = F[r + 3 * c] * s1 + (Finv[c + 3 * r] * J) * s2;
    return true;
}

// build_stable_neohookean_dp_df (fenris/src/solid/materials.rs:506-557) with
// add_volume_hessian_3d (:763-776). dp: column-major 9x9.
template <class T>
__device__ inline bool snh_dpdf(T mu, T lambda, const T* F, T* dp) {
    const T alpha = T(1) + mu / lambda * T(0.75);
    T I_C = T(0);
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const T J = det3(F);
    const T beta = I_C + T(1);
    const T beta2 = beta * beta;
    T Finv[9];
    if (!inv3(F, Finv)) return false;
    T g[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) g[3 * c + r] = Finv[c + 3 * r] * J;
    const T diag = mu * (T(1) - T(1) / beta);
    const T cf = mu * T(2) / beta2;
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r) {
            T v = (r == c) ? diag : T(0);
            v += cf * F[r] * F[c];
            v += lambda * g[r] * g[c];
            dp[r + 9 * c] = v;
        }
    const T scale = lambda * (J - alpha);
    // cross_product_matrix(F_col_k) * scale placed in the off-diagonal 3x3 blocks
    for (int k = 0; k < 3; ++k) {
        const T x0 = F[3 * k] * scale, x1 = F[3 * k + 1] * scale, x2 = F[3 * k + 2] * scale;
        const T h[9] = {T(0), x2, -x1, -x2, T(0), x0, x1, -x0, T(0)};  // column-major
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
