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
constexpr signed char HEX_NODES[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {-1, 0, -1}, {-1, -1, 0}, {1, 0, -1}, {1, -1, 0}, {0, 1, -1},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1},  {-1, 0, 1},  {1, 0, 1},  {0, 1, 1},  {0, 0, -1},
    {0, -1, 0},   {-1, 0, 0},  {1, 0, 0},   {0, 1, 0},   {0, 0, 1},  {0, 0, 0}};
// Per axis, coordinate + 1 of node I in bits 2I and 2I + 1: a lane decodes its node's coordinates
// with a shift (a lane-dependent index into a table in memory is a load on the critical path).
constexpr unsigned long long hex_axis_bits(int d) {
    unsigned long long m = 0;
    for (int I = 0; I < 27; ++I) m |= (unsigned long long)(HEX_NODES[I][d] + 1) << (2 * I);
    return m;
}
__device__ inline int hexn(int I, int d) {
    constexpr unsigned long long B0 = hex_axis_bits(0), B1 = hex_axis_bits(1), B2 = hex_axis_bits(2);
    const unsigned long long b = d == 0 ? B0 : (d == 1 ? B1 : B2);
    return (int)((b >> (2 * I)) & 3ull) - 1;
}
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
        const T a = hexn(I, 0), b = hexn(I, 1), c = hexn(I, 2);
        g[0] = lin1g(a) * lin1(b, xi[1]) * lin1(c, xi[2]);
        g[1] = lin1(a, xi[0]) * lin1g(b) * lin1(c, xi[2]);
        g[2] = lin1(a, xi[0]) * lin1(b, xi[1]) * lin1g(c);
    } else if constexpr (ET == 5) {
        const T a = hexn(I, 0), b = hexn(I, 1), c = hexn(I, 2);
        g[0] = quad1g(a, xi[0]) * quad1(b, xi[1]) * quad1(c, xi[2]);
        g[1] = quad1(a, xi[0]) * quad1g(b, xi[1]) * quad1(c, xi[2]);
        g[2] = quad1(a, xi[0]) * quad1(b, xi[1]) * quad1g(c, xi[2]);
    } else {  // Hex20, fenris/src/element.rs:1556-1625
        const T a = hexn(I, 0), b = hexn(I, 1), c = hexn(I, 2);
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
        return lin1(T(hexn(I, 0)), xi[0]) * lin1(T(hexn(I, 1)), xi[1]) * lin1(T(hexn(I, 2)), xi[2]);
    } else if constexpr (ET == 5) {
        return quad1(T(hexn(I, 0)), xi[0]) * quad1(T(hexn(I, 1)), xi[1]) * quad1(T(hexn(I, 2)), xi[2]);
    } else {
        const T a = hexn(I, 0), b = hexn(I, 1), c = hexn(I, 2);
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
            const T a = hexn(k, 0), b = hexn(k, 1), c = hexn(k, 2);
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

// Cofactor matrix cof(m) = det(m) m^-T (column-major), computed from the 2x2 minors without a
// division: J F^-T enters the stress and tangent, and on a nearly singular F the two
// determinants of inverse-times-determinant would not cancel exactly once contracted to FMAs.
template <class T>
__device__ inline void cof3(const T* m, T* C) {
    const T m11 = m[0], m12 = m[3], m13 = m[6];
    const T m21 = m[1], m22 = m[4], m23 = m[7];
    const T m31 = m[2], m32 = m[5], m33 = m[8];
    C[0] = m22 * m33 - m32 * m23;
    C[3] = -(m21 * m33 - m31 * m23);
    C[6] = m21 * m32 - m31 * m22;
    C[1] = m13 * m32 - m33 * m12;
    C[4] = m11 * m33 - m31 * m13;
    C[7] = m12 * m31 - m32 * m11;
    C[2] = m12 * m23 - m22 * m13;
    C[5] = m13 * m21 - m23 * m11;
    C[8] = m11 * m22 - m21 * m12;
}

// reparametrize_lame_for_stable_neo_hookean (fenris/src/solid/materials.rs:243-253), d = 3
__host__ __device__ inline void snh_reparam(double mu, double lambda, double* mu_r, double* lambda_r) {
    const double d = 3.0;
    *mu_r = (d + 1.0) / d * mu;
    *lambda_r = lambda + (1.0 - 2.0 / (d * (d + 1.0))) * mu;
}

// compute_strain_energy_density (fenris/src/solid/materials.rs:269-297): the energy density at F with
// the Lame parameters as given (mu, lambda) and alpha = 1 + mu_r / lambda_r * d / (d + 1) from the
// reparametrized ones (d = 3). A diagnostic: the Newton-CG step never evaluates it.
template <class T>
__device__ inline T snh_energy(T mu, T lambda, T mu_r, T lambda_r, const T* F) {
    const T d = T(3);
    const T a = mu_r / lambda_r * (d / (d + T(1)));
    const T alpha = T(1) + a;
    T I_C = T(0);
    for (int k = 0; k < 9; ++k) I_C += F[k] * F[k];
    const T Jma = det3(F) - alpha;
    return T(0.5) * mu * (I_C - d) + T(0.5) * lambda * (Jma * Jma - a * a) -
           T(0.5) * mu * log((I_C + T(1)) / (d + T(1)));
}

// compute_stress_tensor (fenris/src/solid/materials.rs:299-322); reparametrized parameters
template <class T>
__device__ inline bool snh_stress(T mu, T lambda, const T* F, T* P) {
    const T alpha = T(1) + mu / lambda * T(0.75);
    T I_C = T(0);
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const T J = det3(F);
    if (J == T(0)) return false;  // singular F (the reference's try_inverse fails)
    T g[9];                       // J F^-T
    cof3(F, g);
    const T s1 = mu * (T(1) - T(1) / (I_C + T(1)));
    const T s2 = lambda * (J - alpha);
    for (int k = 0; k < 9; ++k) P[k] = F[k] * s1 + g[k] * s2;
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
    if (J == T(0)) return false;
    T g[9];  // vec(J F^-T)
    cof3(F, g);
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


// Entry (R, C) of snh_dpdf's 9x9 matrix from the quantities every entry shares: F, g = vec(J F^-T),
// diag = mu (1 - 1 / beta), cf = 2 mu / beta^2 and scale = lambda (J - alpha). The cross-product
// blocks of add_volume_hessian_3d (fenris/src/solid/materials.rs:763-776) contribute +-F(m, k) scale
// to the entries of off-diagonal 3x3 blocks with r != c (m the third index, k the third block).
template <class T>
__device__ inline T snh_dpdf_entry(int R, int C, const T* F, const T* g, T diag, T cf, T lambda, T scale) {
    T v = (R == C) ? diag : T(0);
    v += cf * F[R] * F[C];
    v += lambda * g[R] * g[C];
    const int br = R / 3, bc = C / 3, r = R % 3, c = C % 3;
    if (br != bc && r != c) {
        const int m = 3 - r - c, k = 3 - br - bc;
        const bool cyclic = (r + 1) % 3 == c;  // (r, c, m) an even permutation
        const bool plus_block = (br == 1 && bc == 0) || (br == 0 && bc == 2) || (br == 2 && bc == 1);
        const T x = F[3 * k + m] * scale;
        v += (plus_block != cyclic) ? x : -x;
    }
    return v;
}

// The 3x3 block (br, bc) of snh_dpdf_entry's matrix, written into dp (column-major 9 x 9): the same
// arithmetic per entry, with the indices inside the block known at compile time, so a lane selects
// only the block's rows and columns of F and g instead of indexing by a lane-dependent entry.
template <class T>
__device__ inline T sel3(const T* v, int b, int i) {
    return b == 0 ? v[i] : (b == 1 ? v[3 + i] : v[6 + i]);
}
template <class T>
__device__ inline void snh_dpdf_block(int br, int bc, const T* F, const T* g, T diag, T cf, T lambda, T scale,
                                      T* dp) {
    const int k = br != bc ? 3 - br - bc : 0;  // the third block index (used when br != bc)
    const bool plus_block = (br == 1 && bc == 0) || (br == 0 && bc == 2) || (br == 2 && bc == 1);
    T Fr[3], Fc[3], gr[3], gc[3], Fk[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Fr[i] = sel3(F, br, i);
        Fc[i] = sel3(F, bc, i);
        gr[i] = sel3(g, br, i);
        gc[i] = sel3(g, bc, i);
        Fk[i] = sel3(F, k, i);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            T v = (br == bc && r == c) ? diag : T(0);
            v += cf * Fr[r] * Fc[c];
            v += lambda * gr[r] * gc[c];
            if (r != c && br != bc) {
                const int m = 3 - r - c;
                const bool cyclic = (r + 1) % 3 == c;
                const T x = Fk[m] * scale;
                v += (plus_block != cyclic) ? x : -x;
            }
            dp[(3 * br + r) + 9 * (3 * bc + c)] = v;
        }
}

// ------------------------------------------------------------------ material variants
// The elastic models of simulation_toolbox/src/fem/fe_model.rs:35-45 that wrap Stable
// Neo-Hookean: plain, InvertibleMaterial<StableNeoHookean> (F projected before the material,
// fenris/src/solid/materials.rs:862-945) and ProjectedStableNeoHookean (dP/dF projected onto
// the positive semi-definite cone, :387-441).
constexpr int MAT_SNH = 0, MAT_INVERTIBLE_SNH = 1, MAT_PROJECTED_SNH = 2;

template <class T>
struct Eps;
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };
template <> struct Eps<float> { static constexpr float v = 1.1920929e-7f; };

// One Jacobi rotation zeroing a(p,q) of the symmetric 3x3 matrix held as a[3][3] (only the
// entries touched are updated), accumulated into the columns p, q of V (column-major).
template <class T>
__device__ inline void jacobi3_rotate(T (&a)[3][3], T* V, int p, int q) {
    const T apq = a[p][q];
    if (apq == T(0)) return;
    const T theta = (a[q][q] - a[p][p]) / (T(2) * apq);
    const T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
    const T c = T(1) / sqrt(t * t + T(1));
    const T s = t * c;
    a[p][p] -= t * apq;
    a[q][q] += t * apq;
    a[p][q] = a[q][p] = T(0);
    const int r = 3 - p - q;
    const T arp = a[r][p], arq = a[r][q];
    a[r][p] = a[p][r] = c * arp - s * arq;
    a[r][q] = a[q][r] = s * arp + c * arq;
    for (int k = 0; k < 3; ++k) {
        const T vp = V[k + 3 * p], vq = V[k + 3 * q];
        V[k + 3 * p] = c * vp - s * vq;
        V[k + 3 * q] = s * vp + c * vq;
    }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi sweeps. A column-major (symmetric);
// w = eigenvalues (unsorted), Q = eigenvectors as columns.
template <class T>
__device__ inline void sym3_eigen(const T* A, T* w, T* Q) {
    T a[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) a[r][c] = T(0.5) * (A[r + 3 * c] + A[c + 3 * r]);
    for (int k = 0; k < 9; ++k) Q[k] = (k % 4 == 0) ? T(1) : T(0);
    for (int sweep = 0; sweep < 16; ++sweep) {
        const T off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const T dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= Eps<T>::v * Eps<T>::v * T(1e-2) * dia || off == T(0)) break;
        jacobi3_rotate(a, Q, 0, 1);
        jacobi3_rotate(a, Q, 0, 2);
        jacobi3_rotate(a, Q, 1, 2);
    }
    w[0] = a[0][0];
    w[1] = a[1][1];
    w[2] = a[2][2];
}

// rotation_svd (fenris/src/util.rs:68-94): F = U diag(s) V^T with U, V proper rotations and
// s sorted by magnitude, s[0] >= s[1] >= |s[2]|, the sign of det F carried by s[2].
// V from the eigenvectors of F^T F; U from F V (Gram-Schmidt on the two largest columns,
// the third as their cross product), so s[2] = u2 . (F v2) keeps its sign and absolute accuracy.
template <class T>
__device__ inline void rotation_svd3(const T* F, T* U, T* s, T* V) {
    T C[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            C[r + 3 * c] = F[3 * r] * F[3 * c] + F[3 * r + 1] * F[3 * c + 1] + F[3 * r + 2] * F[3 * c + 2];
    T w[3];
    sym3_eigen(C, w, V);
    // sort eigenpairs descending
    auto swap_cols = [&](int i, int j) {
        const T tw = w[i];
        w[i] = w[j];
        w[j] = tw;
        for (int k = 0; k < 3; ++k) {
            const T tv = V[k + 3 * i];
            V[k + 3 * i] = V[k + 3 * j];
            V[k + 3 * j] = tv;
        }
    };
    if (w[0] < w[1]) swap_cols(0, 1);
    if (w[0] < w[2]) swap_cols(0, 2);
    if (w[1] < w[2]) swap_cols(1, 2);
    if (det3(V) < T(0))
        for (int k = 0; k < 3; ++k) V[k + 6] = -V[k + 6];
    T b[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            b[r + 3 * c] = F[r] * V[3 * c] + F[r + 3] * V[3 * c + 1] + F[r + 6] * V[3 * c + 2];
    // u0
    T n0 = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (n0 > T(0)) {
        for (int k = 0; k < 3; ++k) U[k] = b[k] / n0;
    } else {
        U[0] = T(1);
        U[1] = T(0);
        U[2] = T(0);
    }
    // u1 = b1 - (u0 . b1) u0, normalized; any unit vector orthogonal to u0 if that vanishes
    const T d01 = U[0] * b[3] + U[1] * b[4] + U[2] * b[5];
    T u1[3] = {b[3] - d01 * U[0], b[4] - d01 * U[1], b[5] - d01 * U[2]};
    T n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    if (!(n1 > Eps<T>::v * n0) || n1 == T(0)) {
        // the coordinate axis least aligned with u0, orthogonalized
        int ax = 0;
        if (fabs(U[1]) < fabs(U[ax])) ax = 1;
        if (fabs(U[2]) < fabs(U[ax])) ax = 2;
        T e[3] = {T(0), T(0), T(0)};
        e[ax] = T(1);
        const T d = U[ax];
        for (int k = 0; k < 3; ++k) u1[k] = e[k] - d * U[k];
        n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    }
    for (int k = 0; k < 3; ++k) U[3 + k] = u1[k] / n1;
    U[6] = U[1] * U[5] - U[2] * U[4];
    U[7] = U[2] * U[3] - U[0] * U[5];
    U[8] = U[0] * U[4] - U[1] * U[3];
    for (int i = 0; i < 3; ++i) s[i] = U[3 * i] * b[3 * i] + U[3 * i + 1] * b[3 * i + 1] + U[3 * i + 2] * b[3 * i + 2];
}

// project_deformation_gradient (fenris/src/solid/materials.rs:876-915): signed singular
// values below the threshold are raised to it (an inverted F has its smallest one negative);
// F is recomposed only when something changed. Returns whether it did.
template <class T>
__device__ inline bool project_deformation_gradient(const T* F, T threshold, T* Fp) {
    T U[9], s[3], V[9];
    rotation_svd3(F, U, s, V);
    bool changed = false;
    for (int i = 0; i < 3; ++i)
        if (s[i] < threshold) {
            s[i] = threshold;
            changed = true;
        }
    if (!changed) {
        for (int k = 0; k < 9; ++k) Fp[k] = F[k];
        return false;
    }
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            Fp[r + 3 * c] = U[r] * s[0] * V[c] + U[r + 3] * s[1] * V[c + 3] + U[r + 6] * s[2] * V[c + 6];
    return true;
}

// ProjectedStableNeoHookeanMaterial tangent (fenris/src/solid/materials.rs:427-441 with the
// analytic eigensystem of :561-687): dp holds build_stable_neohookean_dp_df(F); every
// eigenpair with a negative eigenvalue is removed, dp -= lambda_i q_i q_i^T. Eigenpairs, with
// F = U diag(s) V^T (rotation SVD), mu_t = mu (1 - 1/(I_C + 1)), k the index not in {i, j}:
//   twist  mu_t + lambda (J - alpha) s_k,  q = vec(u_i v_j^T - u_j v_i^T) / sqrt 2
//   flip   mu_t - lambda (J - alpha) s_k,  q = vec(u_i v_j^T + u_j v_i^T) / sqrt 2
//   scale  eigenpairs (w, e) of the 3x3 matrix A of the scaling directions, q = vec(U diag(e) V^T)
template <class T>
__device__ inline void snh_dpdf_project_psd(T mu, T lambda, const T* F, T* dp) {
    T U[9], s[3], V[9];
    rotation_svd3(F, U, s, V);
    const T alpha = T(1) + mu / lambda * T(0.75);
    T I_C = T(0);
    for (int i = 0; i < 9; ++i) I_C += F[i] * F[i];
    const T J = det3(F);
    const T mu_t = mu * (T(1) - T(1) / (I_C + T(1)));
    const T sv_scale = lambda * (J - alpha);
    auto remove = [&](T ev, const T* q) {
        for (int c = 0; c < 9; ++c) {
            const T f = ev * q[c];
            for (int r = 0; r < 9; ++r) dp[r + 9 * c] -= f * q[r];
        }
    };
    const int PI[3] = {1, 0, 0}, PJ[3] = {2, 2, 1};
    const T h = T(0.70710678118654752440);
    for (int k = 0; k < 3; ++k) {
        const int i = PI[k], j = PJ[k];
        for (int sign = 0; sign < 2; ++sign) {
            const T ev = sign == 0 ? mu_t + s[k] * sv_scale : mu_t - s[k] * sv_scale;
            if (!(ev < T(0))) continue;
            T q[9];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) {
                    const T a = U[r + 3 * i] * V[c + 3 * j], b = U[r + 3 * j] * V[c + 3 * i];
                    q[r + 3 * c] = h * (sign == 0 ? a - b : a + b);
                }
            remove(ev, q);
        }
    }
    const T beta = I_C + T(1);
    const T m2 = T(2) * mu / (beta * beta);
    const T gamma = lambda * (T(2) * J - alpha);
    T A[9];
    A[0] = mu_t + m2 * s[0] * s[0] + lambda * s[1] * s[1] * s[2] * s[2];
    A[4] = mu_t + m2 * s[1] * s[1] + lambda * s[0] * s[0] * s[2] * s[2];
    A[8] = mu_t + m2 * s[2] * s[2] + lambda * s[0] * s[0] * s[1] * s[1];
    A[3] = A[1] = s[2] * gamma + m2 * s[0] * s[1];
    A[6] = A[2] = s[1] * gamma + m2 * s[0] * s[2];
    A[7] = A[5] = s[0] * gamma + m2 * s[1] * s[2];
    T w[3], E[9];
    sym3_eigen(A, w, E);
    for (int m = 0; m < 3; ++m) {
        if (!(w[m] < T(0))) continue;
        T q[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r)
                q[r + 3 * c] = U[r] * E[3 * m] * V[c] + U[r + 3] * E[3 * m + 1] * V[c + 3] +
                               U[r + 6] * E[3 * m + 2] * V[c + 6];
        remove(w[m], q);
    }
}

// The stress of the selected model (kind as in MAT_*).
template <class T>
__device__ inline bool material_stress(int kind, T threshold, T mu, T lambda, const T* F, T* P) {
    if (kind == MAT_INVERTIBLE_SNH) {
        T Fp[9];
        project_deformation_gradient(F, threshold, Fp);
        return snh_stress(mu, lambda, Fp, P);
    }
    return snh_stress(mu, lambda, F, P);
}

// The tangent dP/dF of the selected model.
template <class T>
__device__ inline bool material_dpdf(int kind, T threshold, T mu, T lambda, const T* F, T* dp) {
    if (kind == MAT_INVERTIBLE_SNH) {
        T Fp[9];
        project_deformation_gradient(F, threshold, Fp);
        return snh_dpdf(mu, lambda, Fp, dp);
    }
    if (!snh_dpdf(mu, lambda, F, dp)) return false;
    if (kind == MAT_PROJECTED_SNH) snh_dpdf_project_psd(mu, lambda, F, dp);
    return true;
}

}  // namespace dev
}  // namespace hoefem
