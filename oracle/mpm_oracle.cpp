// mpm_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY; see mpm_oracle.h).
//
// C++17 + OpenMP restatement of the reference MPM snow substep,
// instantiated for float (the reference build, MPM_Example.cpp:1081-1082)
// and double (COMPILE_WITH_DOUBLE_SUPPORT, MPM_Example.cpp:1083-1086).
// Each function cites the reference file:line it follows.  Nova pieces the
// reference calls (SPGrid storage, Grid geometry, the 3x3 SVD, the 3x3
// solve, the Krylov CG) are restated from mathematics:
//   * SPGrid     -> a dense node array over the bounding-box cells plus one
//                   ghost layer, zeroed at every Initialize_SPGrid.
//   * Grid       -> cell-centred nodes, Closest_Cell = floor(X/dx),
//                   Center(c) = (c+1/2) dx, domain [0,1]^d.
//   * SVD        -> cyclic Jacobi on F^T F (Golub & Van Loan, Matrix
//                   Computations, Alg. 8.5.2), U by Gram-Schmidt of F V,
//                   singular values descending, U,V in SO(d), the sign of
//                   det F carried by the smallest singular value.
//   * 3x3 solve  -> adjugate over determinant (Cramer's rule).
//   * CG         -> Hestenes-Stiefel CG in the M inner product with a
//                   restart at iteration 0, Project after every Multiply,
//                   stop when Convergence_Norm(r) <= tolerance.
#include "mpm_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Small fixed-size vector / matrix types.  Matrices are column major like
// Nova's Matrix<T,d> (element (i,j) at x[j*d+i]).
// ---------------------------------------------------------------------------
template <class T, int d>
struct Vec {
    T v[d];
    Vec() { for (int i = 0; i < d; ++i) v[i] = T(0); }
    T& operator()(int i) { return v[i]; }
    const T& operator()(int i) const { return v[i]; }
    T Dot(const Vec& o) const { T s = T(0); for (int i = 0; i < d; ++i) s += v[i] * o.v[i]; return s; }
    T Norm_Squared() const { return Dot(*this); }
    T Norm() const { return std::sqrt(Norm_Squared()); }
};

template <class T, int d>
struct Mat {
    T x[d * d];
    Mat() { for (int i = 0; i < d * d; ++i) x[i] = T(0); }
    static Mat Identity() { Mat m; for (int i = 0; i < d; ++i) m(i, i) = T(1); return m; }
    T& operator()(int i, int j) { return x[j * d + i]; }
    const T& operator()(int i, int j) const { return x[j * d + i]; }
    Mat operator*(const Mat& b) const {
        Mat r;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                T s = T(0);
                for (int k = 0; k < d; ++k) s += (*this)(i, k) * b(k, j);
                r(i, j) = s;
            }
        return r;
    }
    Vec<T, d> operator*(const Vec<T, d>& b) const {
        Vec<T, d> r;
        for (int i = 0; i < d; ++i) {
            T s = T(0);
            for (int k = 0; k < d; ++k) s += (*this)(i, k) * b(k);
            r(i) = s;
        }
        return r;
    }
    Mat operator*(T a) const { Mat r; for (int i = 0; i < d * d; ++i) r.x[i] = x[i] * a; return r; }
    Mat operator+(const Mat& b) const { Mat r; for (int i = 0; i < d * d; ++i) r.x[i] = x[i] + b.x[i]; return r; }
    Mat operator-(const Mat& b) const { Mat r; for (int i = 0; i < d * d; ++i) r.x[i] = x[i] - b.x[i]; return r; }
    Mat& operator+=(const Mat& b) { for (int i = 0; i < d * d; ++i) x[i] += b.x[i]; return *this; }
    Mat& operator*=(T a) { for (int i = 0; i < d * d; ++i) x[i] *= a; return *this; }
    Mat Transposed() const { Mat r; for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) r(i, j) = (*this)(j, i); return r; }
    // A^T B
    Mat Transpose_Times(const Mat& b) const { return Transposed() * b; }
    Vec<T, d> Transpose_Times(const Vec<T, d>& b) const { return Transposed() * b; }
    // A B^T
    Mat Times_Transpose(const Mat& b) const { return (*this) * b.Transposed(); }
    T Trace() const { T s = T(0); for (int i = 0; i < d; ++i) s += (*this)(i, i); return s; }
    static Mat Outer_Product(const Vec<T, d>& a, const Vec<T, d>& b) {
        Mat r;
        for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) r(i, j) = a(i) * b(j);
        return r;
    }
};

template <class T>
T Determinant(const Mat<T, 2>& m) { return m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0); }
template <class T>
T Determinant(const Mat<T, 3>& m) {
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) - m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
// Cofactor matrix C = det(F) F^{-T}: columns are (f2 x f3, f3 x f1, f1 x f2)
// for F = [f1 f2 f3] (3D); [[f22,-f21],[-f12,f11]] (2D).
template <class T>
Mat<T, 2> Cofactor(const Mat<T, 2>& F) {
    Mat<T, 2> C;
    C(0, 0) = F(1, 1); C(0, 1) = -F(1, 0);
    C(1, 0) = -F(0, 1); C(1, 1) = F(0, 0);
    return C;
}
template <class T>
void Cross(const T* a, const T* b, T* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T>
Mat<T, 3> Cofactor(const Mat<T, 3>& F) {
    Mat<T, 3> C;
    const T* f1 = &F.x[0]; const T* f2 = &F.x[3]; const T* f3 = &F.x[6];
    Cross(f2, f3, &C.x[0]);
    Cross(f3, f1, &C.x[3]);
    Cross(f1, f2, &C.x[6]);
    return C;
}
// Directional derivative of the cofactor matrix (MPM_Constitutive_Model.h:28-46).
template <class T>
Mat<T, 2> Cofactor_Derivative(const Mat<T, 2>&, const Mat<T, 2>& dF) { return Cofactor(dF); }
template <class T>
Mat<T, 3> Cofactor_Derivative(const Mat<T, 3>& F, const Mat<T, 3>& dF) {
    Mat<T, 3> C;
    T a[3], b[3];
    const T* f[3] = {&F.x[0], &F.x[3], &F.x[6]};
    const T* g[3] = {&dF.x[0], &dF.x[3], &dF.x[6]};
    for (int c = 0; c < 3; ++c) {
        const int p = (c + 1) % 3, q = (c + 2) % 3;
        Cross(g[p], f[q], a);
        Cross(f[p], g[q], b);
        for (int i = 0; i < 3; ++i) C.x[3 * c + i] = a[i] + b[i];
    }
    return C;
}
// 3x3 solve by the adjugate (Nova Matrix::Solve_Linear_System stand-in).
template <class T>
Vec<T, 3> Solve_Linear_System(const Mat<T, 3>& K, const Vec<T, 3>& L) {
    const Mat<T, 3> C = Cofactor(K);  // adj(K) = C^T
    const T det = Determinant(K);
    Vec<T, 3> r = C.Transpose_Times(L);
    for (int i = 0; i < 3; ++i) r(i) /= det;
    return r;
}

// ---------------------------------------------------------------------------
// SVD (Nova Matrix::Fast_Singular_Value_Decomposition stand-in).
// Computed in double (ORACLE_SVD_FP64) or in T.
// ---------------------------------------------------------------------------
template <class S>
void jacobi_eigen3(S A[3][3], S V[3][3]) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? S(1) : S(0);
    for (int sweep = 0; sweep < 50; ++sweep) {
        const S off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const S diag = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (!(off > std::numeric_limits<S>::min()) || off <= diag * std::numeric_limits<S>::epsilon() * std::numeric_limits<S>::epsilon() * S(1e-4))
            break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == S(0)) continue;
                // Golub & Van Loan Alg. 8.4.1 (sym.schur2)
                const S tau = (A[q][q] - A[p][p]) / (S(2) * A[p][q]);
                const S t = (tau >= S(0) ? S(1) : S(-1)) / (std::fabs(tau) + std::sqrt(S(1) + tau * tau));
                const S c = S(1) / std::sqrt(S(1) + t * t), s = t * c;
                for (int k = 0; k < 3; ++k) {  // A <- A J
                    const S akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- J^T A
                    const S apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {  // V <- V J
                    const S vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
}

// Row-major F (3x3) -> U, sigma, V (row-major), F = U diag(sigma) V^T.
template <class S>
void svd3(const S F[9], S U[9], S sig[3], S Vout[9]) {
    S A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            S s = S(0);
            for (int k = 0; k < 3; ++k) s += F[3 * k + i] * F[3 * k + j];
            A[i][j] = s;
        }
    jacobi_eigen3(A, V);
    int order[3] = {0, 1, 2};
    const S lam[3] = {A[0][0], A[1][1], A[2][2]};
    std::sort(order, order + 3, [&](int a, int b) { return lam[a] > lam[b]; });
    S v[3][3];  // v[k] = k-th right singular vector
    for (int k = 0; k < 3; ++k) for (int i = 0; i < 3; ++i) v[k][i] = V[i][order[k]];
    S c[3];
    Cross(v[0], v[1], c);
    if (c[0] * v[2][0] + c[1] * v[2][1] + c[2] * v[2][2] < S(0)) for (int i = 0; i < 3; ++i) v[2][i] = -v[2][i];
    S Fv[3][3];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) Fv[k][i] = F[3 * i + 0] * v[k][0] + F[3 * i + 1] * v[k][1] + F[3 * i + 2] * v[k][2];
    S u[3][3];
    S n0 = std::sqrt(Fv[0][0] * Fv[0][0] + Fv[0][1] * Fv[0][1] + Fv[0][2] * Fv[0][2]);
    sig[0] = n0;
    if (n0 > S(0)) for (int i = 0; i < 3; ++i) u[0][i] = Fv[0][i] / n0;
    else { u[0][0] = S(1); u[0][1] = S(0); u[0][2] = S(0); }
    S w[3];
    const S p = u[0][0] * Fv[1][0] + u[0][1] * Fv[1][1] + u[0][2] * Fv[1][2];
    for (int i = 0; i < 3; ++i) w[i] = Fv[1][i] - p * u[0][i];
    S n1 = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (n1 > std::numeric_limits<S>::epsilon() * (n0 > S(0) ? n0 : S(1))) {
        for (int i = 0; i < 3; ++i) u[1][i] = w[i] / n1;
    } else {
        // any unit vector orthogonal to u0
        const int m = (std::fabs(u[0][0]) < std::fabs(u[0][1])) ? ((std::fabs(u[0][0]) < std::fabs(u[0][2])) ? 0 : 2)
                                                                 : ((std::fabs(u[0][1]) < std::fabs(u[0][2])) ? 1 : 2);
        S e[3] = {S(0), S(0), S(0)};
        e[m] = S(1);
        S t[3];
        Cross(u[0], e, t);
        const S nt = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int i = 0; i < 3; ++i) u[1][i] = t[i] / nt;
        n1 = u[1][0] * Fv[1][0] + u[1][1] * Fv[1][1] + u[1][2] * Fv[1][2];
    }
    sig[1] = n1;
    Cross(u[0], u[1], u[2]);
    sig[2] = u[2][0] * Fv[2][0] + u[2][1] * Fv[2][1] + u[2][2] * Fv[2][2];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            U[3 * i + k] = u[k][i];
            Vout[3 * i + k] = v[k][i];
        }
}

template <class S>
void svd2(const S F[4], S U[4], S sig[2], S Vout[4]) {
    // Eigenvectors of the 2x2 normal matrix by one exact Jacobi rotation.
    const S a = F[0] * F[0] + F[2] * F[2], b = F[0] * F[1] + F[2] * F[3], c = F[1] * F[1] + F[3] * F[3];
    S cs = S(1), sn = S(0);
    if (b != S(0)) {
        const S tau = (c - a) / (S(2) * b);
        const S t = (tau >= S(0) ? S(1) : S(-1)) / (std::fabs(tau) + std::sqrt(S(1) + tau * tau));
        cs = S(1) / std::sqrt(S(1) + t * t);
        sn = t * cs;
    }
    const S l0 = cs * cs * a - S(2) * cs * sn * b + sn * sn * c;
    const S l1 = sn * sn * a + S(2) * cs * sn * b + cs * cs * c;
    S v0[2] = {cs, -sn}, v1[2] = {sn, cs};
    if (l1 > l0) { S t0[2] = {v0[0], v0[1]}; v0[0] = v1[0]; v0[1] = v1[1]; v1[0] = -t0[0]; v1[1] = -t0[1]; }
    // det [v0 v1] = v0x v1y - v0y v1x must be +1
    if (v0[0] * v1[1] - v0[1] * v1[0] < S(0)) { v1[0] = -v1[0]; v1[1] = -v1[1]; }
    const S f0[2] = {F[0] * v0[0] + F[1] * v0[1], F[2] * v0[0] + F[3] * v0[1]};
    const S f1[2] = {F[0] * v1[0] + F[1] * v1[1], F[2] * v1[0] + F[3] * v1[1]};
    const S n0 = std::sqrt(f0[0] * f0[0] + f0[1] * f0[1]);
    S u0[2] = {S(1), S(0)};
    if (n0 > S(0)) { u0[0] = f0[0] / n0; u0[1] = f0[1] / n0; }
    const S u1[2] = {-u0[1], u0[0]};
    sig[0] = n0;
    sig[1] = u1[0] * f1[0] + u1[1] * f1[1];
    U[0] = u0[0]; U[1] = u1[0]; U[2] = u0[1]; U[3] = u1[1];
    Vout[0] = v0[0]; Vout[1] = v1[0]; Vout[2] = v0[1]; Vout[3] = v1[1];
}

template <class S, int d>
void svd_rowmajor(const S* F, S* U, S* sig, S* V) {
    if (d == 3) svd3<S>(F, U, sig, V);
    else svd2<S>(F, U, sig, V);
}

// ---------------------------------------------------------------------------
// Constitutive model (MPM_Constitutive_Model.{h,cpp})
// ---------------------------------------------------------------------------
template <class T, int d>
struct Model {
    T lambda = 0, mu = 0, lambda0 = 0, mu0 = 0;
    Mat<T, d> Fe = Mat<T, d>::Identity(), Ue, Ve;
    Vec<T, d> Se;
    T Je = T(1);
    Mat<T, d> Fp = Mat<T, d>::Identity();
    T Jp = T(1);  // used with ORACLE_JP_SCALAR
    Mat<T, d> Re, He;
    bool plastic = false;
    T stretching_yield = FLT_MAX, compression_yield = -FLT_MAX, hardening_factor = 0;
    T hardening_max = FLT_MAX;
};

template <class T, int d>
void SVD(const Mat<T, d>& F, Mat<T, d>& U, Vec<T, d>& S, Mat<T, d>& V, bool fp64) {
    if (fp64) {
        double Fr[d * d], Ur[d * d], Vr[d * d], s[d];
        for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) Fr[i * d + j] = (double)F(i, j);
        svd_rowmajor<double, d>(Fr, Ur, s, Vr);
        for (int i = 0; i < d; ++i) {
            S(i) = (T)s[i];
            for (int j = 0; j < d; ++j) { U(i, j) = (T)Ur[i * d + j]; V(i, j) = (T)Vr[i * d + j]; }
        }
    } else {
        T Fr[d * d], Ur[d * d], Vr[d * d], s[d];
        for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) Fr[i * d + j] = F(i, j);
        svd_rowmajor<T, d>(Fr, Ur, s, Vr);
        for (int i = 0; i < d; ++i) {
            S(i) = s[i];
            for (int j = 0; j < d; ++j) { U(i, j) = Ur[i * d + j]; V(i, j) = Vr[i * d + j]; }
        }
    }
}

template <class T, int d>
Mat<T, d> Diag(const Vec<T, d>& s) { Mat<T, d> m; for (int i = 0; i < d; ++i) m(i, i) = s(i); return m; }

// Precompute (MPM_Constitutive_Model.cpp:32-51)
template <class T, int d>
void Precompute(Model<T, d>& m, bool svd_fp64, bool jp_scalar) {
    m.Ue = Mat<T, d>();
    m.Ve = Mat<T, d>();
    const Mat<T, d> Fe_old = m.Fe;
    SVD(m.Fe, m.Ue, m.Se, m.Ve, svd_fp64);
    if (m.plastic) {
        T det_old = T(1);
        for (int i = 0; i < d; ++i) det_old *= m.Se(i);
        for (int i = 0; i < d; ++i) if (m.Se(i) > m.stretching_yield) m.Se(i) = m.stretching_yield;
        for (int i = 0; i < d; ++i) if (m.Se(i) < m.compression_yield) m.Se(i) = m.compression_yield;
        T detFp;
        if (jp_scalar) {
            T det_new = T(1);
            for (int i = 0; i < d; ++i) det_new *= m.Se(i);
            m.Jp = m.Jp * (det_old / det_new);
            detFp = m.Jp;
        } else {
            Vec<T, d> inv;
            for (int i = 0; i < d; ++i) inv(i) = T(1) / m.Se(i);
            m.Fp = m.Ve * Diag(inv) * m.Ue.Transposed() * Fe_old * m.Fp;  // :40
            detFp = Determinant(m.Fp);
        }
        m.Fe = m.Ue * Diag(m.Se) * m.Ve.Transposed();                          // :41
        const T power = std::min(m.hardening_factor * (T(1) - detFp), (T)m.hardening_max);  // :42
        m.lambda = std::exp(power) * m.lambda0;
        m.mu = std::exp(power) * m.mu0;
    } else {
        m.lambda = m.lambda0;
        m.mu = m.mu0;
    }
    T Je = T(1);
    for (int i = 0; i < d; ++i) Je *= m.Se(i);  // Se.Determinant() (:48)
    m.Je = Je;
    m.Re = m.Ue * m.Ve.Transposed();             // :49
    m.He = Cofactor(m.Fe);                        // :50
}

// Psi (MPM_Constitutive_Model.cpp:55-59)
template <class T, int d>
T Psi(const Model<T, d>& m) {
    T s = T(0);
    for (int i = 0; i < d; ++i) s += (m.Se(i) - T(1)) * (m.Se(i) - T(1));
    return m.mu * s + m.lambda / T(2) * (m.Je - T(1)) * (m.Je - T(1));
}

// P (MPM_Constitutive_Model.cpp:63-67): 2 mu (Fe - Re) + lambda (Je-1) He
template <class T, int d>
Mat<T, d> P(const Model<T, d>& m) { return (m.Fe - m.Re) * (T(2) * m.mu) + m.He * (m.lambda * (m.Je - T(1))); }

// dR (MPM_Constitutive_Model.h:47-69): R^T dR = W skew with W S + S W = M - M^T,
// M = R^T dF, S = R^T F.
template <class T>
Mat<T, 2> dR(const Mat<T, 2>& dF, const Mat<T, 2>& R, const Mat<T, 2>& F) {
    const Mat<T, 2> M = R.Transpose_Times(dF), S = R.Transpose_Times(F);
    const T a = (M(1, 0) - M(0, 1)) / (S(1, 1) + S(0, 0));
    Mat<T, 2> W;
    W(1, 0) = a;
    W(0, 1) = -a;
    return R * W;
}
template <class T>
Mat<T, 3> dR(const Mat<T, 3>& dF, const Mat<T, 3>& R, const Mat<T, 3>& F) {
    const Mat<T, 3> M = R.Transpose_Times(dF), S = R.Transpose_Times(F);
    // W = [[0,-a,-b],[a,0,-c],[b,c,0]];  K (a,b,c) = L
    Mat<T, 3> K;
    K(0, 0) = S(0, 0) + S(1, 1); K(0, 1) = S(2, 1);          K(0, 2) = -S(2, 0);
    K(1, 0) = S(2, 1);          K(1, 1) = S(0, 0) + S(2, 2); K(1, 2) = S(1, 0);
    K(2, 0) = -S(2, 0);         K(2, 1) = S(1, 0);          K(2, 2) = S(1, 1) + S(2, 2);
    Vec<T, 3> L;
    L(0) = M(1, 0) - M(0, 1);
    L(1) = M(2, 0) - M(0, 2);
    L(2) = M(2, 1) - M(1, 2);
    const Vec<T, 3> w = Solve_Linear_System(K, L);
    Mat<T, 3> W;
    W(1, 0) = w(0); W(0, 1) = -w(0);
    W(2, 0) = w(1); W(0, 2) = -w(1);
    W(2, 1) = w(2); W(1, 2) = -w(2);
    return R * W;
}

// Times_dP_dF (MPM_Constitutive_Model.cpp:71-76)
template <class T, int d>
Mat<T, d> Times_dP_dF(const Model<T, d>& m, const Mat<T, d>& dF) {
    const Mat<T, d> dH = Cofactor_Derivative(m.Fe, dF);
    const T HdF = m.He.Times_Transpose(dF).Trace();
    return (dF - dR(dF, m.Re, m.Fe)) * (T(2) * m.mu) + m.He * (m.lambda * HdF) + dH * (m.lambda * (m.Je - T(1)));
}

// ---------------------------------------------------------------------------
// Particle (MPM_Particle.h:14-101)
// ---------------------------------------------------------------------------
template <class T, int d>
struct Particle {
    bool valid = true;
    Vec<T, d> X, V;
    int closest_cell[d];
    T mass = 0, volume = 0;
    T weights[3][d], dweights[3][d];  // weights(k, axis)
    Model<T, d> model;
    Mat<T, d> scp;
    // EOS fluid particle (MPM_Particle.h:33-37, defaults :51-56)
    bool eos = false;
    T density = 1, bulk_modulus = T(1e7), gamma = 7;
    Mat<T, d> eos_scp;

    // Compute_Weights (MPM_Particle.h:90-100)
    void Compute_Weights(T x, const T one_over_dx, const int k) {
        weights[0][k] = std::fabs(T(0.5) * x * x - T(1.5) * x + T(1.125));
        dweights[0][k] = (x - T(1.5)) * one_over_dx;
        x -= T(1);
        weights[1][k] = std::fabs(-x * x + T(0.75));
        dweights[1][k] = T(-2) * x * one_over_dx;
        x -= T(1);
        weights[2][k] = std::fabs(T(0.5) * x * x + T(1.5) * x + T(1.125));
        dweights[2][k] = (x + T(1.5)) * one_over_dx;
    }
    // Update_Weights (MPM_Particle.h:79-88) with Closest_Cell = floor(X/dx) and
    // Center(c) = (c+1/2) dx on the [0,1]^d domain.
    void Update_Weights(int counts) {
        const T dx = T(1) / T(counts), one_over_dx = T(counts);
        for (int i = 0; i < d; ++i) closest_cell[i] = (int)std::floor(X(i) * one_over_dx);
        for (int i = 0; i < d; ++i) {
            const T center = (T(closest_cell[i]) + T(0.5)) * dx;
            const T X_eval = X(i) - (center - dx);
            Compute_Weights(X_eval * one_over_dx, one_over_dx, i);
        }
    }
    // Influence_Iterator::Weight / Weight_Gradient (Tools/Influence_Iterator.h:85-99);
    // idx in {-1,0,1}^d.
    T Weight(const int* idx) const {
        T w = T(1);
        for (int i = 0; i < d; ++i) w *= weights[idx[i] + 1][i];
        return w;
    }
    Vec<T, d> Weight_Gradient(const int* idx) const {
        Vec<T, d> g;
        for (int i = 0; i < d; ++i) g(i) = T(1);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) g(i) *= (i == j) ? dweights[idx[j] + 1][j] : weights[idx[j] + 1][j];
        return g;
    }
};

// Visit the 3^d stencil in Influence_Iterator order (last axis fastest,
// Tools/Influence_Iterator.h:66-83), x cropped to [xlo, xhi] relative offsets
// (Tools/Cropped_Influence_Iterator.h:31-42).
template <int d, class F>
inline void for_stencil(int xlo, int xhi, F&& f) {
    int idx[3] = {0, 0, 0};
    for (int a = std::max(-1, xlo); a <= std::min(1, xhi); ++a) {
        idx[0] = a;
        if (d == 2) {
            for (int b = -1; b <= 1; ++b) { idx[1] = b; f(idx); }
        } else {
            for (int b = -1; b <= 1; ++b)
                for (int c = -1; c <= 1; ++c) { idx[1] = b; idx[2] = c; f(idx); }
        }
    }
}

// Grid node record (MPM_Data.h:12-46): flags + 25 channels.
template <class T>
struct Node {
    uint32_t flags;
    T ch[25];
};
enum { CH_M = 0, CH_V = 1, CH_VS = 4, CH_F = 7, CH_RHS = 10, CH_Q = 13, CH_S = 16, CH_R = 19, CH_Z = 22 };
const uint32_t FLAG_INTERIOR = MPM_FLAG_INTERIOR;
const uint32_t FLAG_DIRICHLET = 2u;

struct Interval {
    int min_corner, max_corner;
    bool Intersection(const Interval& o) const { return min_corner <= o.max_corner && o.min_corner <= max_corner; }
};

struct OracleBase {
    virtual ~OracleBase() {}
    virtual void set_particles(int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                               const float* mass, const float* volume, const int32_t* material) = 0;
    virtual void add_particles(int64_t n, const float* x, const float* v, const float* mass,
                               const int32_t* material) = 0;
    virtual int64_t num_particles() const = 0;
    virtual int64_t num_simulated() const = 0;
    virtual void get_particles(double* x, double* v, double* F, double* Jp, double* mass, double* volume,
                               uint8_t* valid) const = 0;
    virtual void initialize() = 0;
    virtual void advance_step(double dt) = 0;
    virtual void run_stage(int stage, double dt) = 0;
    virtual void populate() = 0;
    virtual double max_velocity() const = 0;
    virtual double advance_to_target_time(double time, double target, int* substeps) = 0;
    virtual void get_grid(int channel, double* out) const = 0;
    virtual void get_flags(uint32_t* out) const = 0;
    virtual void apply_operator(double dt, const double* x, double* q) = 0;
    virtual void set_grid(int channel, const double* in) = 0;
    virtual double inner_product(const double* a, const double* b) const = 0;
    virtual void cg_stats(int* it, int* mult, int* conv, double* res) const = 0;
    virtual void bbox_cells(int* lo, int* hi) const = 0;
};

template <class T, int d>
struct Oracle : OracleBase {
    using TV = Vec<T, d>;
    using TM = Mat<T, d>;
    mpm_scene_desc desc;
    int threads;
    bool dot_fp64, svd_fp64, jp_scalar;
    int counts;
    T dx, one_over_dx;
    std::vector<Particle<T, d>> particles;
    std::vector<int> simulated_particles, waiting_particles, invalid_particles;
    // grid: dense over node indices [lo-1, hi+1]^d
    int glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0}, gsize[3] = {1, 1, 1};
    int bbox_lo[3] = {0, 0, 0}, bbox_hi[3] = {0, 0, 0};
    std::vector<Node<T>> grid;
    std::vector<Interval> x_intervals;
    std::vector<std::vector<int>> particle_bins;  // (tid_process, tid_collect) -> tid_process*threads+tid_collect
    int cg_iterations = 0, cg_multiplies = 0, cg_converged = 0;
    double cg_residual = 0;

    Oracle(const mpm_scene_desc& dsc, int nthreads, int options)
        : desc(dsc), threads(std::max(1, nthreads)) {
        dot_fp64 = options & ORACLE_DOT_FP64;
        svd_fp64 = options & ORACLE_SVD_FP64;
        jp_scalar = options & ORACLE_JP_SCALAR;
        counts = desc.counts;
        dx = T(1) / T(counts);
        one_over_dx = T(counts);
        particle_bins.resize((size_t)threads * threads);
    }

    // ---- particles --------------------------------------------------------
    // A Standard_Tests scene sets each particle's constitutive model and EOS
    // fields (mpm_2d/Standard_Tests/Standard_Tests.h:55-283); here every
    // particle names an entry of the scene's material table.  An EOS
    // particle's Jp input is its density (its material's density if Jp is NULL).
    void set_particles(int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                       const float* mass, const float* volume, const int32_t* material) override {
        particles.assign((size_t)n, Particle<T, d>());
        for (int64_t i = 0; i < n; ++i) {
            Particle<T, d>& p = particles[(size_t)i];
            for (int a = 0; a < d; ++a) { p.X(a) = (T)x[i * d + a]; p.V(a) = (T)v[i * d + a]; }
            p.mass = (T)mass[i];
            p.volume = volume ? (T)volume[i] : T(0);
            const int id = material ? material[i] : 0;
            const mpm_material& mat = id == 0 ? desc.material : desc.extra_materials[id - 1];
            Model<T, d>& m = p.model;
            m.mu0 = (T)mat.mu0;
            m.lambda0 = (T)mat.lambda0;
            m.plastic = mat.plastic != 0;
            m.stretching_yield = (T)mat.stretching_yield;
            m.compression_yield = (T)mat.compression_yield;
            m.hardening_factor = (T)mat.hardening_factor;
            m.hardening_max = (T)mat.hardening_max;
            if (F) for (int a = 0; a < d; ++a) for (int b = 0; b < d; ++b) m.Fe(a, b) = (T)F[(i * d + a) * d + b];
            p.eos = mat.eos != 0;
            p.bulk_modulus = (T)mat.bulk_modulus;
            p.gamma = (T)mat.gamma;
            p.density = p.eos ? (Jp ? (T)Jp[i] : (T)mat.density) : T(1);
            const T jp = Jp && !p.eos ? (T)Jp[i] : T(1);
            m.Jp = jp;
            m.Fp = TM::Identity();
            m.Fp(0, 0) = jp;  // only det(Fp) is ever read (MPM_Constitutive_Model.cpp:42)
        }
        simulated_particles.clear();
        waiting_particles.clear();
        invalid_particles.clear();
        for (int64_t i = 0; i < n; ++i) simulated_particles.push_back((int)i);
        if (!volume) waiting_particles = simulated_particles;
    }
    // Add_Fluid_Source with Allocate_Particle (MPM_Example.cpp:925-1027): new
    // particles join the simulated and waiting lists with volume 0 (their
    // volume comes from the grid in the next Process_Waiting_Particles).  The
    // reference recycles the storage of removed particles; here they are
    // appended, which changes storage order only.
    void add_particles(int64_t n, const float* x, const float* v, const float* mass,
                       const int32_t* material) override {
        for (int64_t i = 0; i < n; ++i) {
            Particle<T, d> p;
            for (int a = 0; a < d; ++a) { p.X(a) = (T)x[i * d + a]; p.V(a) = (T)v[i * d + a]; }
            p.mass = (T)mass[i];
            p.volume = T(0);
            const int id = material ? material[i] : 0;
            const mpm_material& mat = id == 0 ? desc.material : desc.extra_materials[id - 1];
            Model<T, d>& m = p.model;
            m.mu0 = (T)mat.mu0;
            m.lambda0 = (T)mat.lambda0;
            m.plastic = mat.plastic != 0;
            m.stretching_yield = (T)mat.stretching_yield;
            m.compression_yield = (T)mat.compression_yield;
            m.hardening_factor = (T)mat.hardening_factor;
            m.hardening_max = (T)mat.hardening_max;
            p.eos = mat.eos != 0;
            p.bulk_modulus = (T)mat.bulk_modulus;
            p.gamma = (T)mat.gamma;
            p.density = p.eos ? (T)mat.density : T(1);
            m.Jp = T(1);
            m.Fp = TM::Identity();
            const int pid = (int)particles.size();
            particles.push_back(p);
            simulated_particles.push_back(pid);
            waiting_particles.push_back(pid);
        }
    }
    int64_t num_particles() const override { return (int64_t)particles.size(); }
    int64_t num_simulated() const override { return (int64_t)simulated_particles.size(); }
    void get_particles(double* x, double* v, double* F, double* Jp, double* mass, double* volume,
                       uint8_t* valid) const override {
        for (size_t i = 0; i < particles.size(); ++i) {
            const Particle<T, d>& p = particles[i];
            for (int a = 0; a < d; ++a) {
                if (x) x[i * d + a] = (double)p.X(a);
                if (v) v[i * d + a] = (double)p.V(a);
            }
            if (F) for (int a = 0; a < d; ++a) for (int b = 0; b < d; ++b) F[(i * d + a) * d + b] = (double)p.model.Fe(a, b);
            if (Jp) Jp[i] = p.eos ? (double)p.density : jp_scalar ? (double)p.model.Jp : (double)Determinant(p.model.Fp);
            if (mass) mass[i] = (double)p.mass;
            if (volume) volume[i] = (double)p.volume;
            if (valid) valid[i] = p.valid ? 1 : 0;
        }
    }

    // ---- grid -------------------------------------------------------------
    int64_t gidx(const int* c) const {
        int64_t r = 0;
        for (int a = 0; a < d; ++a) {
            const int o = c[a] - glo[a];
            if (o < 0 || o >= gsize[a]) return -1;
            r = r * gsize[a] + o;
        }
        return r;
    }
    Node<T>& node(const int* c) {
        const int64_t i = gidx(c);
        if (i < 0) { std::fprintf(stderr, "oracle: node outside grid\n"); std::abort(); }
        return grid[(size_t)i];
    }
    void node_coord(int64_t lin, int* c) const {
        for (int a = d - 1; a >= 0; --a) { c[a] = (int)(lin % gsize[a]) + glo[a]; lin /= gsize[a]; }
    }
    T center(int c) const { return (T(c) + T(0.5)) * dx; }

    // Compute_Bounding_Box (MPM_Example.cpp:236-265)
    void Compute_Bounding_Box(T* bmin, T* bmax) {
        std::vector<TV> mins((size_t)threads), maxs((size_t)threads);
        for (int t = 0; t < threads; ++t) for (int a = 0; a < d; ++a) { mins[t](a) = T(1); maxs[t](a) = T(0); }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            const int tid = omp_get_thread_num();
            const Particle<T, d>& p = particles[(size_t)simulated_particles[(size_t)i]];
            for (int a = 0; a < d; ++a) {
                const T dd = T(2) / T(counts);
                mins[tid](a) = std::min(mins[tid](a), p.X(a) - dd);
                maxs[tid](a) = std::max(maxs[tid](a), p.X(a) + dd);
            }
        }
        for (int a = 0; a < d; ++a) { bmin[a] = mins[0](a); bmax[a] = maxs[0](a); }
        for (int t = 1; t < threads; ++t)
            for (int a = 0; a < d; ++a) { bmin[a] = std::min(bmin[a], mins[t](a)); bmax[a] = std::max(bmax[a], maxs[t](a)); }
        for (int a = 0; a < d; ++a) { bmin[a] = std::max(T(0), bmin[a]); bmax[a] = std::min(T(1), bmax[a]); }
    }
    int Clamp_To_Cell(T X) const {
        int c = (int)std::floor(X * one_over_dx);
        return std::min(std::max(c, 0), counts - 1);
    }
    // Initialize_SPGrid (MPM_Example.cpp:325-331) + Rasterize_Voxels (:291-311): the
    // hierarchy is recreated (zeroed) and every bounding-box cell is activated.
    void Initialize_SPGrid_And_Voxels() {
        T bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
        Compute_Bounding_Box(bmin, bmax);
        for (int a = 0; a < d; ++a) {
            bbox_lo[a] = Clamp_To_Cell(bmin[a]);
            bbox_hi[a] = Clamp_To_Cell(bmax[a]);
            if (simulated_particles.empty()) { bbox_lo[a] = 0; bbox_hi[a] = -1; }
            glo[a] = bbox_lo[a] - 1;
            ghi[a] = bbox_hi[a] + 1;
            gsize[a] = std::max(0, ghi[a] - glo[a] + 1);
        }
        size_t total = 1;
        for (int a = 0; a < d; ++a) total *= (size_t)gsize[a];
        grid.assign(total, Node<T>());
        std::memset(grid.data(), 0, total * sizeof(Node<T>));
        // activate bbox cells as Dirichlet (:300-302)
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long lin = 0; lin < (long)total; ++lin) {
            int c[3];
            node_coord(lin, c);
            bool in = true;
            for (int a = 0; a < d; ++a) in = in && c[a] >= bbox_lo[a] && c[a] <= bbox_hi[a];
            if (in) grid[(size_t)lin].flags |= FLAG_DIRICHLET;
        }
    }

    // Update_Particle_Weights (MPM_Example.cpp:418-427)
    void Update_Particle_Weights() {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i)
            particles[(size_t)simulated_particles[(size_t)i]].Update_Weights(counts);
    }

    // Group_Particles (MPM_Example.cpp:464-485)
    void Group_Particles() {
        const int min_corner = 0, max_corner = counts - 1;
        x_intervals.assign((size_t)threads, Interval{0, 0});
        x_intervals[0].min_corner = min_corner;
        x_intervals[(size_t)threads - 1].max_corner = max_corner;
        for (int i = 0; i < threads - 1; ++i) {
            const int n = min_corner + (max_corner - min_corner + 1) * (i + 1) / threads;
            x_intervals[(size_t)i].max_corner = n - 1;
            x_intervals[(size_t)i + 1].min_corner = n;
        }
        for (auto& b : particle_bins) b.clear();
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            const int id = simulated_particles[(size_t)i];
            const int cx = particles[(size_t)id].closest_cell[0];
            const int tid_collect = omp_get_thread_num();
            const Interval px{cx - 1, cx + 1};
            for (int tp = 0; tp < threads; ++tp)
                if (px.Intersection(x_intervals[(size_t)tp])) particle_bins[(size_t)tp * threads + tid_collect].push_back(id);
        }
    }

    // x crop of a particle's stencil for slab tp; with crop_ghost_x = 0 the
    // outermost slabs extend over the ghost nodes.
    void crop(int tp, int cx, int& xlo, int& xhi) const {
        int smin = x_intervals[(size_t)tp].min_corner, smax = x_intervals[(size_t)tp].max_corner;
        if (!desc.crop_ghost_x) {
            if (tp == 0) smin = -1;
            if (tp == threads - 1) smax = counts;
        }
        xlo = smin - cx;
        xhi = smax - cx;
    }
    template <class F>
    void slab_scatter(F&& f) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int tp = 0; tp < threads; ++tp)
            for (int tc = 0; tc < threads; ++tc) {
                const std::vector<int>& index = particle_bins[(size_t)tp * threads + tc];
                for (size_t i = 0; i < index.size(); ++i) {
                    Particle<T, d>& p = particles[(size_t)index[i]];
                    int xlo, xhi;
                    crop(tp, p.closest_cell[0], xlo, xhi);
                    f(p, xlo, xhi);
                }
            }
    }
    void cell_of(const Particle<T, d>& p, const int* idx, int* c) const {
        for (int a = 0; a < d; ++a) c[a] = p.closest_cell[a] + idx[a];
    }
    template <class F>
    void for_nodes(F&& f) {
        const long total = (long)grid.size();
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long lin = 0; lin < total; ++lin) {
            int c[3] = {0, 0, 0};
            node_coord(lin, c);
            f(grid[(size_t)lin], c);
        }
    }

    // Rasterize (MPM_Example.cpp:520-547)
    void Rasterize() {
        slab_scatter([&](Particle<T, d>& p, int xlo, int xhi) {
            for_stencil<d>(xlo, xhi, [&](const int* idx) {
                const TV V = p.V;
                const T p_mass = p.mass;
                const T weight = p.Weight(idx);
                int c[3];
                cell_of(p, idx, c);
                Node<T>& n = node(c);
                n.ch[CH_M] += weight * p_mass;
                for (int a = 0; a < d; ++a) n.ch[CH_V + a] += weight * (p_mass * V(a));
            });
        });
        // Flag_Setup_Helper (Flag_Setup_Helper.h:25-35)
        for_nodes([&](Node<T>& n, const int*) { if (n.ch[CH_M] > T(0)) n.flags |= FLAG_INTERIOR; });
        // Velocity_Normalization_Helper (Velocity_Normalization_Helper.h:39-52)
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR) {
                const T mass_inverse = T(1) / n.ch[CH_M];
                for (int a = 0; a < d; ++a) n.ch[CH_V + a] *= mass_inverse;
            }
        });
    }

    // Process_Waiting_Particles (MPM_Example.cpp:827-840)
    void Process_Waiting_Particles() {
        T vol = T(1);
        for (int a = 0; a < d; ++a) vol *= dx;
        const T one_over_volume_per_cell = T(1) / vol;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)waiting_particles.size(); ++i) {
            Particle<T, d>& p = particles[(size_t)waiting_particles[(size_t)i]];
            T particle_density = T(0);
            for_stencil<d>(-1, 1, [&](const int* idx) {
                int c[3];
                cell_of(p, idx, c);
                particle_density += p.Weight(idx) * node(c).ch[CH_M];
            });
            particle_density *= one_over_volume_per_cell;
            p.volume = p.mass / particle_density;
        }
        waiting_particles.clear();
    }

    // Update_Constitutive_Model_State (MPM_Example.cpp:563-571)
    void Update_Constitutive_Model_State() {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            Particle<T, d>& p = particles[(size_t)simulated_particles[(size_t)i]];
            if (!p.eos) Precompute(p.model, svd_fp64, jp_scalar);
        }
    }

    // Apply_Explicit_Force (MPM_Example.cpp:757-790) + Explicit_Force_Helper (Explicit_Force_Helper.h:43-56)
    void Apply_Explicit_Force(T dt) {
        TV gravity;
        for (int a = 0; a < d; ++a) gravity(a) = (T)desc.gravity[a];
        slab_scatter([&](Particle<T, d>& p, int xlo, int xhi) {
            const T V0 = p.volume;
            // EOS: f += V0/rho K rho^(gamma-1) grad w (MPM_Example.cpp:774-782)
            T eos_coefficient = T(0);
            TM V0_P_FT;
            if (p.eos) eos_coefficient = V0 / p.density * p.bulk_modulus * std::pow(p.density, p.gamma - T(1));
            else V0_P_FT = P(p.model).Times_Transpose(p.model.Fe) * V0;
            for_stencil<d>(xlo, xhi, [&](const int* idx) {
                int c[3];
                cell_of(p, idx, c);
                Node<T>& n = node(c);
                const TV weight_gradient = p.Weight_Gradient(idx);
                const T w = p.Weight(idx);
                TV body_force;
                for (int a = 0; a < d; ++a) body_force(a) = gravity(a) * p.mass * w;
                if (p.eos) {
                    for (int a = 0; a < d; ++a) n.ch[CH_F + a] += eos_coefficient * weight_gradient(a);
                } else {
                    const TV inner_force = V0_P_FT * weight_gradient;
                    for (int a = 0; a < d; ++a) n.ch[CH_F + a] -= inner_force(a);
                }
                for (int a = 0; a < d; ++a) n.ch[CH_F + a] += body_force(a);
            });
        });
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR)
                for (int a = 0; a < d; ++a) n.ch[CH_VS + a] = n.ch[CH_V + a] + dt / n.ch[CH_M] * n.ch[CH_F + a];
        });
    }

    // Grid_Based_Collision_Helper::Run (Grid_Based_Collision_Helper.h:57-85)
    void Grid_Based_Collision() {
        for_nodes([&](Node<T>& n, const int* c) {
            if (!(n.flags & FLAG_INTERIOR)) return;
            TV cell_location;
            for (int a = 0; a < d; ++a) cell_location(a) = center(c[a]);
            for (int id = 0; id < desc.nplanes; ++id) {
                const mpm_plane& b = desc.planes[id];
                TV normal_vector, surface;
                for (int a = 0; a < d; ++a) { normal_vector(a) = (T)b.normal[a]; surface(a) = (T)b.surface[a]; }
                const T mu = (T)b.mu;
                TV rel;
                for (int a = 0; a < d; ++a) rel(a) = cell_location(a) - surface(a);
                const T result = rel.Dot(normal_vector);
                if (result < T(0)) {
                    TV vel;
                    for (int a = 0; a < d; ++a) vel(a) = n.ch[CH_VS + a];
                    const T projection = vel.Dot(normal_vector);
                    if (projection < T(0)) {
                        for (int a = 0; a < d; ++a) vel(a) -= normal_vector(a) * projection;
                        const T norm = vel.Norm();
                        if (-projection * mu < norm) {
                            for (int a = 0; a < d; ++a) vel(a) += vel(a) / norm * projection * mu;
                        } else {
                            vel = TV();
                        }
                    }
                    for (int a = 0; a < d; ++a) n.ch[CH_VS + a] = vel(a);
                }
            }
        });
    }

    // ---- CG system (Implicit_Force_Helper/MPM_CG_System.h) -----------------
    // Force (MPM_CG_System.h:150-215): f = sum_p V0 dP(T^T F) F^T grad w with
    // T = sum_i grad w_ip (x) x_i.
    void Force(int f_ch, int x_ch) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            Particle<T, d>& p = particles[(size_t)simulated_particles[(size_t)i]];
            if (p.eos) {
                // eos_scp = V0 K gamma rho^(gamma-1) (sum grad w . x) (MPM_CG_System.h:169-174);
                // the matrix-plus-scalar is taken as adding to the diagonal
                T s = T(0);
                for_stencil<d>(-1, 1, [&](const int* idx) {
                    int c[3];
                    cell_of(p, idx, c);
                    const Node<T>& n = node(c);
                    const TV g = p.Weight_Gradient(idx);
                    for (int a = 0; a < d; ++a) s += g(a) * n.ch[x_ch + a];
                });
                const T coefficient = p.volume * p.bulk_modulus * p.gamma * std::pow(p.density, p.gamma - T(1));
                p.eos_scp = TM::Identity() * (s * coefficient);
                continue;
            }
            TM tmp;
            for_stencil<d>(-1, 1, [&](const int* idx) {
                int c[3];
                cell_of(p, idx, c);
                const Node<T>& n = node(c);
                TV v_vec;
                for (int a = 0; a < d; ++a) v_vec(a) = n.ch[x_ch + a];
                tmp += TM::Outer_Product(p.Weight_Gradient(idx), v_vec);
            });
            const TM F = p.model.Fe;
            tmp = F.Times_Transpose(Times_dP_dF(p.model, tmp.Transpose_Times(F)));
            tmp *= p.volume;
            p.scp = tmp;
        }
        for_nodes([&](Node<T>& n, const int*) { for (int a = 0; a < d; ++a) n.ch[f_ch + a] = T(0); });  // :191
        slab_scatter([&](Particle<T, d>& p, int xlo, int xhi) {
            for_stencil<d>(xlo, xhi, [&](const int* idx) {
                int c[3];
                cell_of(p, idx, c);
                Node<T>& n = node(c);
                // EOS: eos_scp grad w (MPM_CG_System.h:205-207)
                const TV tmp_vec = p.eos ? p.eos_scp * p.Weight_Gradient(idx) : p.scp.Transpose_Times(p.Weight_Gradient(idx));
                for (int a = 0; a < d; ++a) n.ch[f_ch + a] += tmp_vec(a);
            });
        });
    }
    // Multiply (MPM_CG_System.h:73-81) + Multiply_Helper (Multiply_Helper.h:40-53)
    void Multiply(int x_ch, int r_ch, T dt) {
        ++cg_multiplies;
        Force(r_ch, x_ch);
        const T scaled_dt_squared = dt * dt / (T(1) + T(0));
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR)
                for (int a = 0; a < d; ++a) n.ch[r_ch + a] = scaled_dt_squared / n.ch[CH_M] * n.ch[r_ch + a] + n.ch[x_ch + a];
        });
    }
    // Project (MPM_CG_System.h:220-225, Project_Helper.h:30-45): zero the vector at
    // every node inside a plane (project_mode 0) -- see mpm_scene_desc.
    void Project(int v_ch) {
        if (desc.project_mode != 0) return;
        for_nodes([&](Node<T>& n, const int* c) {
            for (int id = 0; id < desc.nplanes; ++id) {
                const mpm_plane& b = desc.planes[id];
                T s = T(0);
                for (int a = 0; a < d; ++a) s += (center(c[a]) - (T)b.surface[a]) * (T)b.normal[a];
                if (s < T(0))
                    for (int a = 0; a < d; ++a) n.ch[v_ch + a] = T(0);
            }
        });
    }
    // Inner_Product (MPM_CG_System.h:227-239, Inner_Product_Helper.h:45-70):
    // sum over Interior nodes of a . (m b); T accumulation (or double).
    double Inner_Product(int a_ch, int b_ch) {
        const long total = (long)grid.size();
        if (dot_fp64) {
            std::vector<double> part((size_t)threads, 0.0);
#pragma omp parallel num_threads(threads)
            {
                double s = 0;
#pragma omp for schedule(static)
                for (long lin = 0; lin < total; ++lin) {
                    const Node<T>& n = grid[(size_t)lin];
                    if (n.flags & FLAG_INTERIOR)
                        for (int a = 0; a < d; ++a) s += (double)n.ch[a_ch + a] * ((double)n.ch[b_ch + a] * (double)n.ch[CH_M]);
                }
                part[(size_t)omp_get_thread_num()] = s;
            }
            double r = 0;
            for (double s : part) r += s;
            return r;
        }
        std::vector<T> part((size_t)threads, T(0));
#pragma omp parallel num_threads(threads)
        {
            T s = T(0);
#pragma omp for schedule(static)
            for (long lin = 0; lin < total; ++lin) {
                const Node<T>& n = grid[(size_t)lin];
                if (n.flags & FLAG_INTERIOR) {
                    T dot = T(0);
                    for (int a = 0; a < d; ++a) dot += n.ch[a_ch + a] * (n.ch[b_ch + a] * n.ch[CH_M]);
                    s += dot;
                }
            }
            part[(size_t)omp_get_thread_num()] = s;
        }
        T r = T(0);
        for (T s : part) r += s;
        return (double)r;
    }
    // Convergence_Norm (MPM_CG_System.h:241-249, Convergence_Norm_Helper.h:43-59)
    double Convergence_Norm(int v_ch) {
        const long total = (long)grid.size();
        double result = 0;
        if (dot_fp64) {
            double mx = 0;
#pragma omp parallel for num_threads(threads) reduction(max : mx)
            for (long lin = 0; lin < total; ++lin) {
                const Node<T>& n = grid[(size_t)lin];
                if (n.flags & FLAG_INTERIOR) {
                    double s = 0;
                    for (int a = 0; a < d; ++a) s += (double)n.ch[v_ch + a] * (double)n.ch[v_ch + a];
                    mx = std::max(mx, s);
                }
            }
            result = std::sqrt(mx);
        } else {
            T mx = T(0);
#pragma omp parallel for num_threads(threads) reduction(max : mx)
            for (long lin = 0; lin < total; ++lin) {
                const Node<T>& n = grid[(size_t)lin];
                if (n.flags & FLAG_INTERIOR) {
                    T s = T(0);
                    for (int a = 0; a < d; ++a) s += n.ch[v_ch + a] * n.ch[v_ch + a];
                    mx = std::max(mx, s);
                }
            }
            result = (double)std::sqrt(mx);
        }
        return result;
    }
    // Interior-masked vector ops (MPM_CG_Vector.h:49-101, Nova Masked_* arithmetic)
    void Copy(T c, int src, int dst) {  // dst = c src
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR) for (int a = 0; a < d; ++a) n.ch[dst + a] = c * n.ch[src + a];
        });
    }
    void Copy(T c, int v1, int v2, int dst) {  // dst = c v1 + v2
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR) for (int a = 0; a < d; ++a) n.ch[dst + a] = c * n.ch[v1 + a] + n.ch[v2 + a];
        });
    }
    void Subtract(int dst, int src) {
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR) for (int a = 0; a < d; ++a) n.ch[dst + a] = n.ch[dst + a] - n.ch[src + a];
        });
    }

    // Conjugate gradient (Nova Conjugate_Gradient<T>::Solve called at MPM_Example.cpp:711
    // with x = v*, b = rhs, min_iterations 0).
    void Solve(T dt) {
        const int x = CH_VS, b = CH_RHS, q = CH_Q, s = CH_S, r = CH_R;
        const double tol = (double)desc.cg_tolerance;
        const int max_it = desc.cg_max_iterations, min_it = desc.cg_min_iterations;
        cg_multiplies = 0;
        cg_converged = 0;
        double rho_old = (double)FLT_MAX, convergence_norm = 0;
        int iterations = 0;
        for (iterations = 0;; ++iterations) {
            const bool restart = iterations == 0;
            if (restart) {
                Copy(T(1), b, r);          // r = b
                Multiply(x, q, dt);        // q = A x
                Subtract(r, q);            // r -= q
                Project(r);
            }
            convergence_norm = Convergence_Norm(r);
            if (convergence_norm <= tol && (iterations >= min_it || convergence_norm < (double)std::numeric_limits<T>::epsilon())) {
                cg_converged = 1;
                break;
            }
            if (iterations == max_it) break;
            double rho = Inner_Product(r, r);
            if (!dot_fp64) rho = (double)(T)rho;
            if (restart) Copy(T(1), r, s);
            else Copy((T)(rho / rho_old), s, r, s);
            Multiply(s, q, dt);
            Project(q);
            double s_dot_q = Inner_Product(s, q);
            if (!dot_fp64) s_dot_q = (double)(T)s_dot_q;
            const double alpha = s_dot_q != 0.0 ? rho / s_dot_q : (double)FLT_MAX;
            Copy((T)alpha, s, x, x);      // x = alpha s + x
            Copy((T)(-alpha), q, r, r);   // r = -alpha q + r
            rho_old = rho;
        }
        cg_iterations = iterations;
        cg_residual = convergence_norm;
    }

    // Reset_Solver_Channels (MPM_Example.cpp:191-199)
    void Reset_Solver_Channels() {
        for_nodes([&](Node<T>& n, const int*) {
            for (int a = 0; a < d; ++a) n.ch[CH_Q + a] = n.ch[CH_R + a] = n.ch[CH_S + a] = n.ch[CH_Z + a] = T(0);
        });
    }
    // Apply_Force (MPM_Example.cpp:695-716) without the explicit part; the stage
    // split of the GPU boundary runs explicit force and collision separately.
    void Implicit_Solve(T dt) {
        // MPM_RHS_Helper (MPM_RHS_Helper.h:39-50)
        for_nodes([&](Node<T>& n, const int*) {
            if (n.flags & FLAG_INTERIOR) for (int a = 0; a < d; ++a) n.ch[CH_RHS + a] = n.ch[CH_VS + a];
        });
        Reset_Solver_Channels();
        Solve(dt);
    }

    // G2P part of Update_Particle_Velocities_And_Positions (MPM_Example.cpp:631-660)
    void G2P(T dt) {
        const T flip = (T)desc.flip;
        const T lo = T(0) + T(0.5) * dx, hi = T(1) - T(0.5) * dx;
        std::vector<std::vector<int>> remove_indices((size_t)threads);
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            const int id = simulated_particles[(size_t)i];
            Particle<T, d>& p = particles[(size_t)id];
            TV V_pic, V_flip = p.V;
            TM grad_Vp;
            for_stencil<d>(-1, 1, [&](const int* idx) {
                int c[3];
                cell_of(p, idx, c);
                const Node<T>& n = node(c);
                const T weight = p.Weight(idx);
                if (weight > T(0)) {
                    TV V_grid, delta_V_grid;
                    for (int a = 0; a < d; ++a) {
                        V_grid(a) = n.ch[CH_VS + a];
                        delta_V_grid(a) = n.ch[CH_VS + a] - n.ch[CH_V + a];
                    }
                    for (int a = 0; a < d; ++a) {
                        V_pic(a) += weight * V_grid(a);
                        V_flip(a) += weight * delta_V_grid(a);
                    }
                    grad_Vp += TM::Outer_Product(V_grid, p.Weight_Gradient(idx));
                }
            });
            // MPM_Example.cpp:645-647
            if (p.eos) p.density /= T(1) + dt * grad_Vp.Trace();
            else p.model.Fe += grad_Vp * dt * p.model.Fe;
            for (int a = 0; a < d; ++a) {
                p.V(a) = V_flip(a) * flip + V_pic(a) * (T(1) - flip);
                p.X(a) += V_pic(a) * dt;
            }
            bool inside = true;
            for (int a = 0; a < d; ++a) inside = inside && p.X(a) >= lo && p.X(a) <= hi;
            if (!inside) {
                remove_indices[(size_t)omp_get_thread_num()].push_back((int)i);
                p.valid = false;
            }
        }
        std::vector<int> all;
        for (auto& r : remove_indices) all.insert(all.end(), r.begin(), r.end());
        std::sort(all.begin(), all.end());
        for (int i = (int)all.size() - 1; i >= 0; --i) {
            const int k = all[(size_t)i];
            invalid_particles.push_back(simulated_particles[(size_t)k]);
            simulated_particles.erase(simulated_particles.begin() + k);
        }
    }

    // ---- driver -------------------------------------------------------------
    void run_stage(int stage, double dt_d) override {
        const T dt = (T)dt_d;
        switch (stage) {
            case MPM_STAGE_BIN:
                Initialize_SPGrid_And_Voxels();
                Update_Particle_Weights();
                Group_Particles();
                break;
            case MPM_STAGE_P2G: Rasterize(); break;
            case MPM_STAGE_WAITING: Process_Waiting_Particles(); break;
            case MPM_STAGE_CONSTITUTIVE: Update_Constitutive_Model_State(); break;
            case MPM_STAGE_EXPLICIT_FORCE: Apply_Explicit_Force(dt); break;
            case MPM_STAGE_COLLISION: Grid_Based_Collision(); break;
            case MPM_STAGE_CG: Implicit_Solve(dt); break;
            case MPM_STAGE_G2P: G2P(dt); break;
            default: break;
        }
    }
    // MPM_Driver::Initialize (MPM_Driver.cpp:18-32)
    void initialize() override {
        populate();
        run_stage(MPM_STAGE_BIN, 0);
        run_stage(MPM_STAGE_P2G, 0);
        run_stage(MPM_STAGE_WAITING, 0);
    }
    // MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116)
    void advance_step(double dt) override {
        for (int s = MPM_STAGE_BIN; s <= MPM_STAGE_G2P; ++s) run_stage(s, dt);
    }
    void populate() override {  // Populate_Simulated_Particles (MPM_Example.cpp:347-355)
        simulated_particles.clear();
        for (size_t i = 0; i < particles.size(); ++i)
            if (particles[i].valid) simulated_particles.push_back((int)i);
    }
    double max_velocity() const override {  // MPM_Example.cpp:375-387
        std::vector<T> r((size_t)threads, T(0));
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long i = 0; i < (long)simulated_particles.size(); ++i) {
            const int tid = omp_get_thread_num();
            r[(size_t)tid] = std::max(r[(size_t)tid], particles[(size_t)simulated_particles[(size_t)i]].V.Norm_Squared());
        }
        T result = T(0);
        for (T v : r) result = std::max(result, v);
        return (double)std::sqrt(result);
    }
    // MPM_Driver::Advance_To_Target_Time (MPM_Driver.cpp:55-80)
    double advance_to_target_time(double time_d, double target_d, int* substeps) override {
        const T min_dt = (T)desc.min_dt, max_dt = (T)desc.max_dt, cfl = (T)desc.cfl, dx_min = dx;
        T time = (T)time_d;
        const T target_time = (T)target_d;
        bool done = false;
        populate();
        int n = 0;
        for (int substep = 1; !done; ++substep) {
            const T max_v = (T)max_velocity();
            T dt = std::max(min_dt, std::min(max_dt, cfl * dx_min / std::max(max_v, (T)1e-2)));
            if (target_time - time < dt * (T)1.001) {
                dt = target_time - time;
                done = true;
            } else if (target_time - time < T(2) * dt) {
                dt = (target_time - time) * T(0.5);
            }
            advance_step((double)dt);
            time += dt;
            ++n;
        }
        if (substeps) *substeps = n;
        return (double)time;
    }

    // ---- inspection -----------------------------------------------------------
    int64_t dense_index(const int* c) const {  // over [-1, counts]^d
        int64_t r = 0;
        for (int a = 0; a < d; ++a) r = r * (counts + 2) + (c[a] + 1);
        return r;
    }
    int channel_base(int channel, int* width) const {
        *width = d;
        switch (channel) {
            case MPM_CH_MASS: *width = 1; return CH_M;
            case MPM_CH_VELOCITY: return CH_V;
            case MPM_CH_VELOCITY_STAR: return CH_VS;
            case MPM_CH_FORCE: return CH_F;
            case MPM_CH_CG_R: return CH_R;
            case MPM_CH_CG_S: return CH_S;
            case MPM_CH_CG_Q: return CH_Q;
            default: *width = 1; return CH_M;
        }
    }
    void get_grid(int channel, double* out) const override {
        int width;
        const int base = channel_base(channel, &width);
        int64_t total = 1;
        for (int a = 0; a < d; ++a) total *= counts + 2;
        std::memset(out, 0, sizeof(double) * (size_t)(total * width));
        for (size_t lin = 0; lin < grid.size(); ++lin) {
            int c[3] = {0, 0, 0};
            node_coord((int64_t)lin, c);
            bool ok = true;
            for (int a = 0; a < d; ++a) ok = ok && c[a] >= -1 && c[a] <= counts;
            if (!ok) continue;
            const int64_t o = dense_index(c);
            for (int w = 0; w < width; ++w) out[o * width + w] = (double)grid[lin].ch[base + w];
        }
    }
    void get_flags(uint32_t* out) const override {
        int64_t total = 1;
        for (int a = 0; a < d; ++a) total *= counts + 2;
        std::memset(out, 0, sizeof(uint32_t) * (size_t)total);
        for (size_t lin = 0; lin < grid.size(); ++lin) {
            int c[3] = {0, 0, 0};
            node_coord((int64_t)lin, c);
            bool ok = true;
            for (int a = 0; a < d; ++a) ok = ok && c[a] >= -1 && c[a] <= counts;
            if (ok) out[dense_index(c)] = grid[lin].flags;
        }
    }
    void load_dense(int ch, const double* src) {
        for (size_t lin = 0; lin < grid.size(); ++lin) {
            int c[3] = {0, 0, 0};
            node_coord((int64_t)lin, c);
            bool ok = true;
            for (int a = 0; a < d; ++a) ok = ok && c[a] >= -1 && c[a] <= counts;
            for (int a = 0; a < d; ++a) grid[lin].ch[ch + a] = ok ? (T)src[dense_index(c) * d + a] : T(0);
        }
    }
    void set_grid(int channel, const double* in) override {
        int width;
        const int base = channel_base(channel, &width);
        for (size_t lin = 0; lin < grid.size(); ++lin) {
            int c[3] = {0, 0, 0};
            node_coord((int64_t)lin, c);
            bool ok = true;
            for (int a = 0; a < d; ++a) ok = ok && c[a] >= -1 && c[a] <= counts;
            if (!ok) continue;
            const int64_t o = dense_index(c);
            for (int w = 0; w < width; ++w) grid[lin].ch[base + w] = (T)in[o * width + w];
        }
    }
    void apply_operator(double dt, const double* x, double* q) override {
        load_dense(CH_Z, x);
        Multiply(CH_Z, CH_Q, (T)dt);
        Project(CH_Q);
        get_grid(MPM_CH_CG_Q, q);
    }
    double inner_product(const double* a, const double* b) const override {
        double s = 0;
        for (size_t lin = 0; lin < grid.size(); ++lin) {
            int c[3] = {0, 0, 0};
            node_coord((int64_t)lin, c);
            if (!(grid[lin].flags & FLAG_INTERIOR)) continue;
            const int64_t o = dense_index(c);
            for (int k = 0; k < d; ++k) s += a[o * d + k] * (b[o * d + k] * (double)grid[lin].ch[CH_M]);
        }
        return s;
    }
    void cg_stats(int* it, int* mult, int* conv, double* res) const override {
        if (it) *it = cg_iterations;
        if (mult) *mult = cg_multiplies;
        if (conv) *conv = cg_converged;
        if (res) *res = cg_residual;
    }
    void bbox_cells(int* lo, int* hi) const override {
        for (int a = 0; a < d; ++a) { lo[a] = bbox_lo[a]; hi[a] = bbox_hi[a]; }
    }
};

OracleBase* make(const mpm_scene_desc& desc, int precision, int threads, int options) {
    if (desc.dim == 3) {
        if (precision == 8) return new Oracle<double, 3>(desc, threads, options);
        return new Oracle<float, 3>(desc, threads, options);
    }
    if (precision == 8) return new Oracle<double, 2>(desc, threads, options);
    return new Oracle<float, 2>(desc, threads, options);
}

template <int d>
void to_mat(const double* rm, Mat<double, d>& m) { for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) m(i, j) = rm[i * d + j]; }
template <int d>
void from_mat(const Mat<double, d>& m, double* rm) { for (int i = 0; i < d; ++i) for (int j = 0; j < d; ++j) rm[i * d + j] = m(i, j); }

template <int d>
void constitutive_d(const mpm_material* mat, const double* F_in, double Jp_in, double* F_out, double* Jp_out,
                    double* R, double* Pout, double* psi) {
    Model<double, d> m;
    m.mu0 = mat->mu0;
    m.lambda0 = mat->lambda0;
    m.plastic = mat->plastic != 0;
    m.stretching_yield = mat->stretching_yield;
    m.compression_yield = mat->compression_yield;
    m.hardening_factor = mat->hardening_factor;
    m.hardening_max = mat->hardening_max;
    to_mat<d>(F_in, m.Fe);
    m.Jp = Jp_in;
    m.Fp = Mat<double, d>::Identity();
    m.Fp(0, 0) = Jp_in;
    Precompute(m, true, false);
    if (F_out) from_mat<d>(m.Fe, F_out);
    if (Jp_out) *Jp_out = Determinant(m.Fp);
    if (R) from_mat<d>(m.Re, R);
    if (Pout) from_mat<d>(P(m), Pout);
    if (psi) *psi = Psi(m);
}

template <int d>
void dP_d(const double* F, const double* R, double lambda, double mu, const double* dF, double* out) {
    Model<double, d> m;
    to_mat<d>(F, m.Fe);
    to_mat<d>(R, m.Re);
    m.He = Cofactor(m.Fe);
    m.Je = Determinant(m.Fe);
    m.lambda = lambda;
    m.mu = mu;
    Mat<double, d> D;
    to_mat<d>(dF, D);
    from_mat<d>(Times_dP_dF(m, D), out);
}

}  // namespace

struct oracle_ctx {
    std::unique_ptr<OracleBase> impl;
};

extern "C" {

oracle_handle oracle_create(const mpm_scene_desc* desc, int32_t precision, int32_t threads, int32_t options) {
    if (!desc || (desc->dim != 2 && desc->dim != 3) || desc->counts < 4) return nullptr;
    oracle_ctx* h = new oracle_ctx;
    h->impl.reset(make(*desc, precision, threads, options));
    return h;
}
void oracle_destroy(oracle_handle h) { delete h; }
void oracle_set_particles(oracle_handle h, int64_t n, const float* x, const float* v, const float* F,
                          const float* Jp, const float* mass, const float* volume) {
    h->impl->set_particles(n, x, v, F, Jp, mass, volume, nullptr);
}

void oracle_add_particles(oracle_handle h, int64_t n, const float* x, const float* v, const float* mass,
                          const int32_t* material) {
    h->impl->add_particles(n, x, v, mass, material);
}

void oracle_set_particles_ex(oracle_handle h, int64_t n, const float* x, const float* v, const float* F,
                             const float* Jp, const float* mass, const float* volume, const int32_t* material) {
    h->impl->set_particles(n, x, v, F, Jp, mass, volume, material);
}
int64_t oracle_num_particles(oracle_handle h) { return h->impl->num_particles(); }
int64_t oracle_num_simulated(oracle_handle h) { return h->impl->num_simulated(); }
void oracle_get_particles(oracle_handle h, double* x, double* v, double* F, double* Jp, double* mass,
                          double* volume, uint8_t* valid) {
    h->impl->get_particles(x, v, F, Jp, mass, volume, valid);
}
void oracle_initialize(oracle_handle h) { h->impl->initialize(); }
void oracle_advance_step(oracle_handle h, double dt) { h->impl->advance_step(dt); }
void oracle_run_stage(oracle_handle h, int32_t stage, double dt) { h->impl->run_stage(stage, dt); }
void oracle_populate_simulated(oracle_handle h) { h->impl->populate(); }
double oracle_max_particle_velocity(oracle_handle h) { return h->impl->max_velocity(); }
double oracle_advance_to_target_time(oracle_handle h, double time, double target_time, int32_t* substeps) {
    int n = 0;
    const double t = h->impl->advance_to_target_time(time, target_time, &n);
    if (substeps) *substeps = n;
    return t;
}
void oracle_get_grid(oracle_handle h, int32_t channel, double* out) { h->impl->get_grid(channel, out); }
void oracle_get_flags(oracle_handle h, uint32_t* out) { h->impl->get_flags(out); }
void oracle_apply_operator(oracle_handle h, double dt, const double* x, double* q) { h->impl->apply_operator(dt, x, q); }
double oracle_inner_product(oracle_handle h, const double* a, const double* b) { return h->impl->inner_product(a, b); }
void oracle_set_grid(oracle_handle h, int32_t channel, const double* in) { h->impl->set_grid(channel, in); }
void oracle_cg_stats(oracle_handle h, int32_t* iterations, int32_t* multiplies, int32_t* converged, double* residual) {
    int it = 0, m = 0, c = 0;
    h->impl->cg_stats(&it, &m, &c, residual);
    if (iterations) *iterations = it;
    if (multiplies) *multiplies = m;
    if (converged) *converged = c;
}
void oracle_bbox_cells(oracle_handle h, int32_t* lo, int32_t* hi) {
    int l[3] = {0, 0, 0}, u[3] = {0, 0, 0};
    h->impl->bbox_cells(l, u);
    for (int a = 0; a < 3; ++a) { lo[a] = l[a]; hi[a] = u[a]; }
}

void oracle_weights(int32_t dim, int32_t counts, const double* X, int32_t* cell, double* w, double* dw) {
    if (dim == 3) {
        Particle<double, 3> p;
        for (int a = 0; a < 3; ++a) p.X(a) = X[a];
        p.Update_Weights(counts);
        for (int a = 0; a < 3; ++a) cell[a] = p.closest_cell[a];
        for (int k = 0; k < 3; ++k) for (int a = 0; a < 3; ++a) { w[k * 3 + a] = p.weights[k][a]; dw[k * 3 + a] = p.dweights[k][a]; }
    } else {
        Particle<double, 2> p;
        for (int a = 0; a < 2; ++a) p.X(a) = X[a];
        p.Update_Weights(counts);
        for (int a = 0; a < 2; ++a) cell[a] = p.closest_cell[a];
        for (int k = 0; k < 3; ++k) for (int a = 0; a < 2; ++a) { w[k * 2 + a] = p.weights[k][a]; dw[k * 2 + a] = p.dweights[k][a]; }
    }
}
void oracle_svd(int32_t dim, const double* F, double* U, double* S, double* V) {
    if (dim == 3) svd3<double>(F, U, S, V);
    else svd2<double>(F, U, S, V);
}
void oracle_constitutive(int32_t dim, const mpm_material* mat, const double* F_in, double Jp_in, double* F_out,
                         double* Jp_out, double* R, double* P, double* psi) {
    if (dim == 3) constitutive_d<3>(mat, F_in, Jp_in, F_out, Jp_out, R, P, psi);
    else constitutive_d<2>(mat, F_in, Jp_in, F_out, Jp_out, R, P, psi);
}
void oracle_dP(int32_t dim, const double* F, const double* R, double lambda, double mu, const double* dF, double* dP) {
    if (dim == 3) dP_d<3>(F, R, lambda, mu, dF, dP);
    else dP_d<2>(F, R, lambda, mu, dF, dP);
}

}  // extern "C"
