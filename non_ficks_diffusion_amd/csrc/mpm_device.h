// Device-side maths of the gfx950 MPM snow substep: B-spline weights,
// small-matrix algebra, the FP64 SVD, the snow plasticity update and the
// fixed-corotated stress and its differential.  Matrices are row-major
// arrays (m[3*i+j] = row i, column j).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

namespace mpm {

template <int D>
struct Dims;
template <>
struct Dims<3> {
    static constexpr int B = 4;          // block edge (nodes)
    static constexpr int NB = 64;        // nodes per block
    static constexpr int TE = 6;         // tile edge: block plus one halo node per side
    static constexpr int TN = 216;       // tile nodes
    // LDS copy of a tile: node (x,y,z) at x*LX + y*LY + z.  With these
    // strides the 64 cells of a block (lanes of a wave) fall on distinct
    // banks of the 32-bank LDS for every stencil offset.
    static constexpr int LX = 52, LY = 8, LTN = 6 * 52;
    static constexpr int ST = 27;        // stencil nodes
    static constexpr int NBR = 27;       // neighbour blocks incl. self
    static constexpr int DD = 9;
};
template <>
struct Dims<2> {
    static constexpr int B = 8;
    static constexpr int NB = 64;
    static constexpr int TE = 10;
    static constexpr int TN = 100;
    static constexpr int LX = 40, LY = 1, LTN = 10 * 40;  // node (x,y) at x*LX + y
    static constexpr int ST = 9;
    static constexpr int NBR = 9;
    static constexpr int DD = 4;
};

struct Material {
    float mu0, lambda0, sy, cy, xi, hmax;
    int plastic;
    int eos;                 // EOS fluid: K, gamma; density in the particle's Jp slot
    float K, gamma, rho0;
};
constexpr int MAX_MATERIALS = 8;

// ---------------------------------------------------------------------------
// Quadratic B-spline weights (MPM_Particle.h:79-100).  The arithmetic is
// written operation for operation as the reference states it, with FMA
// contraction off, so keys, weights and the m > 0 test agree bit for bit
// with a scalar CPU evaluation.
// ---------------------------------------------------------------------------
// The cell c of coordinate X and X's offset from the stencil's first node in
// cells (in [0.5, 1.5)).
__device__ inline float stencil_offset(float X, float dx, float inv_dx, int& c) {
#pragma clang fp contract(off)
    c = (int)floorf(X * inv_dx);
    const float center = ((float)c + 0.5f) * dx;
    return (X - (center - dx)) * inv_dx;
}
// The three weights and their derivatives at stencil offset x.
__device__ inline void bspline_from_offset(float x, float inv_dx, float w[3], float dw[3]) {
#pragma clang fp contract(off)
    w[0] = fabsf(0.5f * x * x - 1.5f * x + 1.125f);
    dw[0] = (x - 1.5f) * inv_dx;
    x -= 1.0f;
    w[1] = fabsf(-x * x + 0.75f);
    dw[1] = -2.0f * x * inv_dx;
    x -= 1.0f;
    w[2] = fabsf(0.5f * x * x + 1.5f * x + 1.125f);
    dw[2] = (x + 1.5f) * inv_dx;
}
__device__ inline void bspline_weights(float X, float dx, float inv_dx, int& c, float w[3], float dw[3]) {
    bspline_from_offset(stencil_offset(X, dx, inv_dx, c), inv_dx, w, dw);
}

__device__ inline int cell_of(float X, float inv_dx) {
#pragma clang fp contract(off)
    return (int)floorf(X * inv_dx);
}

// ---------------------------------------------------------------------------
// 3x3 / 2x2 helpers (float and double)
// ---------------------------------------------------------------------------
template <class S>
__device__ inline S det3(const S* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// cofactor matrix C = det(F) F^{-T}: C column j = (column j+1) x (column j+2) of F
template <class S>
__device__ inline void cof3(const S* F, S* C) {
    // columns f0=(F0,F3,F6) f1=(F1,F4,F7) f2=(F2,F5,F8)
    C[0] = F[4] * F[8] - F[7] * F[5];
    C[3] = F[7] * F[2] - F[1] * F[8];
    C[6] = F[1] * F[5] - F[4] * F[2];
    C[1] = F[5] * F[6] - F[8] * F[3];
    C[4] = F[8] * F[0] - F[2] * F[6];
    C[7] = F[2] * F[3] - F[5] * F[0];
    C[2] = F[3] * F[7] - F[6] * F[4];
    C[5] = F[6] * F[1] - F[0] * F[7];
    C[8] = F[0] * F[4] - F[3] * F[1];
}
// directional derivative of cof3 at F along G
template <class S>
__device__ inline void dcof3(const S* F, const S* G, S* C) {
    C[0] = G[4] * F[8] + F[4] * G[8] - G[7] * F[5] - F[7] * G[5];
    C[3] = G[7] * F[2] + F[7] * G[2] - G[1] * F[8] - F[1] * G[8];
    C[6] = G[1] * F[5] + F[1] * G[5] - G[4] * F[2] - F[4] * G[2];
    C[1] = G[5] * F[6] + F[5] * G[6] - G[8] * F[3] - F[8] * G[3];
    C[4] = G[8] * F[0] + F[8] * G[0] - G[2] * F[6] - F[2] * G[6];
    C[7] = G[2] * F[3] + F[2] * G[3] - G[5] * F[0] - F[5] * G[0];
    C[2] = G[3] * F[7] + F[3] * G[7] - G[6] * F[4] - F[6] * G[4];
    C[5] = G[6] * F[1] + F[6] * G[1] - G[0] * F[7] - F[0] * G[7];
    C[8] = G[0] * F[4] + F[0] * G[4] - G[3] * F[1] - F[3] * G[1];
}

// ---------------------------------------------------------------------------
// FP64 SVD.  Eigen-decomposition of the normal matrix F^T F by cyclic
// Jacobi rotations, singular values descending, V made a rotation, U built
// by Gram-Schmidt of F V (u3 = u1 x u2) so U is a rotation too; the sign
// of det F ends up on the smallest singular value.
// ---------------------------------------------------------------------------
__device__ inline void jacobi_pair(double& app, double& aqq, double& apq, double& arp, double& arq,
                                   double* V, int p, int q) {
    if (apq == 0.0) return;
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = copysign(1.0, tau) / (fabs(tau) + sqrt(1.0 + tau * tau));
    const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
    app -= t * apq;
    aqq += t * apq;
    apq = 0.0;
    const double rp = arp, rq = arq;
    arp = c * rp - s * rq;
    arq = s * rp + c * rq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double vp = V[3 * k + p], vq = V[3 * k + q];
        V[3 * k + p] = c * vp - s * vq;
        V[3 * k + q] = s * vp + c * vq;
    }
}

__device__ inline void svd3_fp64(const double* F, double* U, double* sig, double* V) {
    // normal matrix
    double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double f0 = F[3 * k], f1 = F[3 * k + 1], f2 = F[3 * k + 2];
        a00 += f0 * f0; a01 += f0 * f1; a02 += f0 * f2;
        a11 += f1 * f1; a12 += f1 * f2; a22 += f2 * f2;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        const double off = a01 * a01 + a02 * a02 + a12 * a12;
        const double dia = a00 * a00 + a11 * a11 + a22 * a22;
        if (!(off > 1e-300) || off <= dia * 1e-36) break;
        jacobi_pair(a00, a11, a01, a02, a12, V, 0, 1);  // (0,1); r = 2
        jacobi_pair(a00, a22, a02, a01, a12, V, 0, 2);  // (0,2); r = 1: a_r0 = a01, a_r2 = a12
        jacobi_pair(a11, a22, a12, a01, a02, V, 1, 2);  // (1,2); r = 0: a_r1 = a01, a_r2 = a02
    }
    double lam[3] = {a00, a11, a22};
    // sort descending (columns of V follow)
    auto swapcol = [&](int i, int j) {
        const double t = lam[i]; lam[i] = lam[j]; lam[j] = t;
#pragma unroll
        for (int k = 0; k < 3; ++k) { const double v = V[3 * k + i]; V[3 * k + i] = V[3 * k + j]; V[3 * k + j] = v; }
    };
    if (lam[0] < lam[1]) swapcol(0, 1);
    if (lam[0] < lam[2]) swapcol(0, 2);
    if (lam[1] < lam[2]) swapcol(1, 2);
    // det V = +1
    const double cx = V[3] * V[7] - V[6] * V[4], cy = V[6] * V[1] - V[0] * V[7], cz = V[0] * V[4] - V[3] * V[1];
    if (cx * V[2] + cy * V[5] + cz * V[8] < 0.0) { V[2] = -V[2]; V[5] = -V[5]; V[8] = -V[8]; }
    // F v_k
    double Fv[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) Fv[k][i] = F[3 * i] * V[k] + F[3 * i + 1] * V[3 + k] + F[3 * i + 2] * V[6 + k];
    double u0[3], u1[3], u2[3];
    const double n0 = sqrt(Fv[0][0] * Fv[0][0] + Fv[0][1] * Fv[0][1] + Fv[0][2] * Fv[0][2]);
    if (n0 > 0.0) { const double r = 1.0 / n0; u0[0] = Fv[0][0] * r; u0[1] = Fv[0][1] * r; u0[2] = Fv[0][2] * r; }
    else { u0[0] = 1.0; u0[1] = 0.0; u0[2] = 0.0; }
    const double pr = u0[0] * Fv[1][0] + u0[1] * Fv[1][1] + u0[2] * Fv[1][2];
    double w0 = Fv[1][0] - pr * u0[0], w1 = Fv[1][1] - pr * u0[1], w2 = Fv[1][2] - pr * u0[2];
    double n1 = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    if (n1 > 2.220446049250313e-16 * (n0 > 0.0 ? n0 : 1.0)) {
        const double r = 1.0 / n1;
        u1[0] = w0 * r; u1[1] = w1 * r; u1[2] = w2 * r;
    } else {
        // any unit vector orthogonal to u0: cross with the least-aligned axis
        const double ax = fabs(u0[0]), ay = fabs(u0[1]), az = fabs(u0[2]);
        double e0 = 0, e1 = 0, e2 = 0;
        if (ax <= ay && ax <= az) e0 = 1; else if (ay <= az) e1 = 1; else e2 = 1;
        double t0 = u0[1] * e2 - u0[2] * e1, t1 = u0[2] * e0 - u0[0] * e2, t2 = u0[0] * e1 - u0[1] * e0;
        const double r = 1.0 / sqrt(t0 * t0 + t1 * t1 + t2 * t2);
        u1[0] = t0 * r; u1[1] = t1 * r; u1[2] = t2 * r;
        n1 = u1[0] * Fv[1][0] + u1[1] * Fv[1][1] + u1[2] * Fv[1][2];
    }
    u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
    u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
    u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
    sig[0] = n0;
    sig[1] = n1;
    sig[2] = u2[0] * Fv[2][0] + u2[1] * Fv[2][1] + u2[2] * Fv[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) { U[3 * i] = u0[i]; U[3 * i + 1] = u1[i]; U[3 * i + 2] = u2[i]; }
}

__device__ inline void svd2_fp64(const double* F, double* U, double* sig, double* V) {
    const double a = F[0] * F[0] + F[2] * F[2], b = F[0] * F[1] + F[2] * F[3], c = F[1] * F[1] + F[3] * F[3];
    double cs = 1.0, sn = 0.0;
    if (b != 0.0) {
        const double tau = (c - a) / (2.0 * b);
        const double t = copysign(1.0, tau) / (fabs(tau) + sqrt(1.0 + tau * tau));
        cs = 1.0 / sqrt(1.0 + t * t);
        sn = t * cs;
    }
    // eigenvalues of the rotated normal matrix
    const double l0 = a - (sn / cs) * b, l1 = c + (sn / cs) * b;
    double v0x = cs, v0y = -sn, v1x = sn, v1y = cs;
    if (l1 > l0) { const double tx = v0x, ty = v0y; v0x = v1x; v0y = v1y; v1x = -tx; v1y = -ty; }
    if (v0x * v1y - v0y * v1x < 0.0) { v1x = -v1x; v1y = -v1y; }
    const double f0x = F[0] * v0x + F[1] * v0y, f0y = F[2] * v0x + F[3] * v0y;
    const double f1x = F[0] * v1x + F[1] * v1y, f1y = F[2] * v1x + F[3] * v1y;
    const double n0 = sqrt(f0x * f0x + f0y * f0y);
    double u0x = 1.0, u0y = 0.0;
    if (n0 > 0.0) { u0x = f0x / n0; u0y = f0y / n0; }
    const double u1x = -u0y, u1y = u0x;
    sig[0] = n0;
    sig[1] = u1x * f1x + u1y * f1y;
    U[0] = u0x; U[1] = u1x; U[2] = u0y; U[3] = u1y;
    V[0] = v0x; V[1] = v1x; V[2] = v0y; V[3] = v1y;
}

// ---------------------------------------------------------------------------
// Snow plasticity (MPM_Constitutive_Model.cpp:32-51) in FP64: clamp the
// singular values of Fe to [compression_yield, stretching_yield], move the
// excess into Jp = det Fp (Jp *= prod sigma / prod sigma_clamped), rebuild
// Fe = U Sigma V^T, and return R = U V^T.
// ---------------------------------------------------------------------------
template <int D>
__device__ inline void plasticity(float* Fe, float& Jp, float* R, const Material& mat, float* Uout = nullptr,
                                  double* Sout = nullptr) {
    constexpr int DD = D * D;
    double F[DD], U[DD], V[DD], s[D];
#pragma unroll
    for (int i = 0; i < DD; ++i) F[i] = (double)Fe[i];
    if constexpr (D == 3) svd3_fp64(F, U, s, V);
    else svd2_fp64(F, U, s, V);
    if (mat.plastic) {
        double det_old = 1.0, det_new = 1.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            det_old *= s[i];
            if (s[i] > (double)mat.sy) s[i] = (double)mat.sy;
            if (s[i] < (double)mat.cy) s[i] = (double)mat.cy;
            det_new *= s[i];
        }
        Jp = (float)((double)Jp * (det_old / det_new));
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < D; ++k) a += U[D * i + k] * s[k] * V[D * j + k];
                Fe[D * i + j] = (float)a;
            }
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) a += U[D * i + k] * V[D * j + k];
            R[D * i + j] = (float)a;
        }
    if (Uout) {
#pragma unroll
        for (int i = 0; i < DD; ++i) Uout[i] = (float)U[i];
    }
    if (Sout) {  // singular values of the (clamped) Fe
#pragma unroll
        for (int i = 0; i < D; ++i) Sout[i] = s[i];
    }
}

// Lame parameters after hardening (MPM_Constitutive_Model.cpp:42-47).
__device__ inline void lame(const Material& mat, float Jp, float& lambda, float& mu) {
    if (mat.plastic) {
        const float power = fminf(mat.xi * (1.0f - Jp), mat.hmax);
        const float e = expf(power);
        lambda = e * mat.lambda0;
        mu = e * mat.mu0;
    } else {
        lambda = mat.lambda0;
        mu = mat.mu0;
    }
}

// Stress state used by P and dP: He = cof(Fe), Je = det Fe.
template <int D>
struct StressState {
    float F[D * D], R[D * D], H[D * D];
    float J, lambda, mu;
};

template <int D>
__device__ inline void stress_state(const float* F, const float* R, float Jp, const Material& mat, StressState<D>& st) {
#pragma unroll
    for (int i = 0; i < D * D; ++i) { st.F[i] = F[i]; st.R[i] = R[i]; }
    if constexpr (D == 3) {
        cof3(F, st.H);
        st.J = det3(F);
    } else {
        st.H[0] = F[3]; st.H[1] = -F[2]; st.H[2] = -F[1]; st.H[3] = F[0];
        st.J = F[0] * F[3] - F[1] * F[2];
    }
    lame(mat, Jp, st.lambda, st.mu);
}

// P = 2 mu (Fe - Re) + lambda (Je - 1) He   (MPM_Constitutive_Model.cpp:63-67)
template <int D>
__device__ inline void first_pk(const StressState<D>& st, float* P) {
    const float a = 2.0f * st.mu, b = st.lambda * (st.J - 1.0f);
#pragma unroll
    for (int i = 0; i < D * D; ++i) P[i] = a * (st.F[i] - st.R[i]) + b * st.H[i];
}

// dP = 2 mu (dF - dR) + lambda (He : dF) He + lambda (Je - 1) dHe[dF]
// (MPM_Constitutive_Model.cpp:71-76, dR from MPM_Constitutive_Model.h:47-69)
template <int D>
__device__ inline void dP_of(const StressState<D>& st, const float* dF, float* dP) {
    float dR[D * D], dH[D * D];
    if constexpr (D == 3) {
        const float* R = st.R;
        const float* F = st.F;
        float M[9], S[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                M[3 * i + j] = R[i] * dF[j] + R[3 + i] * dF[3 + j] + R[6 + i] * dF[6 + j];
                S[3 * i + j] = R[i] * F[j] + R[3 + i] * F[3 + j] + R[6 + i] * F[6 + j];
            }
        // K (a,b,c) = L with R^T dR = [[0,-a,-b],[a,0,-c],[b,c,0]]
        const float k00 = S[0] + S[4], k11 = S[0] + S[8], k22 = S[4] + S[8];
        const float k01 = S[7], k02 = -S[6], k12 = S[3];
        const float L0 = M[3] - M[1], L1 = M[6] - M[2], L2 = M[7] - M[5];
        const float c00 = k11 * k22 - k12 * k12, c01 = k02 * k12 - k01 * k22, c02 = k01 * k12 - k02 * k11;
        const float c11 = k00 * k22 - k02 * k02, c12 = k01 * k02 - k00 * k12, c22 = k00 * k11 - k01 * k01;
        const float inv = 1.0f / (k00 * c00 + k01 * c01 + k02 * c02);
        const float a = (c00 * L0 + c01 * L1 + c02 * L2) * inv;
        const float b = (c01 * L0 + c11 * L1 + c12 * L2) * inv;
        const float c = (c02 * L0 + c12 * L1 + c22 * L2) * inv;
        // dR = R W
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float r0 = R[3 * i], r1 = R[3 * i + 1], r2 = R[3 * i + 2];
            dR[3 * i + 0] = r1 * a + r2 * b;
            dR[3 * i + 1] = -r0 * a + r2 * c;
            dR[3 * i + 2] = -r0 * b - r1 * c;
        }
        dcof3(F, dF, dH);
    } else {
        const float* R = st.R;
        const float* F = st.F;
        const float m01 = R[0] * dF[1] + R[2] * dF[3], m10 = R[1] * dF[0] + R[3] * dF[2];
        const float s00 = R[0] * F[0] + R[2] * F[2], s11 = R[1] * F[1] + R[3] * F[3];
        const float a = (m10 - m01) / (s11 + s00);
        // R W with W = [[0,-a],[a,0]]
        dR[0] = R[1] * a;  dR[1] = -R[0] * a;
        dR[2] = R[3] * a;  dR[3] = -R[2] * a;
        dH[0] = dF[3]; dH[1] = -dF[2]; dH[2] = -dF[1]; dH[3] = dF[0];
    }
    float HdF = 0.0f;
#pragma unroll
    for (int i = 0; i < D * D; ++i) HdF += st.H[i] * dF[i];
    const float a2 = 2.0f * st.mu, lh = st.lambda * HdF, lj = st.lambda * (st.J - 1.0f);
#pragma unroll
    for (int i = 0; i < D * D; ++i) dP[i] = a2 * (dF[i] - dR[i]) + lh * st.H[i] + lj * dH[i];
}

}  // namespace mpm
