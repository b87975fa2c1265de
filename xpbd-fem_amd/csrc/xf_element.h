// xf_element.h -- the element arithmetic of the linear-tet projection, shared
// by the forward (xf_kernels.hip), adjoint (xf_adjoint.hip) and vertex-ring
// (xf_rings.hip) kernels, so that one definition fixes the rounding sequence
// everywhere.
//
// Templated on the element precision R: float is the reference's FP32
// element math (P = Vec(X_i - X_3) in Fem.cpp:453, deltas back to the FP64
// node state in Xpbd.h:212); double is the FP64 element model of the
// adjoint's FP64 mode. Constants are written R(literal): for R = float that
// is the nearest float of the double literal, which for every constant used
// here equals the float literal (checked below), and matches how the CPU
// restatement in oracle/ writes them. Every expression keeps the reference's
// operation order; the kernels are compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

namespace xf {

static_assert((float)0.0001 == 0.0001f && (float)1.0e-22 == 1.0e-22f && (float)0.00000001 == 0.00000001f,
              "R(literal) gives the float literals of the FP32 element math");

template <class R>
__device__ __forceinline__ R dot3(const R* a, const R* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// F = Qi * MatMul(dN, P), dN = e_x, e_y, e_z (Fem.cpp:312-323).
template <class R>
__device__ __forceinline__ void defgrad(const R Qi[3][3], const R P[3][3], R F[3][3]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) F[j][r] = Qi[0][r] * P[0][j] + Qi[1][r] * P[1][j] + Qi[2][r] * P[2][j];
}

// adjugate + determinantFromAdjugate (Fem.cpp:284-306).
template <class R>
__device__ __forceinline__ R adjdet(const R F[3][3], R adj[3][3]) {
    R m00 = F[0][0], m01 = F[0][1], m02 = F[0][2];
    R m10 = F[1][0], m11 = F[1][1], m12 = F[1][2];
    R m20 = F[2][0], m21 = F[2][1], m22 = F[2][2];
    adj[0][0] = +(m11 * m22 - m12 * m21);
    adj[0][1] = -(m01 * m22 - m02 * m21);
    adj[0][2] = +(m01 * m12 - m02 * m11);
    adj[1][0] = -(m10 * m22 - m12 * m20);
    adj[1][1] = +(m00 * m22 - m02 * m20);
    adj[1][2] = -(m00 * m12 - m02 * m10);
    adj[2][0] = +(m10 * m21 - m11 * m20);
    adj[2][1] = -(m00 * m21 - m01 * m20);
    adj[2][2] = +(m00 * m11 - m01 * m10);
    return F[0][0] * adj[0][0] + F[0][1] * adj[1][0] + F[0][2] * adj[2][0];
}

// ApplyEnergyGradient at the single point (Fem.cpp:339-354): g[n] += column n of G*Qi.
template <class R>
__device__ __forceinline__ void apply_grad(const R G[3][3], const R Qi[3][3], R g[4][3]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) g[j][r] = g[j][r] + (G[0][r] * Qi[j][0] + G[1][r] * Qi[j][1] + G[2][r] * Qi[j][2]);
}

// The root node's gradient, -(g0 + g1 + g2) in node order.
template <class R>
__device__ __forceinline__ void root_grad(R g[4][3]) {
#pragma unroll
    for (int r = 0; r < 3; r++) g[3][r] = ((R(0.0) - g[0][r]) - g[1][r]) - g[2][r];
}

// Prefactored I1 (Fem.cpp:164-192): U = sum QQ_i |P_i|^2 + sum QR_k P_i.P_j,
// gradients of the three free nodes and the root.
template <class R>
__device__ __forceinline__ void neo_prefactored(const R QQ[3], const R QR[3], const R P[3][3], R& U, R g[4][3]) {
    R u = R(0.0);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        u = u + QQ[i] * dot3(P[i], P[i]);
#pragma unroll
        for (int r = 0; r < 3; r++) g[i][r] = (R(2.0) * QQ[i]) * P[i][r];
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            u = u + QR[k] * dot3(P[i], P[j]);
#pragma unroll
            for (int r = 0; r < 3; r++) g[i][r] = g[i][r] + QR[k] * P[j][r];
#pragma unroll
            for (int r = 0; r < 3; r++) g[j][r] = g[j][r] + QR[k] * P[i][r];
            ++k;
        }
    }
    root_grad(g);
    U = u;
}

// P_i = X_i - X_3 in the element precision (Fem.cpp:453).
template <class R>
__device__ __forceinline__ void gather_p(const double x[4][3], R P[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) P[i][r] = (R)(x[i][r] - x[3][r]);
}

// EnergyXpbdConstrainSimultaneous<2> without damping, Cramer branch
// (Xpbd.h:123-214); w in FP32 as the reference stores it.
template <class R>
__device__ __forceinline__ void constrain2(double x[4][3], const float w[4], const R U[2], const R g0[4][3],
                                           const R g1[4][3], R dt, R comp0, R comp1) {
    R alpha0 = comp0 / (dt * dt);
    R alpha1 = comp1 / (dt * dt);
    R A0 = R(1.0e-22), A1 = R(1.0e-22), A2 = R(1.0e-22);
#pragma unroll
    for (int n = 0; n < 4; n++) A0 = A0 + w[n] * dot3(g0[n], g0[n]);
    A0 = A0 + R(2.0) * U[0] * alpha0;
    R b0 = R(-2.0) * U[0];
#pragma unroll
    for (int n = 0; n < 4; n++) A1 = A1 + w[n] * dot3(g1[n], g0[n]);
#pragma unroll
    for (int n = 0; n < 4; n++) A2 = A2 + w[n] * dot3(g1[n], g1[n]);
    A2 = A2 + R(2.0) * U[1] * alpha1;
    R b1 = R(-2.0) * U[1];
    R invA00 = R(1.0) / A0;
    R invA11 = R(1.0) / A2;
    R invDet = R(1.0) / fmax(R(0.00000001), R(1.0) - (A1 * invA00) * (A1 * invA11));
    R l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
    R l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
#pragma unroll
    for (int n = 0; n < 4; n++) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            R acc = R(0.0) + l0 * g0[n][r];
            acc = acc + l1 * g1[n][r];
            x[n][r] = x[n][r] + (double)(w[n] * acc);
        }
    }
}

// (J - a)^2 volume term (Fem.cpp:537-547): U1 and the root-completed g1.
template <class R>
__device__ __forceinline__ void volume_term(const R Qi[3][3], const R P[3][3], R a, R& U1, R g1[4][3]) {
    R F[3][3], adj[3][3], GJ[3][3];
    defgrad(Qi, P, F);
    R J = adjdet(F, adj);
    U1 = U1 + R(1.0) * ((J - a) * (J - a));
    R s = R(2.0) * (J - a);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) GJ[c][r] = adj[c][r] * s;
    apply_grad(GJ, Qi, g1);
}

// SolveElementMixed for a linear tet with the simultaneous solve, damping
// off (Fem.cpp:437-564): MixedSel (prefactored I1) or YeohSkinFast (the I1
// polynomial with the 1e-4 floor) and the (J - a)^2 volume term, projected
// by constrain2.
template <class R>
__device__ __forceinline__ void solve_mixed_sim(bool yeohFast, const R C[3], double x[4][3], const float w[4],
                                                const R Qi[3][3], const R QQ[3], const R QR[3], R comp0, R comp1,
                                                R a, R dt) {
    R P[3][3];
    gather_p(x, P);
    R U[2] = {R(0.0), R(0.0)};
    R g0[4][3], g1[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            g0[n][r] = R(0.0);
            g1[n][r] = R(0.0);
        }
    neo_prefactored(QQ, QR, P, U[0], g0);
    if (yeohFast) {
        R IM = U[0] - R(3.0);
        U[0] = fmax(R(0.0001), C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM);
        R gScale = C[0] + R(2.0) * C[1] * IM + R(3.0) * C[2] * IM * IM;
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) g0[n][r] = g0[n][r] * gScale;
    }
    volume_term(Qi, P, a, U[1], g1);
    root_grad(g1);
    constrain2(x, w, U, g0, g1, dt, comp0, comp1);
}

}  // namespace xf
