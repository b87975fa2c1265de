// xf_kernels.hip -- gfx950 kernels of the small-step XPBD substep.
//
// One tetrahedron per lane. Tets of one colour share no node, so a colour
// is swept by one launch with plain (non-atomic) gathers and scatters;
// colours run in order, which equals the reference's serial Gauss-Seidel
// sweep (GeoLinear3d::Constrain, Geo.cpp:774-788) over the colour-major
// order. Node kernels fuse pin + collisions + velocity with the next
// substep's prediction (Geo.cpp:305-344).
//
// Node state is FP64 and element arithmetic FP32, as in the reference
// (dvec3 X, O, V in Geo.h:61-64; P = Vec(X_i - X_3) in Fem.cpp:453; deltas
// added back as Dvec in Xpbd.h:212), in the reference's operation order,
// compiled with -ffp-contract=off and IEEE division/sqrt: a serial sweep of
// the same colour order on the CPU reproduces the result bit-for-bit.
//
// The file is compiled twice. As itself it holds every kernel and launcher
// of this part of the library. xf_blocks_contract.hip includes it inside
// namespace xf_contract with XF_BLOCKS_ONLY defined, compiled with
// -ffp-contract=fast: that copy keeps only the sweep-block launcher, named
// XF_BLOCKS_LAUNCH, whose kernels then fuse multiplies and adds (v_fma) and
// no longer round like the reference.
#include <hip/hip_runtime.h>

#include "xf_device.h"
#include "xf_element.h"
#include "xf_sched.h"

#ifndef XF_BLOCKS_LAUNCH
#define XF_BLOCKS_LAUNCH xf_launch_blocks
#endif

using xf::colour_grid;
using xf::colour_slot;

namespace {

using xf::adjdet;
using xf::apply_grad;
using xf::constrain2;
using xf::defgrad;
using xf::dot3;
using xf::gather_p;
using xf::neo_prefactored;
using xf::root_grad;
using xf::volume_term;

// I1 term by one-point quadrature (energies 3/5/6, Fem.cpp:461-477).
__device__ __forceinline__ void i1_quad_term(bool yeoh, const float C[3], const float Qi[3][3], const float P[3][3],
                                             float& U0, float g0[4][3]) {
    float F[3][3], G[3][3];
    defgrad(Qi, P, F);
    float I1 = dot3(F[0], F[0]) + dot3(F[1], F[1]) + dot3(F[2], F[2]);
    float s;
    if (!yeoh) {
        U0 = U0 + 1.0f * I1;
        s = 1.0f * 2.0f;
    } else {
        float IM = I1 - 3.0f;
        U0 = U0 + 1.0f * (C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM);
        s = (1.0f * (C[0] + 2.0f * C[1] * IM + 3.0f * C[2] * IM * IM)) * 2.0f;
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) G[c][r] = F[r][c] * s;
    apply_grad(G, Qi, g0);
}

// EnergyXpbdConstrain without damping (Xpbd.h:110-119).
__device__ __forceinline__ void constrain1(double x[4][3], const float w[4], float U, const float g[4][3], float dt,
                                           float comp) {
    float alpha = comp / (dt * dt);
    float wgg = 1.0e-22f;
#pragma unroll
    for (int n = 0; n < 4; n++) wgg = wgg + w[n] * dot3(g[n], g[n]);
    float lp = (-2.0f * U) / (wgg + (2.0f * U * alpha));
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float s = w[n] * lp;
#pragma unroll
        for (int r = 0; r < 3; r++) x[n][r] = x[n][r] + (double)(s * g[n][r]);
    }
}

// EnergyXpbdConstrain with in-projection Rayleigh damping (Xpbd.h:87-120):
// rtype Rayleigh_Paper (0) or Rayleigh_Limit (1) damp along the substep
// displacement x - o; without damping this is constrain1.
__device__ __forceinline__ void constrain1_damped(double x[4][3], const double o[4][3], const float w[4], float U,
                                                  const float g[4][3], float dt, float comp, float damp, uint32_t rtype) {
    if (!(damp > 0.0f && rtype < 2u)) {
        constrain1(x, w, U, g, dt, comp);
        return;
    }
    float alpha = comp / (dt * dt);
    float gamma = damp / dt;
    float wgg = 1.0e-22f, gV = 0.0f;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float v[3] = {(float)(x[n][0] - o[n][0]), (float)(x[n][1] - o[n][1]), (float)(x[n][2] - o[n][2])};
        wgg = wgg + w[n] * dot3(g[n], g[n]);
        gV = gV + dot3(v, g[n]);
    }
    float lp;
    if (rtype == 0u) {
        float A = wgg * (1.0f + gamma) + (2.0f * U * alpha);
        lp = (-2.0f * U - gamma * gV) / A;
    } else {
        lp = (-2.0f * U) / (wgg + (2.0f * U * alpha));
        float invBeta = comp / (dt * damp);
        float A = wgg + 2.0f * U * invBeta;
        float b = -gV - wgg * lp;
        b = b * (A / fmaxf(A, 4.0f * wgg));
        lp = lp + b / A;
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float s = w[n] * lp;
#pragma unroll
        for (int r = 0; r < 3; r++) x[n][r] = x[n][r] + (double)(s * g[n][r]);
    }
}

// EnergyXpbdConstrainSimultaneous<2> with in-projection Rayleigh damping,
// gamma = damp / dt on both constraints (Xpbd.h:123-214).
__device__ __forceinline__ void constrain2_damped(double x[4][3], const double o[4][3], const float w[4],
                                                  const float U[2], const float g0[4][3], const float g1[4][3],
                                                  float dt, float comp0, float comp1, float damp, uint32_t rtype) {
    if (!(damp > 0.0f && rtype < 2u)) {
        constrain2(x, w, U, g0, g1, dt, comp0, comp1);
        return;
    }
    float comp[2] = {comp0, comp1};
    float alpha[2] = {comp0 / (dt * dt), comp1 / (dt * dt)};
    float gamma = damp / dt;
    float gV[2] = {0.0f, 0.0f};
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float v[3] = {(float)(x[n][0] - o[n][0]), (float)(x[n][1] - o[n][1]), (float)(x[n][2] - o[n][2])};
        gV[0] = gV[0] + dot3(g0[n], v);
        gV[1] = gV[1] + dot3(g1[n], v);
    }
    float wgg[3] = {1.0e-22f, 1.0e-22f, 1.0e-22f};
#pragma unroll
    for (int n = 0; n < 4; n++) wgg[0] = wgg[0] + w[n] * dot3(g0[n], g0[n]);
#pragma unroll
    for (int n = 0; n < 4; n++) wgg[1] = wgg[1] + w[n] * dot3(g1[n], g0[n]);
#pragma unroll
    for (int n = 0; n < 4; n++) wgg[2] = wgg[2] + w[n] * dot3(g1[n], g1[n]);
    float A[3] = {wgg[0], wgg[1], wgg[2]}, b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int d = 2 * i;
        if (rtype == 0u && gamma > 0.0f) {
            A[d] = A[d] + (2.0f * U[i] * alpha[i]) / (1.0f + gamma);
            b[i] = (-2.0f * U[i] - gamma * gV[i]) / (1.0f + gamma);
        } else {
            A[d] = A[d] + 2.0f * U[i] * alpha[i];
            b[i] = -2.0f * U[i];
        }
    }
    float invA00 = 1.0f / A[0];
    float invA11 = 1.0f / A[2];
    float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A[1] * invA00) * (A[1] * invA11));
    float l0 = invDet * ((b[0] * invA00) - (A[1] * invA00) * (b[1] * invA11));
    float l1 = invDet * ((b[1] * invA11) - (b[0] * invA00) * (A[1] * invA11));
    if (rtype == 1u && gamma > 0.0f) {
        float lam[2] = {l0, l1};
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int d = 2 * i;
            float invBeta = comp[i] / (dt * damp);
            A[d] = wgg[d] + 2.0f * U[i] * invBeta;
            b[i] = b[i] * (-1.0f - alpha[i] * lam[i]);
            b[i] = b[i] + -gV[i];
            b[i] = b[i] * (A[d] / fmaxf(A[d], 8.0f * wgg[d]));
        }
        invA00 = 1.0f / A[0];
        invA11 = 1.0f / A[2];
        invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A[1] * invA00) * (A[1] * invA11));
        l0 = l0 + invDet * ((b[0] * invA00) - (A[1] * invA00) * (b[1] * invA11));
        l1 = l1 + invDet * ((b[1] * invA11) - (b[0] * invA00) * (A[1] * invA11));
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            float acc = 0.0f + l0 * g0[n][r];
            acc = acc + l1 * g1[n][r];
            x[n][r] = x[n][r] + (double)(w[n] * acc);
        }
    }
}

// SolveElementMixed for a linear tet (Fem.cpp:437-564), damping off.
// SolveElementPixar's energy and root-completed gradient for a linear tet
// (Fem.cpp:362-432; one quadrature point of weight 1, so Full and Reduced
// coincide): mu = 2, lambda = 2 / (1 - 2 nu) and a = 1 + mu / lambda come in
// the per-tet material slots. SEL: PixarOptions::Selective (the J term at the
// point plus the prefactored I1 term).
template <bool SEL>
__device__ __forceinline__ void pixar_ug(const float Qi[3][3], const float QQ[3], const float QR[3],
                                         const float P[3][3], float lambda, float a, float& U, float g[4][3]) {
    const float mu = 2.0f, weight = 1.0f;
    float F[3][3], adj[3][3], G[3][3];
    defgrad(Qi, P, F);
    float u = 0.0f;
    if constexpr (!SEL) {
        float I1 = dot3(F[0], F[0]) + dot3(F[1], F[1]) + dot3(F[2], F[2]);
        u = u + weight * (mu / 2.0f) * I1;
    }
    float J = adjdet(F, adj);
    u = u + weight * (lambda / 2.0f) * ((J - a) * (J - a));
    float sJ = weight * (lambda / 2.0f) * 2.0f * (J - a);
    float sI = weight * (mu / 2.0f) * 2.0f;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) G[c][r] = SEL ? sJ * adj[c][r] : sI * F[r][c] + sJ * adj[c][r];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) g[n][r] = 0.0f;
    apply_grad(G, Qi, g);
    root_grad(g);
    if constexpr (SEL) {
        float UI1, gI1[4][3];
        neo_prefactored(QQ, QR, P, UI1, gI1);
        u = u + (mu / 2.0f) * UI1;
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) g[n][r] = g[n][r] + (mu / 2.0f) * gI1[n][r];
    }
    U = u;
}

// E: 3 Mixed, 4 MixedSel, 5 YeohSkin, 6 YeohSkinSel, 7 YeohSkinFast,
//    0 = MixedSel or YeohSkinFast chosen per tet by `yeohFast`,
//    20 = Pixar / PixarReduced, 22 = PixarSel (comp0 = compliance / V,
//    comp1 = lambda, a from the Pixar material).
// DAMP: in-projection Rayleigh damping of the body (damp, rtype) along x - o.
template <int E, bool SIM, bool DAMP = false>
__device__ __forceinline__ void solve_tet(bool yeohFast, const float C[3], double x[4][3], const float w[4],
                                          const float Qi[3][3], const float QQ[3], const float QR[3], float comp0,
                                          float comp1, float a, float dt, const double (*o)[3] = nullptr,
                                          float damp = 0.0f, uint32_t rtype = 0u) {
    float P[3][3];
    gather_p(x, P);
    float U[2] = {0.0f, 0.0f};
    float g0[4][3], g1[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) { g0[n][r] = 0.0f; g1[n][r] = 0.0f; }

    if constexpr (E == 20 || E == 22) {
        pixar_ug<E == 22>(Qi, QQ, QR, P, comp1, a, U[0], g0);
        if constexpr (DAMP) constrain1_damped(x, o, w, U[0], g0, dt, comp0, damp, rtype);
        else constrain1(x, w, U[0], g0, dt, comp0);
    } else if constexpr (E == 3 || E == 5 || E == 6) {
        i1_quad_term(E != 3, C, Qi, P, U[0], g0);
        if constexpr (SIM && E != 6) volume_term(Qi, P, a, U[1], g1);
        U[0] = fmaxf(0.0001f, U[0]);
        if constexpr (SIM) {
            if constexpr (E == 6) volume_term(Qi, P, a, U[1], g1);
            root_grad(g0);
            root_grad(g1);
            if constexpr (DAMP) constrain2_damped(x, o, w, U, g0, g1, dt, comp0, comp1, damp, rtype);
            else constrain2(x, w, U, g0, g1, dt, comp0, comp1);
        } else {
            root_grad(g0);
            if constexpr (DAMP) constrain1_damped(x, o, w, U[0], g0, dt, comp0, damp, rtype);
            else constrain1(x, w, U[0], g0, dt, comp0);
            gather_p(x, P);
            volume_term(Qi, P, a, U[1], g1);
            root_grad(g1);
            if constexpr (DAMP) constrain1_damped(x, o, w, U[1], g1, dt, comp1, damp, rtype);
            else constrain1(x, w, U[1], g1, dt, comp1);
        }
    } else {
        neo_prefactored(QQ, QR, P, U[0], g0);
        bool yf = (E == 7) || (E == 0 && yeohFast);
        if (yf) {
            float IM = U[0] - 3.0f;
            U[0] = fmaxf(0.0001f, C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM);
            float gScale = C[0] + 2.0f * C[1] * IM + 3.0f * C[2] * IM * IM;
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int r = 0; r < 3; r++) g0[n][r] = g0[n][r] * gScale;
        }
        if constexpr (!SIM) {
            if constexpr (DAMP) constrain1_damped(x, o, w, U[0], g0, dt, comp0, damp, rtype);
            else constrain1(x, w, U[0], g0, dt, comp0);
            gather_p(x, P);
        }
        volume_term(Qi, P, a, U[1], g1);
        root_grad(g1);
        if constexpr (!SIM) {
            if constexpr (DAMP) constrain1_damped(x, o, w, U[1], g1, dt, comp1, damp, rtype);
            else constrain1(x, w, U[1], g1, dt, comp1);
        } else {
            if constexpr (DAMP) constrain2_damped(x, o, w, U, g0, g1, dt, comp0, comp1, damp, rtype);
            else constrain2(x, w, U, g0, g1, dt, comp0, comp1);
        }
    }
}

// SolveVolumeOnly for tet i of the sweep order (Fem.cpp:840-867): the
// (J - 1)^2 term alone with compliance = settings.compliance * volume (taux.w)
// and a scalar XPBD update. Bodies with fewer than pass + 1 volume passes
// skip the tet.
__device__ __forceinline__ void volume_one(uint32_t i, uint32_t pass, const uint4* __restrict__ tidx,
                                           const float4* __restrict__ tg0, const float4* __restrict__ tg1,
                                           const float4* __restrict__ tg2, const float4* __restrict__ taux,
                                           double4* __restrict__ X, const XfBodyParams* __restrict__ bodies,
                                           float dt) {
    float4 ax = taux[i];
    if (pass >= bodies[__float_as_uint(ax.z)].volume_passes) return;
    uint4 id = tidx[i];
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    float Qi[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    float P[3][3];
    gather_p(x, P);
    float U = 0.0f;
    float g[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) g[n][r] = 0.0f;
    volume_term(Qi, P, 1.0f, U, g);
    root_grad(g);
    constrain1(x, w, U, g, dt, ax.w);
    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
}

// U and root-completed gradients of SolveElementMixed in DampingMode::On
// (Fem.cpp:437-564 with the solve forced to Simultaneous), no projection.
template <int E>
__device__ __forceinline__ void element_ug(bool yeohFast, const float C[3], const double x[4][3], const float Qi[3][3],
                                           const float QQ[3], const float QR[3], float lambda, float a, float U[2],
                                           float g0[4][3], float g1[4][3]) {
    float P[3][3];
    gather_p(x, P);
    U[0] = 0.0f;
    U[1] = 0.0f;
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) { g0[n][r] = 0.0f; g1[n][r] = 0.0f; }
    if constexpr (E == 20 || E == 22) {
        pixar_ug<E == 22>(Qi, QQ, QR, P, lambda, a, U[0], g0);
    } else if constexpr (E == 3 || E == 5 || E == 6) {
        i1_quad_term(E != 3, C, Qi, P, U[0], g0);
        if constexpr (E != 6) volume_term(Qi, P, a, U[1], g1);
        U[0] = fmaxf(0.0001f, U[0]);
        if constexpr (E == 6) volume_term(Qi, P, a, U[1], g1);
        root_grad(g0);
        root_grad(g1);
    } else {
        neo_prefactored(QQ, QR, P, U[0], g0);
        if ((E == 7) || (E == 0 && yeohFast)) {
            float IM = U[0] - 3.0f;
            U[0] = fmaxf(0.0001f, C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM);
            float gScale = C[0] + 2.0f * C[1] * IM + 3.0f * C[2] * IM * IM;
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int r = 0; r < 3; r++) g0[n][r] = g0[n][r] * gScale;
        }
        volume_term(Qi, P, a, U[1], g1);
        root_grad(g1);
    }
}

// RayleighDamp, one constraint (Xpbd.h:216-230).
__device__ __forceinline__ void rayleigh1(double v[4][3], const float w[4], float U, const float g[4][3], float dt,
                                          float comp, float damp) {
    float invBeta = comp / (dt * damp);
    float wgg = 1.0e-22f, gV = 0.0f;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float vf[3] = {(float)v[n][0], (float)v[n][1], (float)v[n][2]};
        wgg = wgg + w[n] * dot3(g[n], g[n]);
        gV = gV + dot3(g[n], vf);
    }
    float lambda = -gV / (2.0f * U * invBeta + wgg);
#pragma unroll
    for (int n = 0; n < 4; n++) {
        float s = w[n] * lambda;
#pragma unroll
        for (int r = 0; r < 3; r++) v[n][r] = v[n][r] + (double)(s * g[n][r]);
    }
}

// RayleighDamp, two coupled constraints (Xpbd.h:232-263).
__device__ __forceinline__ void rayleigh2(double v[4][3], const float w[4], const float U[2], const float g0[4][3],
                                          const float g1[4][3], float dt, float comp0, float comp1, float damp) {
    float invBeta0 = comp0 / (dt * damp), invBeta1 = comp1 / (dt * damp);
    float A0 = 1.0e-22f, A1 = 1.0e-22f, A2 = 1.0e-22f, b0 = 0.0f, b1 = 0.0f;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        A0 = A0 + w[n] * dot3(g0[n], g0[n]);
        A1 = A1 + w[n] * dot3(g0[n], g1[n]);
        A2 = A2 + w[n] * dot3(g1[n], g1[n]);
        float vf[3] = {(float)v[n][0], (float)v[n][1], (float)v[n][2]};
        b0 = b0 - dot3(g0[n], vf);
        b1 = b1 - dot3(g1[n], vf);
    }
    A0 = A0 + 2.0f * U[0] * invBeta0;
    A2 = A2 + 2.0f * U[1] * invBeta1;
    float invA00 = 1.0f / A0;
    float invA11 = 1.0f / A2;
    float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A1 * invA00) * (A1 * invA11));
    float l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
    float l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) v[n][r] = v[n][r] + (double)(w[n] * (l0 * g0[n][r] + l1 * g1[n][r]));
}

// cross(x, y) of vectormath.h:364.
__device__ __forceinline__ void cross3(const float x[3], const float y[3], float o[3]) {
    o[0] = x[1] * y[2] - y[1] * x[2];
    o[1] = x[2] * y[0] - y[2] * x[0];
    o[2] = x[0] * y[1] - y[0] * x[1];
}

// PbdDamp<4> for dvec3 (Xpbd.h:306-348): pull the node velocities toward the
// tet's mass-weighted rigid motion by `damping`; inertia inverted as
// inverse(mat3) does (vectormath.cpp:41-60: adjugate in FP64, det and the
// 1/det scale in FP32).
__device__ __forceinline__ void pbd_damp(const double x[4][3], double v[4][3], const float W[4], float damping) {
    float X[4][3], V[4][3], M[4];
    float Xcm[3] = {0.0f, 0.0f, 0.0f}, Vcm[3] = {0.0f, 0.0f, 0.0f};
    float Msum = 0.0f;
    float Wsum = 1.0e-8f;
#pragma unroll
    for (int i = 0; i < 4; i++) Wsum = Wsum + W[i];
    float Waverage = Wsum * (1.0f / 4.0f);
    float Wmin = 0.0001f * Waverage;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            X[i][r] = (float)(x[i][r] - x[3][r]);
            V[i][r] = (float)v[i][r];
        }
        M[i] = Waverage / fmaxf(Wmin, W[i]);
#pragma unroll
        for (int r = 0; r < 3; r++) Xcm[r] = Xcm[r] + X[i][r] * M[i];
#pragma unroll
        for (int r = 0; r < 3; r++) Vcm[r] = Vcm[r] + V[i][r] * M[i];
        Msum = Msum + M[i];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        Xcm[r] = Xcm[r] / Msum;
        Vcm[r] = Vcm[r] / Msum;
    }
    float rv[4][3], L[3] = {0.0f, 0.0f, 0.0f};
    float I[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int r = 0; r < 3; r++) rv[i][r] = X[i][r] - Xcm[r];
        float c[3];
        cross3(rv[i], V[i], c);
#pragma unroll
        for (int r = 0; r < 3; r++) L[r] = L[r] + M[i] * c[r];
        const float* q = rv[i];
        float rr[3] = {q[0] * q[0], q[1] * q[1], q[2] * q[2]};
        float rp[3] = {q[0] * q[1], q[1] * q[2], q[2] * q[0]};
        float K[3][3] = {{rr[2] + rr[1], -rp[0], -rp[2]}, {-rp[0], rr[2] + rr[0], -rp[1]}, {-rp[2], -rp[1], rr[1] + rr[0]}};
#pragma unroll
        for (int cc = 0; cc < 3; cc++)
#pragma unroll
            for (int r = 0; r < 3; r++) I[cc][r] = I[cc][r] + K[cc][r] * M[i];
    }
    double m00 = I[0][0], m01 = I[0][1], m02 = I[0][2];
    double m10 = I[1][0], m11 = I[1][1], m12 = I[1][2];
    double m20 = I[2][0], m21 = I[2][1], m22 = I[2][2];
    float adj[3][3];
    adj[0][0] = (float)(+(m11 * m22 - m12 * m21));
    adj[0][1] = (float)(-(m01 * m22 - m02 * m21));
    adj[0][2] = (float)(+(m01 * m12 - m02 * m11));
    adj[1][0] = (float)(-(m10 * m22 - m12 * m20));
    adj[1][1] = (float)(+(m00 * m22 - m02 * m20));
    adj[1][2] = (float)(-(m00 * m12 - m02 * m10));
    adj[2][0] = (float)(+(m10 * m21 - m11 * m20));
    adj[2][1] = (float)(-(m00 * m21 - m01 * m20));
    adj[2][2] = (float)(+(m00 * m11 - m01 * m10));
    float det = I[0][0] * adj[0][0] + I[0][1] * adj[1][0] + I[0][2] * adj[2][0];
    float inv = 1.0f / det;
    float om[3];
#pragma unroll
    for (int r = 0; r < 3; r++) om[r] = (adj[0][r] * inv) * L[0] + (adj[1][r] * inv) * L[1] + (adj[2][r] * inv) * L[2];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float c[3];
        cross3(om, rv[i], c);
#pragma unroll
        for (int r = 0; r < 3; r++) v[i][r] = v[i][r] + (double)(damping * ((Vcm[r] + c[r]) - V[i][r]));
    }
}

// Amortized damping (Rayleigh_PostAmortized) covers sweep positions
// [T k / 8, T (k + 1) / 8) of the body, k = tickId % 8 (Geo.cpp:794-798).
__device__ __forceinline__ bool in_damp_range(const XfBodyParams& bp, float2 dm, uint32_t tick) {
    if (bp.rtype != 3u) return true;
    uint32_t k = tick % 8u;
    uint32_t pos = __float_as_uint(dm.y);
    return pos >= bp.n_tets * k / 8u && pos < bp.n_tets * (k + 1u) / 8u;
}

// Rayleigh post damping of tet i (DampElement, Fem.cpp:910-929).
template <int E, bool AUX>
__device__ __forceinline__ void damp_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
                                         const float4* __restrict__ tg1, const float4* __restrict__ tg2,
                                         const float4* __restrict__ tg3, const float4* __restrict__ taux,
                                         const float2* __restrict__ tdmp, const double4* __restrict__ X,
                                         double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
                                         float comp1_u, float a_u, float dt, uint32_t tick) {
    float comp1 = comp1_u, a = a_u;
    uint32_t body = 0;
    if constexpr (AUX) {
        float4 ax = taux[i];
        comp1 = ax.x;
        a = ax.y;
        body = __float_as_uint(ax.z);
    }
    const XfBodyParams& bp = bodies[body];
    if (!(bp.rtype >= 2u && bp.damping > 0.0f) || !in_damp_range(bp, tdmp[i], tick)) return;
    float damp = bp.rtype == 3u ? bp.damping * 8.0f : bp.damping;
    uint4 id = tidx[i];
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i], q3 = tg3[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    double4 va = V[id.x], vb = V[id.y], vc = V[id.z], vd = V[id.w];
    float C[3] = {bp.yeoh[0], bp.yeoh[1], bp.yeoh[2]};
    float Qi[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    float QQ[3] = {q2.y, q2.z, q2.w};
    float QR[3] = {q3.x, q3.y, q3.z};
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    double v[4][3] = {{va.x, va.y, va.z}, {vb.x, vb.y, vb.z}, {vc.x, vc.y, vc.z}, {vd.x, vd.y, vd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    float U[2], g0[4][3], g1[4][3];
    element_ug<E>(bp.energy == 7u, C, x, Qi, QQ, QR, comp1, a, U, g0, g1);
    if constexpr (E == 20 || E == 22) {
        rayleigh1(v, w, U[0], g0, dt, q3.w, damp);  // SolveElementPixar in DampingMode::On
    } else if (bp.simultaneous) {
        rayleigh2(v, w, U, g0, g1, dt, q3.w, comp1, damp);
    } else {
        rayleigh1(v, w, U[0], g0, dt, q3.w, damp);
        rayleigh1(v, w, U[1], g1, dt, comp1, damp);
    }
    V[id.x] = make_double4(v[0][0], v[0][1], v[0][2], 0.0);
    V[id.y] = make_double4(v[1][0], v[1][1], v[1][2], 0.0);
    V[id.z] = make_double4(v[2][0], v[2][1], v[2][2], 0.0);
    V[id.w] = make_double4(v[3][0], v[3][1], v[3][2], 0.0);
}

// PBD damping of tet i (Geo.cpp:806-809), factor min(1, pbd / surfaceArea).
template <bool AUX>
__device__ __forceinline__ void pbd_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ taux,
                                        const float2* __restrict__ tdmp, const double4* __restrict__ X,
                                        double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
                                        uint32_t tick) {
    uint32_t body = 0;
    if constexpr (AUX) body = __float_as_uint(taux[i].z);
    const XfBodyParams& bp = bodies[body];
    float2 dm = tdmp[i];
    if (!(bp.pbd > 0.0f) || !in_damp_range(bp, dm, tick)) return;
    float f = (bp.rtype == 3u ? bp.pbd_amortized : bp.pbd) / dm.x;
    f = f < 1.0f ? f : 1.0f;
    uint4 id = tidx[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    double4 va = V[id.x], vb = V[id.y], vc = V[id.z], vd = V[id.w];
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    double v[4][3] = {{va.x, va.y, va.z}, {vb.x, vb.y, vb.z}, {vc.x, vc.y, vc.z}, {vd.x, vd.y, vd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    pbd_damp(x, v, w, f);
    V[id.x] = make_double4(v[0][0], v[0][1], v[0][2], 0.0);
    V[id.y] = make_double4(v[1][0], v[1][1], v[1][2], 0.0);
    V[id.z] = make_double4(v[2][0], v[2][1], v[2][2], 0.0);
    V[id.w] = make_double4(v[3][0], v[3][1], v[3][2], 0.0);
}

}  // namespace

// ---------------------------------------------------------------- kernels

#ifndef XF_BLOCKS_ONLY
// Volume pass `pass` over one colour: tets [begin, begin+count).
__global__ __launch_bounds__(256) void xf_volume_colour(uint32_t begin, uint32_t count, uint32_t slabs, uint32_t pass,
                                                        const uint4* __restrict__ tidx,
                                                        const float4* __restrict__ tg0,
                                                        const float4* __restrict__ tg1,
                                                        const float4* __restrict__ tg2,
                                                        const float4* __restrict__ taux, double4* __restrict__ X,
                                                        const XfBodyParams* __restrict__ bodies, float dt) {
    uint32_t i = colour_slot(begin, count, slabs);
    if (i == xf::kNoSlot) return;
    volume_one(i, pass, tidx, tg0, tg1, tg2, taux, X, bodies, dt);
}
#endif  // XF_BLOCKS_ONLY

// Projects one colour: tets [begin, begin+count) of the sweep order.
// AUX: per-tet (comp1, a, body) from taux; otherwise comp1/a are uniform and
// the body is 0. The first colour launch of a substep also advances the
// simulation clock (clock != nullptr); no other kernel of the colour sweep
// reads it.
// Rayleigh post damping (Geo.cpp:797-802) of one colour.
template <int E, bool AUX>
__global__ __launch_bounds__(256) void xf_damp_colour(uint32_t begin, uint32_t count, uint32_t slabs,
                                                      const uint4* __restrict__ tidx,
                                                      const float4* __restrict__ tg0, const float4* __restrict__ tg1,
                                                      const float4* __restrict__ tg2, const float4* __restrict__ tg3,
                                                      const float4* __restrict__ taux, const float2* __restrict__ tdmp,
                                                      const double4* __restrict__ X, double4* __restrict__ V,
                                                      const XfBodyParams* __restrict__ bodies, float comp1_u,
                                                      float a_u, float dt, const float* __restrict__ clock) {
    if constexpr (E == 30) return;  // no constraint on tets, nothing to damp
    uint32_t i = colour_slot(begin, count, slabs);
    if (i == xf::kNoSlot) return;
    uint32_t tick = reinterpret_cast<const uint32_t*>(clock)[1] - 1u;
    damp_one<E, AUX>(i, tidx, tg0, tg1, tg2, tg3, taux, tdmp, X, V, bodies, comp1_u, a_u, dt, tick);
}

// PBD damping (Geo.cpp:806-809) of one colour.
template <bool AUX>
__global__ __launch_bounds__(256) void xf_pbd_colour(uint32_t begin, uint32_t count, uint32_t slabs,
                                                     const uint4* __restrict__ tidx,
                                                     const float4* __restrict__ taux, const float2* __restrict__ tdmp,
                                                     const double4* __restrict__ X, double4* __restrict__ V,
                                                     const XfBodyParams* __restrict__ bodies,
                                                     const float* __restrict__ clock) {
    uint32_t i = colour_slot(begin, count, slabs);
    if (i == xf::kNoSlot) return;
    uint32_t tick = reinterpret_cast<const uint32_t*>(clock)[1] - 1u;
    pbd_one<AUX>(i, tidx, taux, tdmp, X, V, bodies, tick);
}

// The projection of one tet from its loaded element data (q0..q3: Qi, QQ,
// QR, comp0; ax: comp1, a, body when AUX) on positions x, inverse masses w.
// DAMP: bodies may damp inside the projection (Rayleigh_Paper / _Limit);
// reads O of the 4 nodes (global ids id).
template <int E, bool SIM, bool AUX, bool DAMP>
__device__ __forceinline__ void project_loaded(const float4& q0, const float4& q1, const float4& q2, const float4& q3,
                                               const float4& ax, const uint4& id, double x[4][3], const float w[4],
                                               const double4* __restrict__ O, const XfBodyParams* __restrict__ bodies,
                                               float comp1_u, float a_u, float dt) {
    float comp1 = comp1_u, a = a_u;
    uint32_t body = 0;
    if constexpr (AUX) {
        comp1 = ax.x;
        a = ax.y;
        body = __float_as_uint(ax.z);
    }
    const XfBodyParams& bp = bodies[body];
    float C[3] = {bp.yeoh[0], bp.yeoh[1], bp.yeoh[2]};
    bool yeohFast = bp.energy == 7;

    float Qi[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    float QQ[3] = {q2.y, q2.z, q2.w};
    float QR[3] = {q3.x, q3.y, q3.z};
    if constexpr (DAMP) {
        double4 oa = O[id.x], ob = O[id.y], oc = O[id.z], od = O[id.w];
        double o[4][3] = {{oa.x, oa.y, oa.z}, {ob.x, ob.y, ob.z}, {oc.x, oc.y, oc.z}, {od.x, od.y, od.z}};
        solve_tet<E, SIM, true>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt, o, bp.damping, bp.rtype);
    } else {
        solve_tet<E, SIM>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt);
    }
}

template <int E, bool SIM, bool AUX, bool DAMP>
__global__ __launch_bounds__(256) void xf_project_colour(
    uint32_t begin, uint32_t count, uint32_t slabs, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, double4* __restrict__ X, const double4* __restrict__ O,
    const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u, float dt, float* __restrict__ clock) {
    if (clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    if constexpr (E == 30) return;  // CubeNeo / CubeSkin: no constraint on tets (Fem.cpp:567-760)
    uint32_t i = colour_slot(begin, count, slabs);
    if (i == xf::kNoSlot) return;

    uint4 id = tidx[i];
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i], q3 = tg3[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    float4 ax = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (AUX) ax = taux[i];

    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    project_loaded<E, SIM, AUX, DAMP>(q0, q1, q2, q3, ax, id, x, w, O, bodies, comp1_u, a_u, dt);

    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
}

// Sweep-block projection: workgroup s of one launch colour sweeps block s
// through all of its phases. The block's nodes live in LDS for the whole
// launch; phases are separated by workgroup barriers and the nodes are
// written back at the end. Blocks of one launch colour share no node, so the
// launch equals the serial sweep of its blocks one after another, each phase
// by phase. Each lane holds one tet of the current phase and loads its
// element data for the next phase before solving the current one.
// rec: per block, nph (first schedule position, count) pairs. bnode: per
// block, nmax global node ids (xf::kNoNode past the block's nodes). tloc: per
// schedule position, the tet's four block-local node indices, 16 bits each
// (x: corners 0 | 1 << 16, y: corners 2 | 3 << 16). The first launch colour
// of a substep advances the simulation clock (clock != nullptr).
struct BlockTet {
    float4 q0, q1, q2, q3, ax;
    uint2 loc;
};

namespace {
// End of a substep and the next prediction for node i of body p, whose
// position after the sweep is xi (defined with the node phase below).
__device__ __forceinline__ void finish_predict(double4 xi, double4 oi, uint32_t i, double4* __restrict__ X,
                                               double4* __restrict__ O, const XfBodyParams& p, float t_end);
}  // namespace

template <bool AUX>
__device__ __forceinline__ BlockTet load_block_tet(uint32_t i, const uint2* __restrict__ tloc,
                                                   const float4* __restrict__ tg0, const float4* __restrict__ tg1,
                                                   const float4* __restrict__ tg2, const float4* __restrict__ tg3,
                                                   const float4* __restrict__ taux) {
    BlockTet t;
    t.q0 = tg0[i];
    t.q1 = tg1[i];
    t.q2 = tg2[i];
    t.q3 = tg3[i];
    t.ax = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (AUX) t.ax = taux[i];
    t.loc = tloc[i];
    return t;
}

// A slot's nodes in LDS, node k at byte k * NS: (x, y) in its first 16
// bytes, (z, w) in the next. NS = 32 packs them; NS = 48 spreads their
// 16-byte halves over every bank group of a 128-byte LDS row (with 32 only
// every other group is reached), so the four-node gathers of a phase meet
// fewer bank conflicts, for half as much LDS again.
template <uint32_t NS>
struct NodeLds {
    static_assert(NS == 32u || NS == 48u, "32 or 48 bytes per node");
    double2* base;  // 16-byte units
    __device__ __forceinline__ double2* at(uint32_t k) const { return base + k * (NS / 16u); }
    __device__ __forceinline__ double4 get(uint32_t k) const {
        const double2* p = at(k);
        const double2 a = p[0], b = p[1];
        return make_double4(a.x, a.y, b.x, b.y);
    }
    __device__ __forceinline__ void put(uint32_t k, double4 v) const {
        double2* p = at(k);
        p[0] = make_double2(v.x, v.y);
        p[1] = make_double2(v.z, v.w);
    }
    // position only (w, the inverse mass, is not changed by a projection)
    __device__ __forceinline__ void put_xyz(uint32_t k, double x, double y, double z) const {
        double2* p = at(k);
        p[0] = make_double2(x, y);
        p[1].x = z;
    }
};

// One tet of a block phase: nodes from LDS, element data loaded, solved,
// nodes written back to LDS.
template <int E, bool SIM, bool AUX, bool DAMP, uint32_t NS>
__device__ __forceinline__ void block_tet(const BlockTet& t, uint32_t i, NodeLds<NS> xs, const uint4* __restrict__ tidx,
                                          const double4* __restrict__ O, const XfBodyParams* __restrict__ bodies,
                                          float comp1_u, float a_u, float dt) {
    uint32_t ia = t.loc.x & 0xFFFFu, ib = t.loc.x >> 16, ic = t.loc.y & 0xFFFFu, id = t.loc.y >> 16;
    double4 na = xs.get(ia), nb = xs.get(ib), nc = xs.get(ic), nd = xs.get(id);
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    uint4 gid = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (DAMP) gid = tidx[i];
    project_loaded<E, SIM, AUX, DAMP>(t.q0, t.q1, t.q2, t.q3, t.ax, gid, x, w, O, bodies, comp1_u, a_u, dt);
    xs.put_xyz(ia, x[0][0], x[0][1], x[0][2]);
    xs.put_xyz(ib, x[1][0], x[1][1], x[1][2]);
    xs.put_xyz(ic, x[2][0], x[2][1], x[2][2]);
    xs.put_xyz(id, x[3][0], x[3][1], x[3][2]);
}

// Value of v in lane L of this lane's quad (lanes 4k .. 4k + 3), by DPP
// quad_perm broadcast.
template <int L>
__device__ __forceinline__ float quad_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), L * 0x55, 0xF, 0xF, false));
}

// One tet of a block phase on the four lanes of a quad: lane q owns node q.
// SolveElementMixed with the simultaneous solve (MixedSel / YeohSkinFast;
// E = 0 picks per body), damping off, in exactly the operations and order of
// solve_tet: the per-tet scalars (I1, F, adjugate, J, the Cramer solve) are
// computed by all four lanes alike, each node's gradients, its terms of the
// sums of the 2x2 system and its position update by its own lane; the root
// node's gradient -(g0 + g1 + g2) and the sums over nodes are assembled in
// node order from the other lanes by DPP. Results equal block_tet bit for
// bit.
template <int E, bool AUX>
__device__ __forceinline__ void block_tet_quad(const BlockTet& t, uint32_t q, NodeLds<32> xs,
                                               const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u,
                                               float dt) {
    static_assert(E == 0 || E == 4 || E == 7, "quad lanes take MixedSel / YeohSkinFast");
    const uint32_t ia = t.loc.x & 0xFFFFu, ib = t.loc.x >> 16, ic = t.loc.y & 0xFFFFu, id = t.loc.y >> 16;
    const double4 na = xs.get(ia), nb = xs.get(ib), nc = xs.get(ic), nd = xs.get(id);
    const double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    const bool q0 = q == 0u, q1 = q == 1u, q2 = q == 2u, q3 = q == 3u;
    const uint32_t own = q0 ? ia : (q1 ? ib : (q2 ? ic : id));
    const double4 xo = q0 ? na : (q1 ? nb : (q2 ? nc : nd));
    const float w = (float)xo.w;

    float comp1 = comp1_u, a = a_u;
    uint32_t body = 0;
    if constexpr (AUX) {
        comp1 = t.ax.x;
        a = t.ax.y;
        body = __float_as_uint(t.ax.z);
    }
    const XfBodyParams& bp = bodies[body];
    const float C[3] = {bp.yeoh[0], bp.yeoh[1], bp.yeoh[2]};
    const bool yf = (E == 7) || (E == 0 && bp.energy == 7);
    const float Qi[3][3] = {{t.q0.x, t.q0.y, t.q0.z}, {t.q0.w, t.q1.x, t.q1.y}, {t.q1.z, t.q1.w, t.q2.x}};
    const float QQ[3] = {t.q2.y, t.q2.z, t.q2.w};
    const float QR[3] = {t.q3.x, t.q3.y, t.q3.z};
    const float comp0 = t.q3.w;

    float P[3][3];
    gather_p(x, P);

    // Prefactored I1 (neo_prefactored): the energy on every lane, the
    // gradient of node q on lane q, g[i] = (2 QQ_i) P_i + QR_a P_a + QR_b P_b
    // in the pair order (0,1), (0,2), (1,2).
    float u = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) u = u + QQ[i] * dot3(P[i], P[i]);
    {
        int k = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i + 1; j < 3; j++) {
                u = u + QR[k] * dot3(P[i], P[j]);
                ++k;
            }
    }
    const float cq = 2.0f * (q0 ? QQ[0] : (q1 ? QQ[1] : QQ[2]));
    const float qa = q2 ? QR[1] : QR[0];
    const float qb = q0 ? QR[1] : QR[2];
    float g0[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float pn = q0 ? P[0][r] : (q1 ? P[1][r] : P[2][r]);
        const float pa = q0 ? P[1][r] : P[0][r];
        const float pb = q2 ? P[1][r] : P[2][r];
        float g = cq * pn;
        g = g + qa * pa;
        g = g + qb * pb;
        g0[r] = g;
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float root = ((0.0f - quad_lane<0>(g0[r])) - quad_lane<1>(g0[r])) - quad_lane<2>(g0[r]);
        g0[r] = q3 ? root : g0[r];
    }
    float U0 = u;
    if (yf) {
        const float IM = U0 - 3.0f;
        U0 = fmaxf(0.0001f, C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM);
        const float gScale = C[0] + 2.0f * C[1] * IM + 3.0f * C[2] * IM * IM;
#pragma unroll
        for (int r = 0; r < 3; r++) g0[r] = g0[r] * gScale;
    }

    // (J - a)^2 volume term (volume_term): F, adjugate and J on every lane,
    // node q's row of G Qi^T on lane q.
    float F[3][3], adj[3][3], GJ[3][3];
    defgrad(Qi, P, F);
    const float J = adjdet(F, adj);
    const float U1 = 0.0f + 1.0f * ((J - a) * (J - a));
    const float sJ = 2.0f * (J - a);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) GJ[c][r] = adj[c][r] * sJ;
    float Qn[3];
#pragma unroll
    for (int c = 0; c < 3; c++) Qn[c] = q0 ? Qi[0][c] : (q1 ? Qi[1][c] : Qi[2][c]);
    float g1[3];
#pragma unroll
    for (int r = 0; r < 3; r++) g1[r] = 0.0f + (GJ[0][r] * Qn[0] + GJ[1][r] * Qn[1] + GJ[2][r] * Qn[2]);
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float root = ((0.0f - quad_lane<0>(g1[r])) - quad_lane<1>(g1[r])) - quad_lane<2>(g1[r]);
        g1[r] = q3 ? root : g1[r];
    }

    // EnergyXpbdConstrainSimultaneous<2>, Cramer branch (constrain2): the
    // node sums in node order, then the solve on every lane.
    const float alpha0 = comp0 / (dt * dt);
    const float alpha1 = comp1 / (dt * dt);
    const float t00 = w * dot3(g0, g0), t10 = w * dot3(g1, g0), t11 = w * dot3(g1, g1);
    float A0 = 1.0e-22f, A1 = 1.0e-22f, A2 = 1.0e-22f;
    A0 = A0 + quad_lane<0>(t00);
    A0 = A0 + quad_lane<1>(t00);
    A0 = A0 + quad_lane<2>(t00);
    A0 = A0 + quad_lane<3>(t00);
    A0 = A0 + 2.0f * U0 * alpha0;
    const float b0 = -2.0f * U0;
    A1 = A1 + quad_lane<0>(t10);
    A1 = A1 + quad_lane<1>(t10);
    A1 = A1 + quad_lane<2>(t10);
    A1 = A1 + quad_lane<3>(t10);
    A2 = A2 + quad_lane<0>(t11);
    A2 = A2 + quad_lane<1>(t11);
    A2 = A2 + quad_lane<2>(t11);
    A2 = A2 + quad_lane<3>(t11);
    A2 = A2 + 2.0f * U1 * alpha1;
    const float b1 = -2.0f * U1;
    const float invA00 = 1.0f / A0;
    const float invA11 = 1.0f / A2;
    const float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A1 * invA00) * (A1 * invA11));
    const float l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
    const float l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
    double xn[3] = {xo.x, xo.y, xo.z};
#pragma unroll
    for (int r = 0; r < 3; r++) {
        float acc = 0.0f + l0 * g0[r];
        acc = acc + l1 * g1[r];
        xn[r] = xn[r] + (double)(w * acc);
    }
    xs.put_xyz(own, xn[0], xn[1], xn[2]);
}

// A slot's nodes into LDS (stage_nodes) and back (write_nodes), kStage list
// entries per lane at a time: the batch's id loads are issued together,
// then its node loads (at clamped in-range positions, so that no branch
// separates them), so a batch costs two memory round trips whatever its
// size, where one entry per iteration waits for each id before its node.
constexpr uint32_t kStage = 4;

template <uint32_t NS>
__device__ __forceinline__ void stage_nodes(uint32_t tid, uint32_t bs, uint32_t n, const uint32_t* __restrict__ bn,
                                            const double4* __restrict__ X, NodeLds<NS> xs) {
    for (uint32_t k0 = tid; k0 < n; k0 += kStage * bs) {
        uint32_t g[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; u++) g[u] = bn[min(k0 + u * bs, n - 1u)];
        double4 v[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; u++) v[u] = X[g[u] != xf::kNoNode ? g[u] : 0u];
#pragma unroll
        for (uint32_t u = 0; u < kStage; u++) {
            const uint32_t k = k0 + u * bs;
            if (k < n && g[u] != xf::kNoNode) xs.put(k, v[u]);
        }
    }
}

template <uint32_t NS>
__device__ __forceinline__ void write_nodes(uint32_t tid, uint32_t bs, uint32_t n, const uint32_t* __restrict__ bn,
                                            double4* __restrict__ X, NodeLds<NS> xs) {
    for (uint32_t k0 = tid; k0 < n; k0 += kStage * bs) {
        uint32_t g[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; u++) g[u] = bn[min(k0 + u * bs, n - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < kStage; u++) {
            const uint32_t k = k0 + u * bs;
            if (k < n && g[u] != xf::kNoNode) X[g[u]] = xs.get(k);
        }
    }
}

// ntet: tets in the schedule. Element loads are issued by every lane for
// every phase pair, at clamped in-range positions (lanes without a tet, and
// the phase after the last, load data they do not use), so that the loads in
// flight are the same on every path and the wait before a phase covers
// exactly that phase's loads. OVER: some block has more tets in a phase than
// the workgroup has lanes (lanes then load and solve further tets in turn).
// FIN: the launch is the substep's last; its workgroups end the substep and
// predict the next one for every node of the scene (the node phase, fused).
// fextra holds per workgroup nmax + fmax node ids (xf::kNoNode padded): its
// block nodes as in bnode, then fmax nodes that no block of this launch
// colour touches. All are staged in LDS with their O and finished from there
// after the last phase. No other workgroup of the launch reads or writes
// those nodes, and O is not read by the projection (FIN excludes DAMP).
// tend: the clock, read for the end time of the substep (advanced by an
// earlier launch). fsingle: the scene has one body.
template <int E, bool SIM, bool AUX, bool DAMP, bool OVER, bool FIN, bool QUAD = false, uint32_t NS = 32>
__global__ __launch_bounds__(256) void xf_project_blocks(
    uint32_t nph, uint32_t nmax, uint32_t ntet, const uint32_t* __restrict__ rec, const uint32_t* __restrict__ bnode,
    const uint2* __restrict__ tloc, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, double4* __restrict__ X, double4* __restrict__ O,
    const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u, float dt, float* __restrict__ clock,
    const uint32_t* __restrict__ fextra, uint32_t fmax, const float* __restrict__ tend, uint32_t fsingle) {
    static_assert(!(FIN && (DAMP || E == 30)), "the fused node phase needs a projection that leaves O unread");
    static_assert(!QUAD || (SIM && !DAMP && (E == 0 || E == 4 || E == 7)),
                  "quad lanes take the simultaneous MixedSel / YeohSkinFast projection without damping");
    static_assert(NS == 32u || !QUAD, "quad lanes take 32-byte nodes");
    if (clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    if constexpr (E == 30) return;  // CubeNeo / CubeSkin: no constraint on tets (Fem.cpp:567-760)
    extern __shared__ double4 xf_block_nodes[];
    const NodeLds<NS> xs{reinterpret_cast<double2*>(xf_block_nodes)};
    const uint32_t tid = threadIdx.x, bs = blockDim.x, last = ntet - 1u;
    const uint32_t* r = rec + (size_t)blockIdx.x * 2 * nph;
    const uint32_t* bn = bnode + (size_t)blockIdx.x * nmax;
    const uint32_t jl = nph - 1u;
    // FIN: LDS holds ftot nodes (the block nodes, then the extra nodes, NS
    // bytes each), then their O, then their ids
    const uint32_t ftot = FIN ? nmax + fmax : 0u;
    double4* os = FIN ? reinterpret_cast<double4*>(xs.base + (NS / 16u) * ftot) : nullptr;
    uint32_t* fs = FIN ? reinterpret_cast<uint32_t*>(os + ftot) : nullptr;
    // QUAD: four lanes per tet (lane tid & 3 owns node tid & 3), bs / 4 tets
    // at a time; else one lane per tet
    const uint32_t lt = QUAD ? tid >> 2 : tid, per = QUAD ? bs >> 2 : bs;
    auto load = [&](uint32_t j) {
        return load_block_tet<AUX>(min(r[2 * j] + lt, last), tloc, tg0, tg1, tg2, tg3, taux);
    };
    auto solve = [&](const BlockTet& t, uint32_t i) {
        if constexpr (QUAD) block_tet_quad<E, AUX>(t, tid & 3u, xs, bodies, comp1_u, a_u, dt);
        else block_tet<E, SIM, AUX, DAMP, NS>(t, i, xs, tidx, O, bodies, comp1_u, a_u, dt);
    };
    // each lane's first tet of the phase is t; with OVER, lanes go on to the
    // tets beyond the workgroup's lanes
    auto phase = [&](uint32_t j, const BlockTet& t) {
        const uint32_t b = r[2 * j], cnt = r[2 * j + 1];
        if (lt < cnt) solve(t, b + lt);
        if constexpr (OVER)
            for (uint32_t k = lt + per; k < cnt; k += per)
                solve(load_block_tet<AUX>(b + k, tloc, tg0, tg1, tg2, tg3, taux), b + k);
        __syncthreads();
    };
    BlockTet ta = load(0), tb;
    if constexpr (FIN) {
        // nodes, their O and their ids staged for the end of the substep
        const uint32_t* fl = fextra + (size_t)blockIdx.x * ftot;
        for (uint32_t k = tid; k < ftot; k += bs) {
            uint32_t g = fl[k];
            fs[k] = g;
            if (g != xf::kNoNode) {
                xs.put(k, X[g]);
                os[k] = O[g];
            }
        }
    } else {
        stage_nodes(tid, bs, nmax, bn, X, xs);
    }
    __syncthreads();
    // phases in pairs, element data of the next phase loaded before solving
    // the current one (ta: even phases, tb: odd phases)
    for (uint32_t j = 0; j < nph; j += 2) {
        tb = load(min(j + 1u, jl));
        phase(j, ta);
        ta = load(min(j + 2u, jl));
        if (j + 1u < nph) phase(j + 1u, tb);
    }
    if constexpr (FIN) {
        // (single-body scenes read the body's parameters once)
        const float t_end = tend[0];
        auto finish_all = [&](auto body_of) {
            for (uint32_t k = tid; k < ftot; k += bs) {
                uint32_t g = fs[k];
                if (g != xf::kNoNode) {
                    double4 oi = os[k];
                    finish_predict(xs.get(k), oi, g, X, O, body_of((uint32_t)oi.w), t_end);
                }
            }
        };
        if (fsingle)
            finish_all([&](uint32_t) -> const XfBodyParams& { return bodies[0]; });
        else
            finish_all([&](uint32_t flags) -> const XfBodyParams& { return bodies[flags >> 1]; });
    } else {
        write_nodes(tid, bs, nmax, bn, X, xs);
    }
}

namespace {

// Collision projection of one unpinned node (no reference counterpart).
// Ground plane y >= ground_y and a kinematic sphere; Coulomb friction on the
// tangential displacement of the substep (static if |d_t| <= mu * depth).
// Corrections are computed in FP32 and added to the FP64 position.
__device__ __forceinline__ void collide_node(const XfBodyParams& p, float t_end, const double o[3], double x[3]) {
    float mu = p.friction;
    if (p.collide & 1u) {
        float gy = p.ground_y;
        float xy = (float)x[1];
        if (xy < gy) {
            float depth = gy - xy;
            x[1] = (double)gy;
            float dx = (float)(x[0] - o[0]);
            float dz = (float)(x[2] - o[2]);
            float t2 = dx * dx + dz * dz;
            float lim = mu * depth;
            if (t2 <= lim * lim) {
                x[0] = o[0];
                x[2] = o[2];
            } else {
                float s = lim / sqrtf(t2);
                x[0] = x[0] - (double)(dx * s);
                x[2] = x[2] - (double)(dz * s);
            }
        }
    }
    if (p.collide & 2u) {
        float c[3], d[3];
#pragma unroll
        for (int r = 0; r < 3; r++) c[r] = p.sc0[r] + p.sv[r] * t_end;
#pragma unroll
        for (int r = 0; r < 3; r++) d[r] = (float)x[r] - c[r];
        float d2 = dot3(d, d);
        float rad = p.sr;
        if (d2 < rad * rad && d2 > 0.0f) {
            float dist = sqrtf(d2);
            float inv = 1.0f / dist;
            float n[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
            float depth = rad - dist;
#pragma unroll
            for (int r = 0; r < 3; r++) x[r] = x[r] + (double)(n[r] * depth);
            float rel[3], rt[3];
#pragma unroll
            for (int r = 0; r < 3; r++) rel[r] = (float)(x[r] - o[r]) - p.sv[r] * p.dt;
            float rn = dot3(rel, n);
#pragma unroll
            for (int r = 0; r < 3; r++) rt[r] = rel[r] - n[r] * rn;
            float t2 = dot3(rt, rt);
            float lim = mu * depth;
            float s = 1.0f;
            if (t2 > lim * lim) s = lim / sqrtf(t2);
#pragma unroll
            for (int r = 0; r < 3; r++) x[r] = x[r] - (double)(rt[r] * s);
        }
    }
}

// Prediction (Geo.cpp:307-312), FP64 state with FP32 settings products.
__device__ __forceinline__ void predict(const XfBodyParams& p, double v[3], const double x[3], double4* Xo,
                                        double4* Oo, double w, double oflags) {
    v[0] = v[0] + (double)p.gdx;
    v[1] = v[1] + (double)p.gdy;
    v[2] = v[2] + (double)0.0f;
    double dragf = (double)p.dragf;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = v[r] * dragf;
    double dt = (double)p.dt;
    *Oo = make_double4(x[0], x[1], x[2], oflags);
    *Xo = make_double4(x[0] + v[0] * dt, x[1] + v[1] * dt, x[2] + v[2] * dt, w);
}

// The lock of Geo3d::Substep (Geo.cpp:318-322) and the contact pass for one
// node: a pinned node of a body whose pins are locked this step returns to
// O and its inverse mass becomes 0 (from then on; it may be finite until its
// first locked substep); other nodes with mass take contact. Returns the
// node's inverse mass after the lock.
__device__ __forceinline__ double lock_contact(const XfBodyParams& p, uint32_t flags, float t_end, double w,
                                               const double o[3], double x[3]) {
    if ((flags & 1u) && !p.unlock) {
        x[0] = o[0];
        x[1] = o[1];
        x[2] = o[2];
        return 0.0;
    }
    if (p.collide && w != 0.0) collide_node(p, t_end, o, x);
    return w;
}

__device__ __forceinline__ void finish_predict(double4 xi, double4 oi, uint32_t i, double4* __restrict__ X,
                                               double4* __restrict__ O, const XfBodyParams& p, float t_end) {
    uint32_t flags = (uint32_t)oi.w;
    double x[3] = {xi.x, xi.y, xi.z};
    double o[3] = {oi.x, oi.y, oi.z};
    const double w = lock_contact(p, flags, t_end, xi.w, o, x);
    double v[3];
    double invdt = (double)p.invdt;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = (x[r] - o[r]) * invdt;
    predict(p, v, x, &X[i], &O[i], w, oi.w);
}

}  // namespace

#ifndef XF_BLOCKS_ONLY
// First prediction of a step call from the stored velocity.
__global__ __launch_bounds__(256) void xf_predict(uint32_t n, double4* __restrict__ X, double4* __restrict__ O,
                                                  const double4* __restrict__ V,
                                                  const XfBodyParams* __restrict__ bodies) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 xi = X[i];
    double4 oi = O[i];
    double4 vi = V[i];
    const XfBodyParams& p = bodies[(uint32_t)oi.w >> 1];
    double v[3] = {vi.x, vi.y, vi.z};
    double x[3] = {xi.x, xi.y, xi.z};
    predict(p, v, x, &X[i], &O[i], xi.w, oi.w);
}

// End of a substep: pin (Geo.cpp:318-322), collisions, velocity
// (Geo.cpp:342-344); then, if PREDICT, the next substep's prediction,
// else the velocity is stored. clock holds the end time of the substep.
// SINGLE: the scene has one body, whose parameters every node reads (loaded
// once per wave, alongside the node loads, instead of through the node's
// body id).
// SNAP: X after the sweep and O are first copied to xs_out / oo_out (the
// adjoint's replay records them for the reverse of the contact pass).
template <bool PREDICT, bool SINGLE, bool SNAP>
__global__ __launch_bounds__(256) void xf_finish(uint32_t n, double4* __restrict__ X, double4* __restrict__ O,
                                                 double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
                                                 const float* __restrict__ clock, double4* __restrict__ xs_out,
                                                 double4* __restrict__ oo_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 xi = X[i];
    double4 oi = O[i];
    if constexpr (SNAP) {
        xs_out[i] = xi;
        oo_out[i] = oi;
    }
    uint32_t flags = (uint32_t)oi.w;
    const XfBodyParams& p = bodies[SINGLE ? 0u : flags >> 1];
    double x[3] = {xi.x, xi.y, xi.z};
    double o[3] = {oi.x, oi.y, oi.z};
    const double w = lock_contact(p, flags, clock[0], xi.w, o, x);
    double v[3];
    double invdt = (double)p.invdt;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = (x[r] - o[r]) * invdt;
    if constexpr (PREDICT) {
        predict(p, v, x, &X[i], &O[i], w, oi.w);
    } else {
        X[i] = make_double4(x[0], x[1], x[2], w);
        V[i] = make_double4(v[0], v[1], v[2], 0.0);
    }
}

// Nodes held at kinematic targets (Settings_LockRight, Geo.cpp:324-331):
// after the sweep, X = O = target and the inverse mass becomes 0.
__global__ __launch_bounds__(256) void xf_kinematic(uint32_t n, const uint32_t* __restrict__ idx,
                                                    const double4* __restrict__ target, double4* __restrict__ X,
                                                    double4* __restrict__ O) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t i = idx[k];
    double4 t = target[k];
    X[i] = make_double4(t.x, t.y, t.z, 0.0);
    O[i] = make_double4(t.x, t.y, t.z, O[i].w);
}

// The manipulation constraint of Geo3d::Substep (Geo.cpp:334-339): m holds
// the target point and, in w, the bits of the node id (xf::kNoNode: none).
// X += (target - X) * w / (max(1e-6, w) + 1.8 / dt^2), in FP32 as there.
__global__ __launch_bounds__(64) void xf_manip(const float4* __restrict__ m, float dt, double4* __restrict__ X) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const float4 t = m[0];
    const uint32_t i = __float_as_uint(t.w);
    if (i == xf::kNoNode) return;
    double4 xi = X[i];
    float w = (float)xi.w;
    float s = w / (fmaxf(0.000001f, w) + 1.8f / (dt * dt));
    float d0 = (t.x - (float)xi.x) * s, d1 = (t.y - (float)xi.y) * s, d2 = (t.z - (float)xi.z) * s;
    X[i] = make_double4(xi.x + (double)d0, xi.y + (double)d1, xi.z + (double)d2, xi.w);
}

// Forward projection of one colour that also records each tet's 4 node
// positions before the projection into tape (12 doubles per schedule
// position), for the adjoint's replay of a checkpoint window.
template <int E, bool AUX>
__global__ __launch_bounds__(256) void xf_project_record(
    uint32_t begin, uint32_t count, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, double4* __restrict__ X, double* __restrict__ tape,
    const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u, float dt, float* __restrict__ clock) {
    if (clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    uint32_t i = begin + k;
    uint4 id = tidx[i];
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i], q3 = tg3[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    float comp1 = comp1_u, a = a_u;
    uint32_t body = 0;
    if constexpr (AUX) {
        float4 ax = taux[i];
        comp1 = ax.x;
        a = ax.y;
        body = __float_as_uint(ax.z);
    }
    const XfBodyParams& bp = bodies[body];
    float C[3] = {bp.yeoh[0], bp.yeoh[1], bp.yeoh[2]};
    bool yeohFast = bp.energy == 7;
    float Qi[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    float QQ[3] = {q2.y, q2.z, q2.w};
    float QR[3] = {q3.x, q3.y, q3.z};
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    double* tp = tape + 12 * (size_t)i;
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) tp[3 * n + r] = x[n][r];
    solve_tet<E, true>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt);
    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
}

// ------------------------------------------------------ persistent frame

namespace {

// Spin bound of one grid-barrier wait (each spin sleeps ~2 x 64 cycles): far
// beyond any real barrier wait, small enough that a stalled grid gives up
// within a fraction of a second.
constexpr uint32_t kBarrierSpinLimit = 1u << 21;

// Grid barrier of a co-resident grid (cooperative launch). `arrive` counts
// block arrivals since the launch (zeroed by the host before it); barrier k
// waits for k * gridDim.x arrivals. On a spin timeout the error flag is set
// and every block returns false from its next barrier, so all blocks leave
// the kernel.
__device__ __forceinline__ bool grid_barrier(uint32_t* arrive, uint32_t target, uint32_t* err) {
    __shared__ uint32_t ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0, good = 1;
        while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++spins > kBarrierSpinLimit) {
                __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        ok = good;
    }
    __syncthreads();
    return ok != 0u;
}

template <int E, bool SIM, bool AUX, bool DAMP>
__device__ __forceinline__ void project_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
                                            const float4* __restrict__ tg1, const float4* __restrict__ tg2,
                                            const float4* __restrict__ tg3, const float4* __restrict__ taux,
                                            double4* __restrict__ X, const double4* __restrict__ O,
                                            const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u,
                                            float dt) {
    uint4 id = tidx[i];
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i], q3 = tg3[i];
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    float comp1 = comp1_u, a = a_u;
    uint32_t body = 0;
    if constexpr (AUX) {
        float4 ax = taux[i];
        comp1 = ax.x;
        a = ax.y;
        body = __float_as_uint(ax.z);
    }
    const XfBodyParams& bp = bodies[body];
    float C[3] = {bp.yeoh[0], bp.yeoh[1], bp.yeoh[2]};
    bool yeohFast = bp.energy == 7;
    float Qi[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    float QQ[3] = {q2.y, q2.z, q2.w};
    float QR[3] = {q3.x, q3.y, q3.z};
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    if constexpr (DAMP) {
        double4 oa = O[id.x], ob = O[id.y], oc = O[id.z], od = O[id.w];
        double o[4][3] = {{oa.x, oa.y, oa.z}, {ob.x, ob.y, ob.z}, {oc.x, oc.y, oc.z}, {od.x, od.y, od.z}};
        solve_tet<E, SIM, true>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt, o, bp.damping, bp.rtype);
    } else {
        solve_tet<E, SIM>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt);
    }
    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
}

// End of substep + next prediction for one node (the xf_finish body).
__device__ __forceinline__ void finish_one(uint32_t i, bool predict_next, double4* __restrict__ X,
                                           double4* __restrict__ O, double4* __restrict__ V,
                                           const XfBodyParams* __restrict__ bodies, float t_end) {
    double4 xi = X[i];
    double4 oi = O[i];
    uint32_t flags = (uint32_t)oi.w;
    const XfBodyParams& p = bodies[flags >> 1];
    double x[3] = {xi.x, xi.y, xi.z};
    double o[3] = {oi.x, oi.y, oi.z};
    const double w = lock_contact(p, flags, t_end, xi.w, o, x);
    double v[3];
    double invdt = (double)p.invdt;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = (x[r] - o[r]) * invdt;
    if (predict_next) {
        predict(p, v, x, &X[i], &O[i], w, oi.w);
    } else {
        X[i] = make_double4(x[0], x[1], x[2], w);
        V[i] = make_double4(v[0], v[1], v[2], 0.0);
    }
}

// Prediction of one node from its stored velocity (the xf_predict body).
__device__ __forceinline__ void predict_one(uint32_t i, double4* __restrict__ X, double4* __restrict__ O,
                                            const double4* __restrict__ V, const XfBodyParams* __restrict__ bodies) {
    double4 xi = X[i];
    double4 oi = O[i];
    double4 vi = V[i];
    const XfBodyParams& p = bodies[(uint32_t)oi.w >> 1];
    double v[3] = {vi.x, vi.y, vi.z};
    double x[3] = {xi.x, xi.y, xi.z};
    predict(p, v, x, &X[i], &O[i], xi.w, oi.w);
}

}  // namespace

// A whole step call (`substeps` substeps) in one cooperative launch: the
// first prediction, then per substep every colour phase, `vpasses` rounds of
// volume colour phases and the fused finish/predict node phase, with a grid
// barrier after each phase. With post damping (post bit 0: Rayleigh post
// pass, bit 1: PBD pass) the node phase stores the velocity instead, the
// damping colour phases follow, and then a separate prediction phase. Work is
// spread over the grid by global thread index (grid-stride loops). The
// results equal the per-colour launches bit-for-bit: each tet and node runs
// the same code on the same inputs in the same phase order.
template <int E, bool SIM, bool AUX, bool DAMP>
__global__ __launch_bounds__(256) void xf_frame_persistent(
    uint32_t substeps, uint32_t ncol, uint32_t vpasses, uint32_t post, const uint32_t* __restrict__ colour_begin,
    uint32_t n_nodes, const uint4* __restrict__ tidx, const float4* __restrict__ tg0, const float4* __restrict__ tg1,
    const float4* __restrict__ tg2, const float4* __restrict__ tg3, const float4* __restrict__ taux,
    const float2* __restrict__ tdmp, double4* __restrict__ X, double4* __restrict__ O, double4* __restrict__ V,
    const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u, float dt, float* __restrict__ clock,
    uint32_t* __restrict__ arrive, uint32_t* __restrict__ err) {
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* tickw = reinterpret_cast<uint32_t*>(clock) + 1;
    uint32_t target = 0;
    for (uint32_t i = gid; i < n_nodes; i += nthreads) predict_one(i, X, O, V, bodies);
    target += gridDim.x;
    if (!grid_barrier(arrive, target, err)) return;
    for (uint32_t s = 0; s < substeps; s++) {
        for (uint32_t c = 0; c < ncol; c++) {
            if (c == 0 && gid == 0) {
                clock[0] = clock[0] + dt;
                tickw[0] = tickw[0] + 1u;
            }
            uint32_t begin = colour_begin[c], count = colour_begin[c + 1] - begin;
            for (uint32_t k = gid; k < count; k += nthreads)
                project_one<E, SIM, AUX, DAMP>(begin + k, tidx, tg0, tg1, tg2, tg3, taux, X, O, bodies, comp1_u, a_u,
                                               dt);
            target += gridDim.x;
            if (!grid_barrier(arrive, target, err)) return;
        }
        for (uint32_t pass = 0; pass < vpasses; pass++)
            for (uint32_t c = 0; c < ncol; c++) {
                uint32_t begin = colour_begin[c], count = colour_begin[c + 1] - begin;
                for (uint32_t k = gid; k < count; k += nthreads)
                    volume_one(begin + k, pass, tidx, tg0, tg1, tg2, taux, X, bodies, dt);
                target += gridDim.x;
                if (!grid_barrier(arrive, target, err)) return;
            }
        float t_end = clock[0];
        bool next = s + 1 < substeps;
        if (post == 0u) {
            for (uint32_t i = gid; i < n_nodes; i += nthreads) finish_one(i, next, X, O, V, bodies, t_end);
            if (next) {
                target += gridDim.x;
                if (!grid_barrier(arrive, target, err)) return;
            }
            continue;
        }
        for (uint32_t i = gid; i < n_nodes; i += nthreads) finish_one(i, false, X, O, V, bodies, t_end);
        target += gridDim.x;
        if (!grid_barrier(arrive, target, err)) return;
        uint32_t tick = tickw[0] - 1u;
        for (uint32_t kind = 0; kind < 2; kind++) {
            if (!(post & (1u << kind))) continue;
            for (uint32_t c = 0; c < ncol; c++) {
                uint32_t begin = colour_begin[c], count = colour_begin[c + 1] - begin;
                for (uint32_t k = gid; k < count; k += nthreads) {
                    if (kind == 0)
                        damp_one<E, AUX>(begin + k, tidx, tg0, tg1, tg2, tg3, taux, tdmp, X, V, bodies, comp1_u, a_u,
                                         dt, tick);
                    else
                        pbd_one<AUX>(begin + k, tidx, taux, tdmp, X, V, bodies, tick);
                }
                target += gridDim.x;
                if (!grid_barrier(arrive, target, err)) return;
            }
        }
        if (next) {
            for (uint32_t i = gid; i < n_nodes; i += nthreads) predict_one(i, X, O, V, bodies);
            target += gridDim.x;
            if (!grid_barrier(arrive, target, err)) return;
        }
    }
}

// Counts nodes whose position is not finite.
__global__ __launch_bounds__(256) void xf_count_nonfinite(uint32_t n, const double4* __restrict__ X,
                                                          uint32_t* __restrict__ count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        double4 xi = X[i];
        bad = !(isfinite(xi.x) && isfinite(xi.y) && isfinite(xi.z));
    }
    unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63u) == 0 && m != 0ull) atomicAdd(count, (uint32_t)__popcll(m));
}

// ------------------------------------------------------ launch wrappers

typedef void (*XfProjectFn)(uint32_t, uint32_t, uint32_t, const uint4*, const float4*, const float4*, const float4*,
                            const float4*, const float4*, double4*, const double4*, const XfBodyParams*, float, float, float,
                            float*);


namespace {
template <int E, bool SIM>
XfProjectFn pick_aux_damp(bool aux, bool damp) {
    if (aux) return damp ? xf_project_colour<E, SIM, true, true> : xf_project_colour<E, SIM, true, false>;
    return damp ? xf_project_colour<E, SIM, false, true> : xf_project_colour<E, SIM, false, false>;
}
template <int E>
XfProjectFn pick_aux(bool sim, bool aux, bool damp) {
    return sim ? pick_aux_damp<E, true>(aux, damp) : pick_aux_damp<E, false>(aux, damp);
}
}  // namespace

// Host-callable: launch one projection kernel. energy 3..7, or 0 for
// per-body MixedSel/YeohSkinFast dispatch. damp: some body damps inside
// the projection (O is then read).
extern "C" hipError_t xf_launch_project(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t block,
                                        uint32_t slabs, uint32_t begin, uint32_t count, const uint4* tidx,
                                        const float4* tg0,
                                        const float4* tg1, const float4* tg2, const float4* tg3, const float4* taux,
                                        double4* X, const double4* O, const XfBodyParams* bodies, float comp1_u,
                                        float a_u, float dt, float* clock) {
    XfProjectFn fn = nullptr;
    switch (energy) {
        case 0: fn = pick_aux<0>(sim, aux, damp); break;
        case 3: fn = pick_aux<3>(sim, aux, damp); break;
        case 4: fn = pick_aux<4>(sim, aux, damp); break;
        case 5: fn = pick_aux<5>(sim, aux, damp); break;
        case 6: fn = pick_aux<6>(sim, aux, damp); break;
        case 7: fn = pick_aux<7>(sim, aux, damp); break;
        case 20: fn = pick_aux<20>(sim, aux, damp); break;
        case 22: fn = pick_aux<22>(sim, aux, damp); break;
        case 30: fn = pick_aux<30>(sim, aux, damp); break;
        default: return hipErrorInvalidValue;
    }
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fn, dim3(colour_grid(count, block, slabs)), dim3(block), 0, s, begin, count, slabs, tidx, tg0,
                       tg1, tg2, tg3, taux, X, O, bodies, comp1_u, a_u, dt, clock);
    return hipGetLastError();
}
#endif  // XF_BLOCKS_ONLY

typedef void (*XfBlocksFn)(uint32_t, uint32_t, uint32_t, const uint32_t*, const uint32_t*, const uint2*,
                           const uint4*, const float4*, const float4*, const float4*, const float4*, const float4*,
                           double4*, double4*, const XfBodyParams*, float, float, float, float*, const uint32_t*,
                           uint32_t, const float*, uint32_t);

namespace {
template <int E, bool OVER>
XfBlocksFn pick_blocks_quad(bool aux, bool fin) {
    if constexpr (E == 0 || E == 4 || E == 7) {
        if (fin)
            return aux ? xf_project_blocks<E, true, true, false, OVER, true, true>
                       : xf_project_blocks<E, true, false, false, OVER, true, true>;
        return aux ? xf_project_blocks<E, true, true, false, OVER, false, true>
                   : xf_project_blocks<E, true, false, false, OVER, false, true>;
    }
    return nullptr;
}
template <int E, bool SIM, bool OVER>
XfBlocksFn pick_blocks_ad(bool aux, bool damp, bool fin) {
    if (fin) {
        if constexpr (E == 30) return nullptr;
        else if (damp) return nullptr;
        else return aux ? xf_project_blocks<E, SIM, true, false, OVER, true>
                        : xf_project_blocks<E, SIM, false, false, OVER, true>;
    }
    if (aux)
        return damp ? xf_project_blocks<E, SIM, true, true, OVER, false> : xf_project_blocks<E, SIM, true, false, OVER, false>;
    return damp ? xf_project_blocks<E, SIM, false, true, OVER, false> : xf_project_blocks<E, SIM, false, false, OVER, false>;
}
// 48-byte nodes: the simultaneous MixedSel / YeohSkinFast projection without
// damping, one lane per tet, no phase beyond the workgroup's lanes; else
// nullptr (the launch keeps 32-byte nodes)
template <int E>
XfBlocksFn pick_blocks_wide(bool sim, bool aux, bool damp, bool over, bool fin, bool quad) {
    if constexpr (E == 0 || E == 4 || E == 7) {
        if (!sim || damp || over || quad) return nullptr;
        if (fin)
            return aux ? xf_project_blocks<E, true, true, false, false, true, false, 48>
                       : xf_project_blocks<E, true, false, false, false, true, false, 48>;
        return aux ? xf_project_blocks<E, true, true, false, false, false, false, 48>
                   : xf_project_blocks<E, true, false, false, false, false, false, 48>;
    }
    return nullptr;
}
template <int E>
XfBlocksFn pick_blocks(bool sim, bool aux, bool damp, bool over, bool fin, bool quad) {
    if (quad) {
        if (!sim || damp) return nullptr;
        return over ? pick_blocks_quad<E, true>(aux, fin) : pick_blocks_quad<E, false>(aux, fin);
    }
    if (over)
        return sim ? pick_blocks_ad<E, true, true>(aux, damp, fin) : pick_blocks_ad<E, false, true>(aux, damp, fin);
    return sim ? pick_blocks_ad<E, true, false>(aux, damp, fin) : pick_blocks_ad<E, false, false>(aux, damp, fin);
}
}  // namespace

#ifndef XF_BLOCKS_ONLY
// LDS bytes of a sweep-block workgroup: its nodes in FP64, ns (32 or 48)
// bytes each.
extern "C" size_t xf_block_lds_bytes(uint32_t nmax, uint32_t ns) { return (size_t)ns * nmax; }

// LDS bytes of a workgroup of the launch that ends the substep: its nodes and
// extra nodes, each with O (32 B) and id (4 B).
extern "C" size_t xf_block_fin_lds_bytes(uint32_t nmax, uint32_t fmax, uint32_t ns) {
    return ((size_t)ns + 36) * ((size_t)nmax + fmax);
}
#else
extern "C" size_t xf_block_lds_bytes(uint32_t nmax, uint32_t ns);
extern "C" size_t xf_block_fin_lds_bytes(uint32_t nmax, uint32_t fmax, uint32_t ns);
#endif

// Host-callable: one sweep-block launch (nslot workgroups of `threads`);
// rec / bnode point at the launch colour's first block; pmax: the most tets
// of one block in one phase. fin: the launch also ends the substep and
// predicts the next for every node (fextra: per workgroup its nmax block
// nodes and fmax extra nodes; tend: the clock, not advanced by this launch;
// single_body: the scene has one body); not with damp or energy 30. quad:
// four lanes per tet (simultaneous MixedSel / YeohSkinFast without damping,
// energies 0, 4 and 7; anything else is hipErrorInvalidValue). lds_limit:
// the device's LDS bytes per workgroup (more than 64 KiB is granted to the
// kernel with hipFuncAttributeMaxDynamicSharedMemorySize). wide: the launch
// may keep 48 bytes per node in LDS (NodeLds) where a 48-byte kernel exists
// for it and its LDS fits; *nstride (if not null) receives the bytes per
// node used.
extern "C" hipError_t XF_BLOCKS_LAUNCH(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t threads,
                                       uint32_t nslot, uint32_t nph, uint32_t nmax, uint32_t ntet, uint32_t pmax,
                                       const uint32_t* rec,
                                       const uint32_t* bnode, const uint2* tloc, const uint4* tidx,
                                       const float4* tg0, const float4* tg1, const float4* tg2, const float4* tg3,
                                       const float4* taux, double4* X, double4* O,
                                       const XfBodyParams* bodies, float comp1_u, float a_u, float dt,
                                       float* clock, bool fin, const uint32_t* fextra, uint32_t fmax,
                                       const float* tend, bool single_body, bool quad, size_t lds_limit,
                                       bool wide, uint32_t* nstride) {
    XfBlocksFn fn = nullptr;
    const bool over = pmax > (quad ? threads / 4 : threads);
    uint32_t ns = 32;
    if (wide && (fin ? xf_block_fin_lds_bytes(nmax, fmax, 48) : xf_block_lds_bytes(nmax, 48)) <= lds_limit) {
        switch (energy) {
            case 0: fn = pick_blocks_wide<0>(sim, aux, damp, over, fin, quad); break;
            case 4: fn = pick_blocks_wide<4>(sim, aux, damp, over, fin, quad); break;
            case 7: fn = pick_blocks_wide<7>(sim, aux, damp, over, fin, quad); break;
            default: break;
        }
        if (fn != nullptr) ns = 48;
    }
    if (nstride != nullptr) *nstride = ns;
    if (fn == nullptr) switch (energy) {
        case 0: fn = pick_blocks<0>(sim, aux, damp, over, fin, quad); break;
        case 3: fn = pick_blocks<3>(sim, aux, damp, over, fin, quad); break;
        case 4: fn = pick_blocks<4>(sim, aux, damp, over, fin, quad); break;
        case 5: fn = pick_blocks<5>(sim, aux, damp, over, fin, quad); break;
        case 6: fn = pick_blocks<6>(sim, aux, damp, over, fin, quad); break;
        case 7: fn = pick_blocks<7>(sim, aux, damp, over, fin, quad); break;
        case 20: fn = pick_blocks<20>(sim, aux, damp, over, fin, quad); break;
        case 22: fn = pick_blocks<22>(sim, aux, damp, over, fin, quad); break;
        case 30: fn = pick_blocks<30>(sim, aux, damp, over, fin, quad); break;
        default: return hipErrorInvalidValue;
    }
    size_t lds = fin ? xf_block_fin_lds_bytes(nmax, fmax, ns) : xf_block_lds_bytes(nmax, ns);
    if (fn == nullptr || threads == 0 || threads > 256 || threads % 64 || nph == 0 || ntet == 0 || lds > lds_limit)
        return hipErrorInvalidValue;
    if (fin && (tend == nullptr || clock != nullptr || fextra == nullptr)) return hipErrorInvalidValue;
    if (nslot == 0) return hipSuccess;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nslot), dim3(threads), lds, s, nph, nmax, ntet, rec, bnode, tloc, tidx, tg0, tg1,
                       tg2, tg3, taux, X, O, bodies, comp1_u, a_u, dt, clock, fextra, fmax, tend,
                       single_body ? 1u : 0u);
    return hipGetLastError();
}

#ifndef XF_BLOCKS_ONLY
typedef void (*XfDampFn)(uint32_t, uint32_t, uint32_t, const uint4*, const float4*, const float4*, const float4*,
                         const float4*, const float4*, const float2*, const double4*, double4*, const XfBodyParams*, float, float,
                         float, const float*);

namespace {
template <int E>
XfDampFn pick_damp(bool aux) {
    return aux ? xf_damp_colour<E, true> : xf_damp_colour<E, false>;
}
}  // namespace

// Rayleigh post damping of one colour (bodies with type Post/PostAmortized).
extern "C" hipError_t xf_launch_damp(hipStream_t s, int energy, bool aux, uint32_t block, uint32_t slabs,
                                     uint32_t begin, uint32_t count, const uint4* tidx, const float4* tg0, const float4* tg1,
                                     const float4* tg2, const float4* tg3, const float4* taux, const float2* tdmp,
                                     const double4* X, double4* V, const XfBodyParams* bodies, float comp1_u,
                                     float a_u, float dt, const float* clock) {
    XfDampFn fn = nullptr;
    switch (energy) {
        case 0: fn = pick_damp<0>(aux); break;
        case 3: fn = pick_damp<3>(aux); break;
        case 4: fn = pick_damp<4>(aux); break;
        case 5: fn = pick_damp<5>(aux); break;
        case 6: fn = pick_damp<6>(aux); break;
        case 7: fn = pick_damp<7>(aux); break;
        case 20: fn = pick_damp<20>(aux); break;
        case 22: fn = pick_damp<22>(aux); break;
        case 30: fn = pick_damp<30>(aux); break;
        default: return hipErrorInvalidValue;
    }
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fn, dim3(colour_grid(count, block, slabs)), dim3(block), 0, s, begin, count, slabs, tidx, tg0,
                       tg1, tg2, tg3, taux, tdmp, X, V, bodies, comp1_u, a_u, dt, clock);
    return hipGetLastError();
}

// PBD damping of one colour.
extern "C" hipError_t xf_launch_pbd(hipStream_t s, bool aux, uint32_t block, uint32_t slabs, uint32_t begin,
                                    uint32_t count,
                                    const uint4* tidx, const float4* taux, const float2* tdmp, const double4* X,
                                    double4* V, const XfBodyParams* bodies, const float* clock) {
    if (count == 0) return hipSuccess;
    uint32_t grid = colour_grid(count, block, slabs);
    if (aux)
        hipLaunchKernelGGL(xf_pbd_colour<true>, dim3(grid), dim3(block), 0, s, begin, count, slabs, tidx, taux, tdmp,
                           X, V, bodies, clock);
    else
        hipLaunchKernelGGL(xf_pbd_colour<false>, dim3(grid), dim3(block), 0, s, begin, count, slabs, tidx, taux, tdmp,
                           X, V, bodies, clock);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_volume(hipStream_t s, uint32_t block, uint32_t slabs, uint32_t begin, uint32_t count,
                                       uint32_t pass,
                                       const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                       const float4* taux, double4* X, const XfBodyParams* bodies, float dt) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_volume_colour, dim3(colour_grid(count, block, slabs)), dim3(block), 0, s, begin, count,
                       slabs, pass, tidx, tg0, tg1, tg2, taux, X, bodies, dt);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_predict(hipStream_t s, uint32_t n, double4* X, double4* O, const double4* V,
                                        const XfBodyParams* bodies) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_predict, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_finish(hipStream_t s, bool predict_next, bool single_body, uint32_t n, double4* X,
                                       double4* O, double4* V, const XfBodyParams* bodies, const float* clock) {
    if (n == 0) return hipSuccess;
    auto fn = predict_next ? (single_body ? xf_finish<true, true, false> : xf_finish<true, false, false>)
                           : (single_body ? xf_finish<false, true, false> : xf_finish<false, false, false>);
    hipLaunchKernelGGL(fn, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, clock, nullptr, nullptr);
    return hipGetLastError();
}

// xf_launch_finish that first copies X (after the sweep) and O to xs_out /
// oo_out.
extern "C" hipError_t xf_launch_finish_snap(hipStream_t s, bool predict_next, bool single_body, uint32_t n,
                                            double4* X, double4* O, double4* V, const XfBodyParams* bodies,
                                            const float* clock, double4* xs_out, double4* oo_out) {
    if (n == 0) return hipSuccess;
    auto fn = predict_next ? (single_body ? xf_finish<true, true, true> : xf_finish<true, false, true>)
                           : (single_body ? xf_finish<false, true, true> : xf_finish<false, false, true>);
    hipLaunchKernelGGL(fn, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, clock, xs_out, oo_out);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_kinematic(hipStream_t s, uint32_t n, const uint32_t* idx, const double4* target,
                                          double4* X, double4* O) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_kinematic, dim3((n + 255) / 256), dim3(256), 0, s, n, idx, target, X, O);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_manip(hipStream_t s, const float4* target, float dt, double4* X) {
    hipLaunchKernelGGL(xf_manip, dim3(1), dim3(64), 0, s, target, dt, X);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_count_nonfinite(hipStream_t s, uint32_t n, const double4* X, uint32_t* count) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_count_nonfinite, dim3((n + 255) / 256), dim3(256), 0, s, n, X, count);
    return hipGetLastError();
}

typedef void (*XfFrameFn)(uint32_t, uint32_t, uint32_t, uint32_t, const uint32_t*, uint32_t, const uint4*,
                          const float4*, const float4*, const float4*, const float4*, const float4*, const float2*,
                          double4*, double4*, double4*, const XfBodyParams*, float, float, float, float*, uint32_t*,
                          uint32_t*);

namespace {
template <int E, bool SIM>
XfFrameFn pick_frame_ad(bool aux, bool damp) {
    if (aux) return damp ? xf_frame_persistent<E, SIM, true, true> : xf_frame_persistent<E, SIM, true, false>;
    return damp ? xf_frame_persistent<E, SIM, false, true> : xf_frame_persistent<E, SIM, false, false>;
}
template <int E>
XfFrameFn pick_frame(bool sim, bool aux, bool damp) {
    return sim ? pick_frame_ad<E, true>(aux, damp) : pick_frame_ad<E, false>(aux, damp);
}
XfFrameFn frame_fn(int energy, bool sim, bool aux, bool damp) {
    switch (energy) {
        case 0: return pick_frame<0>(sim, aux, damp);
        case 3: return pick_frame<3>(sim, aux, damp);
        case 4: return pick_frame<4>(sim, aux, damp);
        case 5: return pick_frame<5>(sim, aux, damp);
        case 6: return pick_frame<6>(sim, aux, damp);
        case 7: return pick_frame<7>(sim, aux, damp);
        default: return nullptr;
    }
}
}  // namespace

// Largest co-resident grid of the persistent kernel at 256 threads per block.
extern "C" hipError_t xf_frame_max_blocks(int energy, bool sim, bool aux, bool damp, int* max_blocks) {
    XfFrameFn fn = frame_fn(energy, sim, aux, damp);
    if (!fn) return hipErrorInvalidValue;
    int dev = 0, per_cu = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, 256, 0);
    if (e == hipSuccess) *max_blocks = per_cu * cus;
    return e;
}

extern "C" hipError_t xf_launch_frame(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t blocks,
                                      uint32_t substeps, uint32_t ncol, uint32_t vpasses, uint32_t post,
                                      const float2* tdmp, const uint32_t* colour_begin,
                                      uint32_t n_nodes, const uint4* tidx, const float4* tg0, const float4* tg1,
                                      const float4* tg2, const float4* tg3, const float4* taux, double4* X,
                                      double4* O, double4* V, const XfBodyParams* bodies, float comp1_u, float a_u,
                                      float dt, float* clock, uint32_t* arrive, uint32_t* err) {
    XfFrameFn fn = frame_fn(energy, sim, aux, damp);
    if (!fn) return hipErrorInvalidValue;
    // same order as the xf_frame_persistent parameters
    void* args[] = {&substeps, &ncol, &vpasses, &post, &colour_begin, &n_nodes, &tidx, &tg0,     &tg1, &tg2, &tg3,
                    &taux,     &tdmp, &X,       &O,    &V,            &bodies,  &comp1_u, &a_u, &dt,  &clock,
                    &arrive,   &err};
    return hipLaunchCooperativeKernel((const void*)fn, dim3(blocks), dim3(256), args, 0, s);
}

extern "C" hipError_t xf_launch_project_record(hipStream_t s, int energy, bool aux, uint32_t begin, uint32_t count,
                                               const uint4* tidx, const float4* tg0, const float4* tg1,
                                               const float4* tg2, const float4* tg3, const float4* taux, double4* X,
                                               double* tape, const XfBodyParams* bodies, float comp1_u, float a_u,
                                               float dt, float* clock) {
    if (count == 0) return hipSuccess;
    dim3 grid((count + 63) / 64), block(64);
#define XF_REC(E, A) hipLaunchKernelGGL((xf_project_record<E, A>), grid, block, 0, s, begin, count, tidx, tg0, tg1, \
                                        tg2, tg3, taux, X, tape, bodies, comp1_u, a_u, dt, clock)
    switch (energy) {
        case 0: if (aux) XF_REC(0, true); else XF_REC(0, false); break;
        case 4: if (aux) XF_REC(4, true); else XF_REC(4, false); break;
        case 7: if (aux) XF_REC(7, true); else XF_REC(7, false); break;
        default: return hipErrorInvalidValue;
    }
#undef XF_REC
    return hipGetLastError();
}
#endif  // XF_BLOCKS_ONLY
