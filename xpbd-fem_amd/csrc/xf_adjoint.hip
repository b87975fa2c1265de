// xf_adjoint.hip -- gfx950 kernels of the adjoint (reverse mode) of the
// substep: gradients of a trajectory loss with respect to per-tet log mu.
// No reference counterpart (the reference has no adjoint).
//
// Precision follows the forward: FP64 node state and adjoint state, FP32
// element arithmetic; per-tet gradients are summed in FP32 over a checkpoint
// window and added to an FP64 accumulator at the window boundary. Every
// forward expression is differentiated as written, clamps included. The
// operation order matches the CPU oracle's adjoint of the same precision.
//
// FP64 element model (xf_scene_adjoint_fp64): the same rollout and adjoint
// with every element, contact and settings product in double (element
// constants and settings widened from their FP32 values, inverse masses kept
// in FP32 as the reference stores them), per-tet window sums in double: the
// trajectory and gradient of the FP64 model rather than of the reference's
// FP32 element arithmetic.
//
// Scope: simultaneous MixedSel / YeohSkinFast (energy 4, 7 or per-body 0),
// pins, ground contact. Sphere contact is not differentiated.
#include <hip/hip_runtime.h>

#include "xf_device.h"
#include "xf_element.h"

namespace {

using xf::adjdet;
using xf::apply_grad;
using xf::defgrad;
using xf::dot3;
using xf::neo_prefactored;
using xf::root_grad;

// Adjoint of one simultaneous projection (forward: xf::solve_mixed_sim, as
// xf_kernels.hip solve_tet for E = 4 / 7) in element precision R. x:
// positions before it; xbar: in, adjoint of the positions after it; out,
// adjoint of the positions before it. Returns d(loss)/d(log comp0) of this
// projection.
template <class R>
__device__ __forceinline__ R adjoint_tet(bool yeoh, const R C[3], const double x[4][3], const float w[4],
                                         const R Qi[3][3], const R QQ[3], const R QR[3], R comp0, R comp1, R a, R dt,
                                         double xbar[4][3]) {
    R P[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) P[i][r] = (R)(x[i][r] - x[3][r]);
    R g0n[4][3], U0n;
    neo_prefactored(QQ, QR, P, U0n, g0n);
    R U0 = U0n, IM = R(0.0), cpoly = R(0.0), s = R(1.0), g0[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) g0[n][r] = g0n[n][r];
    if (yeoh) {
        IM = U0n - R(3.0);
        cpoly = C[0] * IM + C[1] * IM * IM + C[2] * IM * IM * IM;
        U0 = fmax(R(0.0001), cpoly);
        s = C[0] + R(2.0) * C[1] * IM + R(3.0) * C[2] * IM * IM;
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) g0[n][r] = g0n[n][r] * s;
    }
    R F[3][3], adj[3][3], G[3][3], g1[4][3];
    defgrad(Qi, P, F);
    R J = adjdet(F, adj);
    R d = J - a;
    R U1 = R(0.0) + R(1.0) * (d * d);
    R sv = R(2.0) * d;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) G[c][r] = adj[c][r] * sv;
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) g1[n][r] = R(0.0);
    apply_grad(G, Qi, g1);
    root_grad(g1);
    R alpha0 = comp0 / (dt * dt), alpha1 = comp1 / (dt * dt);
    R A0 = R(1.0e-22), A1 = R(1.0e-22), A2 = R(1.0e-22);
#pragma unroll
    for (int n = 0; n < 4; n++) A0 = A0 + w[n] * dot3(g0[n], g0[n]);
    A0 = A0 + R(2.0) * U0 * alpha0;
    R b0 = -R(2.0) * U0;
#pragma unroll
    for (int n = 0; n < 4; n++) A1 = A1 + w[n] * dot3(g1[n], g0[n]);
#pragma unroll
    for (int n = 0; n < 4; n++) A2 = A2 + w[n] * dot3(g1[n], g1[n]);
    A2 = A2 + R(2.0) * U1 * alpha1;
    R b1 = -R(2.0) * U1;
    R u = R(1.0) / A0, v = R(1.0) / A2;
    R p = A1 * u, q = A1 * v, be0 = b0 * u, be1 = b1 * v;
    R Dr = R(1.0) - p * q;
    R invDet = R(1.0) / fmax(R(0.00000001), Dr);
    R r0 = be0 - p * be1, r1 = be1 - be0 * q;
    R l0 = invDet * r0, l1 = invDet * r1;

    R g0b[4][3], g1b[4][3];
    R l0b = R(0.0), l1b = R(0.0);
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            R accb = w[n] * (R)xbar[n][r];
            l0b = l0b + g0[n][r] * accb;
            l1b = l1b + g1[n][r] * accb;
            g0b[n][r] = l0 * accb;
            g1b[n][r] = l1 * accb;
        }
    R invDetb = l0b * r0 + l1b * r1;
    R r0b = l0b * invDet, r1b = l1b * invDet;
    R be0b = r0b - r1b * q, be1b = r1b - r0b * p;
    R pb = R(0.0) - r0b * be1, qb = R(0.0) - r1b * be0;
    R Drb = Dr > R(0.00000001) ? (R(0.0) - invDetb * invDet * invDet) : R(0.0);
    pb = pb - Drb * q;
    qb = qb - Drb * p;
    R A1b = pb * u + qb * v;
    R ub = pb * A1 + be0b * b0, vb = qb * A1 + be1b * b1;
    R b0b = be0b * u, b1b = be1b * v;
    R A0b = R(0.0) - ub * u * u, A2b = R(0.0) - vb * v * v;
    R U0b = -R(2.0) * b0b + R(2.0) * alpha0 * A0b;
    R U1b = -R(2.0) * b1b + R(2.0) * alpha1 * A2b;
    R alpha0b = R(2.0) * U0 * A0b;
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            g0b[n][r] = g0b[n][r] + R(2.0) * A0b * w[n] * g0[n][r] + A1b * w[n] * g1[n][r];
            g1b[n][r] = g1b[n][r] + A1b * w[n] * g0[n][r] + R(2.0) * A2b * w[n] * g1[n][r];
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) g1b[i][r] = g1b[i][r] - g1b[3][r];
    R adjb[3][3], Fb[3][3], Pb[3][3];
    R svb = R(0.0);
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            R Gb = g1b[0][r] * Qi[0][m] + g1b[1][r] * Qi[1][m] + g1b[2][r] * Qi[2][m];
            adjb[m][r] = Gb * sv;
            svb = svb + Gb * adj[m][r];
        }
    R db = R(2.0) * svb + R(2.0) * d * U1b;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) Fb[c][r] = db * adj[r][c];
    {
        R m00 = F[0][0], m01 = F[0][1], m02 = F[0][2];
        R m10 = F[1][0], m11 = F[1][1], m12 = F[1][2];
        R m20 = F[2][0], m21 = F[2][1], m22 = F[2][2];
        R t;
        t = adjb[0][0]; Fb[1][1] += t * m22; Fb[2][2] += t * m11; Fb[1][2] -= t * m21; Fb[2][1] -= t * m12;
        t = adjb[0][1]; Fb[0][1] -= t * m22; Fb[2][2] -= t * m01; Fb[0][2] += t * m21; Fb[2][1] += t * m02;
        t = adjb[0][2]; Fb[0][1] += t * m12; Fb[1][2] += t * m01; Fb[0][2] -= t * m11; Fb[1][1] -= t * m02;
        t = adjb[1][0]; Fb[1][0] -= t * m22; Fb[2][2] -= t * m10; Fb[1][2] += t * m20; Fb[2][0] += t * m12;
        t = adjb[1][1]; Fb[0][0] += t * m22; Fb[2][2] += t * m00; Fb[0][2] -= t * m20; Fb[2][0] -= t * m02;
        t = adjb[1][2]; Fb[0][0] -= t * m12; Fb[1][2] -= t * m00; Fb[0][2] += t * m10; Fb[1][0] += t * m02;
        t = adjb[2][0]; Fb[1][0] += t * m21; Fb[2][1] += t * m10; Fb[1][1] -= t * m20; Fb[2][0] -= t * m11;
        t = adjb[2][1]; Fb[0][0] -= t * m21; Fb[2][1] -= t * m00; Fb[0][1] += t * m20; Fb[2][0] += t * m01;
        t = adjb[2][2]; Fb[0][0] += t * m11; Fb[1][1] += t * m00; Fb[0][1] -= t * m10; Fb[1][0] -= t * m01;
    }
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int j = 0; j < 3; j++) Pb[m][j] = Qi[m][0] * Fb[j][0] + Qi[m][1] * Fb[j][1] + Qi[m][2] * Fb[j][2];
    R g0nb[4][3], U0nb;
    if (yeoh) {
        R polyb = cpoly > R(0.0001) ? U0b : R(0.0);
        R sb = R(0.0);
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                sb = sb + g0b[n][r] * g0n[n][r];
                g0nb[n][r] = g0b[n][r] * s;
            }
        U0nb = polyb * s + sb * (R(2.0) * C[1] + R(6.0) * C[2] * IM);
    } else {
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) g0nb[n][r] = g0b[n][r];
        U0nb = U0b;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) g0nb[i][r] = g0nb[i][r] - g0nb[3][r];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) Pb[i][r] = Pb[i][r] + U0nb * g0n[i][r] + R(2.0) * QQ[i] * g0nb[i][r];
    {
        int k = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i + 1; j < 3; j++) {
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    Pb[j][r] = Pb[j][r] + QR[k] * g0nb[i][r];
                    Pb[i][r] = Pb[i][r] + QR[k] * g0nb[j][r];
                }
                ++k;
            }
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            xbar[i][r] = xbar[i][r] + (double)Pb[i][r];
            xbar[3][r] = xbar[3][r] - (double)Pb[i][r];
        }
    return alpha0b * alpha0;
}

struct TetData {
    float Qi[3][3], QQ[3], QR[3], comp0, comp1, a;
    uint32_t body;
};

template <bool AUX>
__device__ __forceinline__ TetData load_tet(uint32_t i, const float4* tg0, const float4* tg1, const float4* tg2,
                                            const float4* tg3, const float4* taux, float comp1_u, float a_u) {
    TetData t;
    float4 q0 = tg0[i], q1 = tg1[i], q2 = tg2[i], q3 = tg3[i];
    t.Qi[0][0] = q0.x; t.Qi[0][1] = q0.y; t.Qi[0][2] = q0.z;
    t.Qi[1][0] = q0.w; t.Qi[1][1] = q1.x; t.Qi[1][2] = q1.y;
    t.Qi[2][0] = q1.z; t.Qi[2][1] = q1.w; t.Qi[2][2] = q2.x;
    t.QQ[0] = q2.y; t.QQ[1] = q2.z; t.QQ[2] = q2.w;
    t.QR[0] = q3.x; t.QR[1] = q3.y; t.QR[2] = q3.z;
    t.comp0 = q3.w;
    t.comp1 = comp1_u;
    t.a = a_u;
    t.body = 0;
    if constexpr (AUX) {
        float4 ax = taux[i];
        t.comp1 = ax.x;
        t.a = ax.y;
        t.body = __float_as_uint(ax.z);
    }
    return t;
}

}  // namespace

// Reverse of the projection of the tet at schedule position i. tape holds
// each tet's 4 node positions before its projection (12 doubles per
// schedule position) for this substep; W holds the inverse masses (.w);
// gacc accumulates d(loss)/d(log comp0) per schedule position. R: element
// precision (float, or double for the FP64 element model: constants
// widened from their FP32 values).
template <bool AUX, class R>
__device__ __forceinline__ void reverse_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
                                            const float4* __restrict__ tg1, const float4* __restrict__ tg2,
                                            const float4* __restrict__ tg3, const float4* __restrict__ taux,
                                            const double* __restrict__ tape, const double4* __restrict__ W,
                                            double4* __restrict__ Xb, R* __restrict__ gacc,
                                            const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u,
                                            float dt) {
    uint4 id = tidx[i];
    TetData t = load_tet<AUX>(i, tg0, tg1, tg2, tg3, taux, comp1_u, a_u);
    const XfBodyParams& bp = bodies[t.body];
    R C[3], Qi[3][3], QQ[3], QR[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        C[c] = (R)bp.yeoh[c];
        QQ[c] = (R)t.QQ[c];
        QR[c] = (R)t.QR[c];
#pragma unroll
        for (int r = 0; r < 3; r++) Qi[c][r] = (R)t.Qi[c][r];
    }
    const double* tp = tape + 12 * (size_t)i;
    double x[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) x[n][r] = tp[3 * n + r];
    uint32_t nid[4] = {id.x, id.y, id.z, id.w};
    float w[4];
    double xb[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        w[n] = (float)W[nid[n]].w;
        double4 b = Xb[nid[n]];
        xb[n][0] = b.x;
        xb[n][1] = b.y;
        xb[n][2] = b.z;
    }
    R gl = adjoint_tet<R>(bp.energy == 7, C, x, w, Qi, QQ, QR, (R)t.comp0, (R)t.comp1, (R)t.a, (R)dt, xb);
    gacc[i] = gacc[i] + gl;
#pragma unroll
    for (int n = 0; n < 4; n++) Xb[nid[n]] = make_double4(xb[n][0], xb[n][1], xb[n][2], 0.0);
}

// Reverse sweep of one colour: tets [begin, begin+count) of the sweep order.
template <bool AUX, class R>
__global__ __launch_bounds__(64) void xf_project_reverse(
    uint32_t begin, uint32_t count, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, const double* __restrict__ tape, const double4* __restrict__ W,
    double4* __restrict__ Xb, R* __restrict__ gacc, const XfBodyParams* __restrict__ bodies, float comp1_u,
    float a_u, float dt) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count)
        reverse_one<AUX, R>(begin + k, tidx, tg0, tg1, tg2, tg3, taux, tape, W, Xb, gacc, bodies, comp1_u, a_u, dt);
}

// Projection of the tet at schedule position i in element precision R
// (xf::solve_mixed_sim: the simultaneous MixedSel / YeohSkinFast solve of
// solve_tet in xf_kernels.hip, constants widened for R = double); REC
// records its four node positions before the projection into tape (12
// doubles per schedule position).
template <bool AUX, class R, bool REC>
__device__ __forceinline__ void project_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
                                            const float4* __restrict__ tg1, const float4* __restrict__ tg2,
                                            const float4* __restrict__ tg3, const float4* __restrict__ taux,
                                            double4* __restrict__ X, double* __restrict__ tape,
                                            const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u,
                                            float dt) {
    uint4 id = tidx[i];
    TetData t = load_tet<AUX>(i, tg0, tg1, tg2, tg3, taux, comp1_u, a_u);
    const XfBodyParams& bp = bodies[t.body];
    R C[3], Qi[3][3], QQ[3], QR[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        C[c] = (R)bp.yeoh[c];
        QQ[c] = (R)t.QQ[c];
        QR[c] = (R)t.QR[c];
#pragma unroll
        for (int r = 0; r < 3; r++) Qi[c][r] = (R)t.Qi[c][r];
    }
    double4 na = X[id.x], nb = X[id.y], nc = X[id.z], nd = X[id.w];
    double x[4][3] = {{na.x, na.y, na.z}, {nb.x, nb.y, nb.z}, {nc.x, nc.y, nc.z}, {nd.x, nd.y, nd.z}};
    float w[4] = {(float)na.w, (float)nb.w, (float)nc.w, (float)nd.w};
    if constexpr (REC) {
        double* tp = tape + 12 * (size_t)i;
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int r = 0; r < 3; r++) tp[3 * n + r] = x[n][r];
    }
    xf::solve_mixed_sim<R>(bp.energy == 7, C, x, w, Qi, QQ, QR, (R)t.comp0, (R)t.comp1, (R)t.a, (R)dt);
    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
}

// One launch of two independent sweeps of the pipelined reverse pass:
// workgroups [0, rblocks) record colour [rbegin, rbegin + rcount) of a
// window replay (on X, into tape_rec); the rest take the reverse of colour
// [vbegin, vbegin + vcount) of the later window (tape_rev, inverse masses
// from W, adjoint Xb, gacc). The two touch disjoint arrays.
template <bool AUX, class R>
__global__ __launch_bounds__(64) void xf_adj_pair(
    uint32_t rbegin, uint32_t rcount, uint32_t rblocks, double4* __restrict__ X, double* __restrict__ tape_rec,
    uint32_t vbegin, uint32_t vcount, const double* __restrict__ tape_rev, const double4* __restrict__ W,
    double4* __restrict__ Xb, R* __restrict__ gacc, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, const XfBodyParams* __restrict__ bodies, float comp1_u, float a_u, float dt) {
    if (blockIdx.x < rblocks) {
        uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k < rcount)
            project_one<AUX, R, true>(rbegin + k, tidx, tg0, tg1, tg2, tg3, taux, X, tape_rec, bodies, comp1_u, a_u,
                                      dt);
    } else {
        uint32_t k = (blockIdx.x - rblocks) * blockDim.x + threadIdx.x;
        if (k < vcount)
            reverse_one<AUX, R>(vbegin + k, tidx, tg0, tg1, tg2, tg3, taux, tape_rev, W, Xb, gacc, bodies, comp1_u,
                                a_u, dt);
    }
}

// ---------------------------------------------- FP64 element model forward
//
// Geo3d::Substep (Geo.cpp:305-356) with the element, contact and settings
// arithmetic in double, as the oracle's f64 variant: predict, colour
// launches of the simultaneous MixedSel / YeohSkinFast projection
// (xf::solve_mixed_sim<double>), then pins, ground contact and velocity.

// Prediction from V (Geo.cpp:307-312): v += g dt, v *= 1 - drag, O = X,
// X += v dt, with g dt and 1 - drag formed in double. Thread 0 advances the
// scene clock by dt (the substep's end time; not read by this model).
__global__ __launch_bounds__(256) void xf_predict64(uint32_t n, double4* __restrict__ X, double4* __restrict__ O,
                                                    const double4* __restrict__ V,
                                                    const XfBodyParams* __restrict__ bodies, float dt,
                                                    float* __restrict__ clock) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && clock != nullptr) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    if (i >= n) return;
    double4 xi = X[i], oi = O[i], vi = V[i];
    const XfBodyParams& p = bodies[(uint32_t)oi.w >> 1];
    const double d = (double)p.dt;
    const double gdx = (double)p.gravity[0] * d, gdy = (double)p.gravity[1] * d;
    const double dragf = 1.0 - (double)p.drag;
    double v[3] = {vi.x + gdx, vi.y + gdy, vi.z + 0.0};
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = v[r] * dragf;
    O[i] = make_double4(xi.x, xi.y, xi.z, oi.w);
    X[i] = make_double4(xi.x + v[0] * d, xi.y + v[1] * d, xi.z + v[2] * d, xi.w);
}

// One colour of the FP64 projection; REC: records each tet's four node
// positions before its projection into tape (12 doubles per schedule
// position), as xf_project_record does for the FP32 model.
template <bool AUX, bool REC>
__global__ __launch_bounds__(64) void xf_project64(uint32_t begin, uint32_t count, const uint4* __restrict__ tidx,
                                                   const float4* __restrict__ tg0, const float4* __restrict__ tg1,
                                                   const float4* __restrict__ tg2, const float4* __restrict__ tg3,
                                                   const float4* __restrict__ taux, double4* __restrict__ X,
                                                   double* __restrict__ tape, const XfBodyParams* __restrict__ bodies,
                                                   float comp1_u, float a_u, float dt) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count)
        project_one<AUX, double, REC>(begin + k, tidx, tg0, tg1, tg2, tg3, taux, X, tape, bodies, comp1_u, a_u, dt);
}

// End of an FP64 substep for one node: a pinned node returns to O; other
// nodes take the ground contact in double (static friction when the
// tangential displacement is within mu * depth, else it shrinks by mu *
// depth); then V = (X - O) / dt with 1 / dt in double.
// xs_out / oo_out (nullable): X after the sweep and O are copied there first.
__global__ __launch_bounds__(256) void xf_finish64(uint32_t n, double4* __restrict__ X, const double4* __restrict__ O,
                                                   double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
                                                   double4* __restrict__ xs_out, double4* __restrict__ oo_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 xi = X[i], oi = O[i];
    if (xs_out) {
        xs_out[i] = xi;
        oo_out[i] = oi;
    }
    const uint32_t flags = (uint32_t)oi.w;
    const XfBodyParams& p = bodies[flags >> 1];
    double x[3] = {xi.x, xi.y, xi.z};
    const double o[3] = {oi.x, oi.y, oi.z};
    if (flags & 1u) {
        x[0] = o[0];
        x[1] = o[1];
        x[2] = o[2];
    } else if (p.collide & 1u) {
        const double gy = (double)p.ground_y;
        if (x[1] < gy) {
            const double mu = (double)p.friction;
            const double depth = gy - x[1];
            x[1] = gy;
            const double dx = x[0] - o[0], dz = x[2] - o[2];
            const double t2 = dx * dx + dz * dz;
            const double lim = mu * depth;
            if (t2 <= lim * lim) {
                x[0] = o[0];
                x[2] = o[2];
            } else {
                const double sc = lim / sqrt(t2);
                x[0] = x[0] - dx * sc;
                x[2] = x[2] - dz * sc;
            }
        }
    }
    const double invdt = 1.0 / (double)p.dt;
    X[i] = make_double4(x[0], x[1], x[2], xi.w);
    V[i] = make_double4((x[0] - o[0]) * invdt, (x[1] - o[1]) * invdt, (x[2] - o[2]) * invdt, 0.0);
}

// Square root of the element precision.
__device__ __forceinline__ float rsqrt_of(float x) { return sqrtf(x); }
__device__ __forceinline__ double rsqrt_of(double x) { return sqrt(x); }

// Reverse of the end of one substep (velocity, pin, ground contact) for one
// node, with the loss injected first when this substep ends a loss frame:
// Xb += lw_i * (x_loss - target). xs / o: the node's position before the
// contact pass and at the substep start (recorded by the forward replay).
// R: element precision of the forward (float: FP32 contact and 1.0f / dt;
// double: the FP64 element model). xbv / vbv: the node's Xb, Vb in, Xb out;
// ob receives Ob.
template <class R>
__device__ __forceinline__ void node_end_one(uint32_t i, uint32_t flags, const XfBodyParams& p,
                                             const double4* __restrict__ xs, const double4* __restrict__ o,
                                             const double4* __restrict__ xloss, const double4* __restrict__ target,
                                             const double* __restrict__ lw, double xbv[3], const double vbv[3],
                                             double ob[3]) {
    if (xloss) {
        double4 xl = xloss[i], tg = target[i];
        double wl = lw[i];
        xbv[0] = xbv[0] + wl * (xl.x - tg.x);
        xbv[1] = xbv[1] + wl * (xl.y - tg.y);
        xbv[2] = xbv[2] + wl * (xl.z - tg.z);
    }
    const double invdt = sizeof(R) == 4 ? (double)p.invdt : 1.0 / (double)p.dt;
    double xcb[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        xcb[r] = xbv[r] + vbv[r] * invdt;
        ob[r] = 0.0 - vbv[r] * invdt;
    }
    double out[3] = {xcb[0], xcb[1], xcb[2]};
    if (flags & 1u) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            ob[r] = ob[r] + xcb[r];
            out[r] = 0.0;
        }
    } else if (p.collide & 1u) {
        double4 s4 = xs[i], o4 = o[i];
        R gy = (R)p.ground_y;
        R xy = (R)s4.y;
        if (xy < gy) {
            R mu = (R)p.friction;
            R depth = gy - xy;
            R dx = (R)(s4.x - o4.x), dz = (R)(s4.z - o4.z);
            R t2 = dx * dx + dz * dz;
            R lim = mu * depth;
            out[1] = 0.0;
            if (t2 <= lim * lim) {
                ob[0] = ob[0] + xcb[0];
                ob[2] = ob[2] + xcb[2];
                out[0] = 0.0;
                out[2] = 0.0;
            } else {
                R L = rsqrt_of(t2);
                R ux = dx / L, uz = dz / L;
                R bx = (R)xcb[0], bz = (R)xcb[2];
                R ubv = ux * bx + uz * bz;
                R kk = lim / L;
                R px = kk * (bx - ux * ubv), pz = kk * (bz - uz * ubv);
                out[0] = xcb[0] - (double)px;
                out[2] = xcb[2] - (double)pz;
                ob[0] = ob[0] + (double)px;
                ob[2] = ob[2] + (double)pz;
                out[1] = (double)(mu * ubv);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) xbv[r] = out[r];
}

// Reverse of the prediction (Geo.cpp:307-312) for one node: Xb <- Xb + Ob,
// Vb <- Xb_pred * dt * (1 - drag) (1 - drag: the FP32 value, or formed in
// double for the FP64 element model).
template <class R>
__device__ __forceinline__ void node_start_one(const XfBodyParams& p, double xb[3], const double ob[3],
                                               double vb[3]) {
    const double dt = (double)p.dt;
    const double dragf = sizeof(R) == 4 ? (double)p.dragf : 1.0 - (double)p.drag;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double x0b = xb[r];
        xb[r] = x0b + ob[r];
        vb[r] = (x0b * dt) * dragf;
    }
}

// Reverse of the node steps around one substep boundary, per node:
// START: the prediction of the later substep (node_start_one, from Ob);
// END: then the end of the earlier substep (node_end_one, recorded xs / o,
// loss at xloss when it ends a loss frame), writing Ob for its sweep.
template <class R, bool START, bool END>
__global__ __launch_bounds__(256) void xf_adj_nodes(uint32_t n, const double4* __restrict__ O,
                                                    const double4* __restrict__ xs, const double4* __restrict__ o,
                                                    double4* __restrict__ Xb, double4* __restrict__ Vb,
                                                    double4* __restrict__ Ob, const XfBodyParams* __restrict__ bodies,
                                                    const double4* __restrict__ xloss,
                                                    const double4* __restrict__ target,
                                                    const double* __restrict__ lw) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t flags = (uint32_t)O[i].w;
    const XfBodyParams& p = bodies[flags >> 1];
    const double4 xb4 = Xb[i], vb4 = Vb[i];
    double xb[3] = {xb4.x, xb4.y, xb4.z}, vb[3] = {vb4.x, vb4.y, vb4.z};
    if constexpr (START) {
        const double4 o4 = Ob[i];
        const double ob[3] = {o4.x, o4.y, o4.z};
        node_start_one<R>(p, xb, ob, vb);
    }
    if constexpr (END) {
        double ob[3];
        node_end_one<R>(i, flags, p, xs, o, xloss, target, lw, xb, vb, ob);
        Ob[i] = make_double4(ob[0], ob[1], ob[2], 0.0);
    }
    Xb[i] = make_double4(xb[0], xb[1], xb[2], 0.0);
    if constexpr (START) Vb[i] = make_double4(vb[0], vb[1], vb[2], 0.0);
}

// Window boundary: glog += gacc (the window sum in element precision into
// the FP64 total); gacc = 0.
template <class R>
__global__ __launch_bounds__(256) void xf_adj_flush(uint32_t t, R* __restrict__ gacc, double* __restrict__ glog) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t) return;
    glog[i] += (double)gacc[i];
    gacc[i] = R(0.0);
}

// Zero-fill of a double4 array (adjoint state reset).
__global__ __launch_bounds__(256) void xf_zero_d4(uint32_t n, double4* __restrict__ a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// Zero-fill of n 32-bit words (gradient accumulators; a kernel rather than
// a memset so that the caller's stream may be under graph capture).
__global__ __launch_bounds__(256) void xf_zero_u32(uint32_t n, uint32_t* __restrict__ a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// Loss targets as the caller passes them (3 doubles a node, n entries over
// all loss substeps) to the double4 layout of the node state.
__global__ __launch_bounds__(256) void xf_expand_targets(uint32_t n, const double* __restrict__ raw,
                                                         double4* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double4(raw[3 * (size_t)i], raw[3 * (size_t)i + 1], raw[3 * (size_t)i + 2], 0.0);
}

// Per-node weight of the loss gradient, 2 / (nodes of the node's body); the
// body id is O.w >> 1.
__global__ __launch_bounds__(256) void xf_loss_weights(uint32_t n, const double4* __restrict__ O,
                                                       const XfBodyParams* __restrict__ bodies,
                                                       double* __restrict__ lw) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t body = (uint32_t)O[i].w >> 1;
    lw[i] = 2.0 / (double)bodies[body].n_nodes;
}

// Squared-error sum of one body at one loss substep: thread (j, b) adds
// |x - target|^2 over the body's nodes in node order, component by
// component, into e[j * nb + b] (one thread per term keeps the summation
// order of a serial loop).
__global__ __launch_bounds__(64) void xf_body_sq_err(uint32_t n_loss, uint32_t nb, uint32_t n,
                                                     const double4* __restrict__ xloss,
                                                     const double4* __restrict__ tgt,
                                                     const XfBodyParams* __restrict__ bodies,
                                                     double* __restrict__ e) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_loss * nb) return;
    uint32_t j = k / nb, b = k % nb;
    const XfBodyParams& bp = bodies[b];
    const double4* x = xloss + (size_t)n * j + bp.node_base;
    const double4* t = tgt + (size_t)n * j + bp.node_base;
    double acc = 0.0;
    for (uint32_t i = 0; i < bp.n_nodes; i++) {
        double d0 = x[i].x - t[i].x, d1 = x[i].y - t[i].y, d2 = x[i].z - t[i].z;
        acc += d0 * d0;
        acc += d1 * d1;
        acc += d2 * d2;
    }
    e[k] = acc;
}

// L = sum over loss substeps, then bodies, of e / n_b, and each body's share,
// in that order by one thread: out[0] = L, out[1 + b] = body b's terms.
__global__ __launch_bounds__(64) void xf_loss_total(uint32_t n_loss, uint32_t nb, const double* __restrict__ e,
                                                    const XfBodyParams* __restrict__ bodies,
                                                    double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (uint32_t b = 0; b < nb; b++) out[1 + b] = 0.0;
    double L = 0.0;
    for (uint32_t j = 0; j < n_loss; j++)
        for (uint32_t b = 0; b < nb; b++) {
            double v = e[(size_t)j * nb + b] / (double)bodies[b].n_nodes;
            L += v;
            out[1 + b] += v;
        }
    out[0] = L;
}

// Gradient in the caller's order: schedule position p holds the tet whose
// output slot is tout[p]; dL/dlog mu = -glog.
__global__ __launch_bounds__(256) void xf_grad_out(uint32_t t, const double* __restrict__ glog,
                                                   const uint32_t* __restrict__ tout, double* __restrict__ g) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < t) g[tout[p]] = -glog[p];
}

// Copy of two double4 arrays of n entries (the replay's per-substep
// snapshot of X and O) in one launch.
__global__ __launch_bounds__(256) void xf_copy2_d4(uint32_t n, double4* __restrict__ d0,
                                                   const double4* __restrict__ s0, double4* __restrict__ d1,
                                                   const double4* __restrict__ s1) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        d0[i] = s0[i];
        d1[i] = s1[i];
    }
}

// ------------------------------------------------------ launch wrappers

// fp64: the FP64 element model (gacc then holds doubles).
extern "C" hipError_t xf_launch_project_reverse(hipStream_t s, bool aux, bool fp64, uint32_t begin, uint32_t count,
                                                const uint4* tidx, const float4* tg0, const float4* tg1,
                                                const float4* tg2, const float4* tg3, const float4* taux,
                                                const double* tape, const double4* W, double4* Xb, void* gacc,
                                                const XfBodyParams* bodies, float comp1_u, float a_u, float dt) {
    if (count == 0) return hipSuccess;
    dim3 grid((count + 63) / 64), block(64);
    float* g32 = static_cast<float*>(gacc);
    double* g64 = static_cast<double*>(gacc);
    if (fp64 && aux)
        hipLaunchKernelGGL((xf_project_reverse<true, double>), grid, block, 0, s, begin, count, tidx, tg0, tg1, tg2,
                           tg3, taux, tape, W, Xb, g64, bodies, comp1_u, a_u, dt);
    else if (fp64)
        hipLaunchKernelGGL((xf_project_reverse<false, double>), grid, block, 0, s, begin, count, tidx, tg0, tg1, tg2,
                           tg3, taux, tape, W, Xb, g64, bodies, comp1_u, a_u, dt);
    else if (aux)
        hipLaunchKernelGGL((xf_project_reverse<true, float>), grid, block, 0, s, begin, count, tidx, tg0, tg1, tg2,
                           tg3, taux, tape, W, Xb, g32, bodies, comp1_u, a_u, dt);
    else
        hipLaunchKernelGGL((xf_project_reverse<false, float>), grid, block, 0, s, begin, count, tidx, tg0, tg1, tg2,
                           tg3, taux, tape, W, Xb, g32, bodies, comp1_u, a_u, dt);
    return hipGetLastError();
}

// Record of colour [rbegin, rbegin + rcount) and reverse of colour
// [vbegin, vbegin + vcount) in one launch (xf_adj_pair); either count may
// be 0.
extern "C" hipError_t xf_launch_adj_pair(hipStream_t s, bool aux, bool fp64, uint32_t rbegin, uint32_t rcount,
                                         double4* X, double* tape_rec, uint32_t vbegin, uint32_t vcount,
                                         const double* tape_rev, const double4* W, double4* Xb, void* gacc,
                                         const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                         const float4* tg3, const float4* taux, const XfBodyParams* bodies,
                                         float comp1_u, float a_u, float dt) {
    const uint32_t rblocks = (rcount + 63) / 64, vblocks = (vcount + 63) / 64;
    if (rblocks + vblocks == 0) return hipSuccess;
    dim3 grid(rblocks + vblocks), block(64);
#define XF_PAIR(A, R, G)                                                                                         \
    hipLaunchKernelGGL((xf_adj_pair<A, R>), grid, block, 0, s, rbegin, rcount, rblocks, X, tape_rec, vbegin,   \
                       vcount, tape_rev, W, Xb, static_cast<R*>(G), tidx, tg0, tg1, tg2, tg3, taux, bodies, comp1_u, \
                       a_u, dt)
    if (fp64) {
        if (aux) XF_PAIR(true, double, gacc);
        else XF_PAIR(false, double, gacc);
    } else {
        if (aux) XF_PAIR(true, float, gacc);
        else XF_PAIR(false, float, gacc);
    }
#undef XF_PAIR
    return hipGetLastError();
}

// start / end: the reverse of the later substep's prediction and of the
// earlier substep's end (xs, o, xloss, target, lw are read only with end).
extern "C" hipError_t xf_launch_adj_nodes(hipStream_t s, bool fp64, bool start, bool end, uint32_t n,
                                          const double4* O, const double4* xs, const double4* o, double4* Xb,
                                          double4* Vb, double4* Ob, const XfBodyParams* bodies,
                                          const double4* xloss, const double4* target, const double* lw) {
    if (n == 0 || (!start && !end)) return hipSuccess;
    dim3 grid((n + 255) / 256), block(256);
#define XF_NODES(R, S, E)                                                                                    \
    hipLaunchKernelGGL((xf_adj_nodes<R, S, E>), grid, block, 0, s, n, O, xs, o, Xb, Vb, Ob, bodies, xloss, \
                       target, lw)
    if (fp64) {
        if (start && end) XF_NODES(double, true, true);
        else if (start) XF_NODES(double, true, false);
        else XF_NODES(double, false, true);
    } else {
        if (start && end) XF_NODES(float, true, true);
        else if (start) XF_NODES(float, true, false);
        else XF_NODES(float, false, true);
    }
#undef XF_NODES
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_adj_flush(hipStream_t s, bool fp64, uint32_t t, void* gacc, double* glog) {
    if (t == 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL(xf_adj_flush<double>, dim3((t + 255) / 256), dim3(256), 0, s, t, static_cast<double*>(gacc),
                           glog);
    else
        hipLaunchKernelGGL(xf_adj_flush<float>, dim3((t + 255) / 256), dim3(256), 0, s, t, static_cast<float*>(gacc),
                           glog);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_predict64(hipStream_t s, uint32_t n, double4* X, double4* O, const double4* V,
                                          const XfBodyParams* bodies, float dt, float* clock) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_predict64, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, dt, clock);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_project64(hipStream_t s, bool aux, bool rec, uint32_t begin, uint32_t count,
                                          const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                          const float4* tg3, const float4* taux, double4* X, double* tape,
                                          const XfBodyParams* bodies, float comp1_u, float a_u, float dt) {
    if (count == 0) return hipSuccess;
    dim3 grid((count + 63) / 64), block(64);
#define XF_P64(A, RC)                                                                                              \
    hipLaunchKernelGGL((xf_project64<A, RC>), grid, block, 0, s, begin, count, tidx, tg0, tg1, tg2, tg3, taux, X, \
                       tape, bodies, comp1_u, a_u, dt)
    if (aux) {
        if (rec) XF_P64(true, true);
        else XF_P64(true, false);
    } else {
        if (rec) XF_P64(false, true);
        else XF_P64(false, false);
    }
#undef XF_P64
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_finish64(hipStream_t s, uint32_t n, double4* X, const double4* O, double4* V,
                                         const XfBodyParams* bodies, double4* xs_out, double4* oo_out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_finish64, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, xs_out, oo_out);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_copy2_d4(hipStream_t s, uint32_t n, double4* d0, const double4* s0, double4* d1,
                                         const double4* s1) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_copy2_d4, dim3((n + 255) / 256), dim3(256), 0, s, n, d0, s0, d1, s1);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_zero_d4(hipStream_t s, uint32_t n, double4* a) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_zero_d4, dim3((n + 255) / 256), dim3(256), 0, s, n, a);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_zero_u32(hipStream_t s, uint32_t n, uint32_t* a) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, n, a);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_expand_targets(hipStream_t s, uint32_t n, const double* raw, double4* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_expand_targets, dim3((n + 255) / 256), dim3(256), 0, s, n, raw, out);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_loss_weights(hipStream_t s, uint32_t n, const double4* O, const XfBodyParams* bodies,
                                             double* lw) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_loss_weights, dim3((n + 255) / 256), dim3(256), 0, s, n, O, bodies, lw);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_loss(hipStream_t s, uint32_t n_loss, uint32_t nb, uint32_t n, const double4* xloss,
                                     const double4* tgt, const XfBodyParams* bodies, double* e, double* out) {
    uint32_t terms = n_loss * nb;
    if (terms == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_body_sq_err, dim3((terms + 63) / 64), dim3(64), 0, s, n_loss, nb, n, xloss, tgt, bodies, e);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(xf_loss_total, dim3(1), dim3(64), 0, s, n_loss, nb, e, bodies, out);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_grad_out(hipStream_t s, uint32_t t, const double* glog, const uint32_t* tout,
                                         double* g) {
    if (t == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_grad_out, dim3((t + 255) / 256), dim3(256), 0, s, t, glog, tout, g);
    return hipGetLastError();
}
