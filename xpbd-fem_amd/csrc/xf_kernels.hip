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
#include <hip/hip_runtime.h>

#include "xf_device.h"

namespace {

__device__ __forceinline__ float dot3(const float* a, const float* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// F = Qi * MatMul(dN, P), dN = e_x, e_y, e_z (Fem.cpp:312-323).
__device__ __forceinline__ void defgrad(const float Qi[3][3], const float P[3][3], float F[3][3]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) F[j][r] = Qi[0][r] * P[0][j] + Qi[1][r] * P[1][j] + Qi[2][r] * P[2][j];
}

// adjugate + determinantFromAdjugate (Fem.cpp:284-306).
__device__ __forceinline__ float adjdet(const float F[3][3], float adj[3][3]) {
    float m00 = F[0][0], m01 = F[0][1], m02 = F[0][2];
    float m10 = F[1][0], m11 = F[1][1], m12 = F[1][2];
    float m20 = F[2][0], m21 = F[2][1], m22 = F[2][2];
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
__device__ __forceinline__ void apply_grad(const float G[3][3], const float Qi[3][3], float g[4][3]) {
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) g[j][r] = g[j][r] + (G[0][r] * Qi[j][0] + G[1][r] * Qi[j][1] + G[2][r] * Qi[j][2]);
}

__device__ __forceinline__ void root_grad(float g[4][3]) {
#pragma unroll
    for (int r = 0; r < 3; r++) g[3][r] = ((0.0f - g[0][r]) - g[1][r]) - g[2][r];
}

// (J - a)^2 volume term (Fem.cpp:537-547).
__device__ __forceinline__ void volume_term(const float Qi[3][3], const float P[3][3], float a, float& U1,
                                            float g1[4][3]) {
    float F[3][3], adj[3][3], GJ[3][3];
    defgrad(Qi, P, F);
    float J = adjdet(F, adj);
    U1 = U1 + 1.0f * ((J - a) * (J - a));
    float s = 2.0f * (J - a);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) GJ[c][r] = adj[c][r] * s;
    apply_grad(GJ, Qi, g1);
}

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

// Prefactored I1 (Fem.cpp:164-192).
__device__ __forceinline__ void neo_prefactored(const float QQ[3], const float QR[3], const float P[3][3], float& U,
                                                float g[4][3]) {
    float u = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        u = u + QQ[i] * dot3(P[i], P[i]);
#pragma unroll
        for (int r = 0; r < 3; r++) g[i][r] = (2.0f * QQ[i]) * P[i][r];
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

__device__ __forceinline__ void gather_p(const double x[4][3], float P[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) P[i][r] = (float)(x[i][r] - x[3][r]);
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

// EnergyXpbdConstrainSimultaneous<2> without damping, Cramer branch (Xpbd.h:123-214).
__device__ __forceinline__ void constrain2(double x[4][3], const float w[4], const float U[2], const float g0[4][3],
                                           const float g1[4][3], float dt, float comp0, float comp1) {
    float alpha0 = comp0 / (dt * dt);
    float alpha1 = comp1 / (dt * dt);
    float A0 = 1.0e-22f, A1 = 1.0e-22f, A2 = 1.0e-22f;
#pragma unroll
    for (int n = 0; n < 4; n++) A0 = A0 + w[n] * dot3(g0[n], g0[n]);
    A0 = A0 + 2.0f * U[0] * alpha0;
    float b0 = -2.0f * U[0];
#pragma unroll
    for (int n = 0; n < 4; n++) A1 = A1 + w[n] * dot3(g1[n], g0[n]);
#pragma unroll
    for (int n = 0; n < 4; n++) A2 = A2 + w[n] * dot3(g1[n], g1[n]);
    A2 = A2 + 2.0f * U[1] * alpha1;
    float b1 = -2.0f * U[1];
    float invA00 = 1.0f / A0;
    float invA11 = 1.0f / A2;
    float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A1 * invA00) * (A1 * invA11));
    float l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
    float l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
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
// E: 3 Mixed, 4 MixedSel, 5 YeohSkin, 6 YeohSkinSel, 7 YeohSkinFast,
//    0 = MixedSel or YeohSkinFast chosen per tet by `yeohFast`.
template <int E, bool SIM>
__device__ __forceinline__ void solve_tet(bool yeohFast, const float C[3], double x[4][3], const float w[4],
                                          const float Qi[3][3], const float QQ[3], const float QR[3], float comp0,
                                          float comp1, float a, float dt) {
    float P[3][3];
    gather_p(x, P);
    float U[2] = {0.0f, 0.0f};
    float g0[4][3], g1[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int r = 0; r < 3; r++) { g0[n][r] = 0.0f; g1[n][r] = 0.0f; }

    if constexpr (E == 3 || E == 5 || E == 6) {
        i1_quad_term(E != 3, C, Qi, P, U[0], g0);
        if constexpr (SIM && E != 6) volume_term(Qi, P, a, U[1], g1);
        U[0] = fmaxf(0.0001f, U[0]);
        if constexpr (SIM) {
            if constexpr (E == 6) volume_term(Qi, P, a, U[1], g1);
            root_grad(g0);
            root_grad(g1);
            constrain2(x, w, U, g0, g1, dt, comp0, comp1);
        } else {
            root_grad(g0);
            constrain1(x, w, U[0], g0, dt, comp0);
            gather_p(x, P);
            volume_term(Qi, P, a, U[1], g1);
            root_grad(g1);
            constrain1(x, w, U[1], g1, dt, comp1);
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
            constrain1(x, w, U[0], g0, dt, comp0);
            gather_p(x, P);
        }
        volume_term(Qi, P, a, U[1], g1);
        root_grad(g1);
        if constexpr (!SIM) constrain1(x, w, U[1], g1, dt, comp1);
        else constrain2(x, w, U, g0, g1, dt, comp0, comp1);
    }
}

}  // namespace

// ---------------------------------------------------------------- kernels

// Projects one colour: tets [begin, begin+count) of the sweep order.
// AUX: per-tet (comp1, a, body) from taux; otherwise comp1/a are uniform and
// the body is 0. The first colour launch of a substep also advances the
// simulation clock (clock != nullptr); no other kernel of the colour sweep
// reads it.
template <int E, bool SIM, bool AUX>
__global__ __launch_bounds__(256) void xf_project_colour(
    uint32_t begin, uint32_t count, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
    const float4* __restrict__ tg1, const float4* __restrict__ tg2, const float4* __restrict__ tg3,
    const float4* __restrict__ taux, double4* __restrict__ X, const XfBodyParams* __restrict__ bodies, float comp1_u,
    float a_u, float dt, float* __restrict__ clock) {
    if (clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) clock[0] = clock[0] + dt;
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

    solve_tet<E, SIM>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt);

    X[id.x] = make_double4(x[0][0], x[0][1], x[0][2], na.w);
    X[id.y] = make_double4(x[1][0], x[1][1], x[1][2], nb.w);
    X[id.z] = make_double4(x[2][0], x[2][1], x[2][2], nc.w);
    X[id.w] = make_double4(x[3][0], x[3][1], x[3][2], nd.w);
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

}  // namespace

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
template <bool PREDICT>
__global__ __launch_bounds__(256) void xf_finish(uint32_t n, double4* __restrict__ X, double4* __restrict__ O,
                                                 double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
                                                 const float* __restrict__ clock) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 xi = X[i];
    double4 oi = O[i];
    uint32_t flags = (uint32_t)oi.w;
    const XfBodyParams& p = bodies[flags >> 1];
    double x[3] = {xi.x, xi.y, xi.z};
    double o[3] = {oi.x, oi.y, oi.z};
    if (flags & 1u) {
        x[0] = o[0];
        x[1] = o[1];
        x[2] = o[2];
    } else if (p.collide) {
        collide_node(p, clock[0], o, x);
    }
    double v[3];
    double invdt = (double)p.invdt;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = (x[r] - o[r]) * invdt;
    if constexpr (PREDICT) {
        predict(p, v, x, &X[i], &O[i], xi.w, oi.w);
    } else {
        X[i] = make_double4(x[0], x[1], x[2], xi.w);
        V[i] = make_double4(v[0], v[1], v[2], 0.0);
    }
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
    if (clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) clock[0] = clock[0] + dt;
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

template <int E, bool SIM, bool AUX>
__device__ __forceinline__ void project_one(uint32_t i, const uint4* __restrict__ tidx, const float4* __restrict__ tg0,
                                            const float4* __restrict__ tg1, const float4* __restrict__ tg2,
                                            const float4* __restrict__ tg3, const float4* __restrict__ taux,
                                            double4* __restrict__ X, const XfBodyParams* __restrict__ bodies,
                                            float comp1_u, float a_u, float dt) {
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
    solve_tet<E, SIM>(yeohFast, C, x, w, Qi, QQ, QR, q3.w, comp1, a, dt);
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
    if (flags & 1u) {
        x[0] = o[0];
        x[1] = o[1];
        x[2] = o[2];
    } else if (p.collide) {
        collide_node(p, t_end, o, x);
    }
    double v[3];
    double invdt = (double)p.invdt;
#pragma unroll
    for (int r = 0; r < 3; r++) v[r] = (x[r] - o[r]) * invdt;
    if (predict_next) {
        predict(p, v, x, &X[i], &O[i], xi.w, oi.w);
    } else {
        X[i] = make_double4(x[0], x[1], x[2], xi.w);
        V[i] = make_double4(v[0], v[1], v[2], 0.0);
    }
}

}  // namespace

// A whole step call (`substeps` substeps) in one cooperative launch: the
// first prediction, then per substep every colour phase and the fused
// finish/predict node phase, with a grid barrier after each phase. Work is
// spread over the grid by global thread index (grid-stride loops). The
// results equal the per-colour launches bit-for-bit: each tet and node runs
// the same code on the same inputs in the same phase order.
template <int E, bool SIM, bool AUX>
__global__ __launch_bounds__(256) void xf_frame_persistent(
    uint32_t substeps, uint32_t ncol, const uint32_t* __restrict__ colour_begin, uint32_t n_nodes,
    const uint4* __restrict__ tidx, const float4* __restrict__ tg0, const float4* __restrict__ tg1,
    const float4* __restrict__ tg2, const float4* __restrict__ tg3, const float4* __restrict__ taux,
    double4* __restrict__ X, double4* __restrict__ O, double4* __restrict__ V, const XfBodyParams* __restrict__ bodies,
    float comp1_u, float a_u, float dt, float* __restrict__ clock, uint32_t* __restrict__ arrive,
    uint32_t* __restrict__ err) {
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t target = 0;
    for (uint32_t i = gid; i < n_nodes; i += nthreads) {
        double4 xi = X[i];
        double4 oi = O[i];
        double4 vi = V[i];
        const XfBodyParams& p = bodies[(uint32_t)oi.w >> 1];
        double v[3] = {vi.x, vi.y, vi.z};
        double x[3] = {xi.x, xi.y, xi.z};
        predict(p, v, x, &X[i], &O[i], xi.w, oi.w);
    }
    target += gridDim.x;
    if (!grid_barrier(arrive, target, err)) return;
    for (uint32_t s = 0; s < substeps; s++) {
        for (uint32_t c = 0; c < ncol; c++) {
            if (c == 0 && gid == 0) clock[0] = clock[0] + dt;
            uint32_t begin = colour_begin[c], count = colour_begin[c + 1] - begin;
            for (uint32_t k = gid; k < count; k += nthreads)
                project_one<E, SIM, AUX>(begin + k, tidx, tg0, tg1, tg2, tg3, taux, X, bodies, comp1_u, a_u, dt);
            target += gridDim.x;
            if (!grid_barrier(arrive, target, err)) return;
        }
        float t_end = clock[0];
        bool next = s + 1 < substeps;
        for (uint32_t i = gid; i < n_nodes; i += nthreads) finish_one(i, next, X, O, V, bodies, t_end);
        if (next) {
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

typedef void (*XfProjectFn)(uint32_t, uint32_t, const uint4*, const float4*, const float4*, const float4*,
                            const float4*, const float4*, double4*, const XfBodyParams*, float, float, float, float*);

namespace {
template <int E>
XfProjectFn pick_aux(bool sim, bool aux) {
    if (sim) return aux ? xf_project_colour<E, true, true> : xf_project_colour<E, true, false>;
    return aux ? xf_project_colour<E, false, true> : xf_project_colour<E, false, false>;
}
}  // namespace

// Host-callable: launch one projection kernel. energy 3..7, or 0 for
// per-body MixedSel/YeohSkinFast dispatch.
extern "C" hipError_t xf_launch_project(hipStream_t s, int energy, bool sim, bool aux, uint32_t block,
                                        uint32_t begin, uint32_t count, const uint4* tidx, const float4* tg0,
                                        const float4* tg1, const float4* tg2, const float4* tg3, const float4* taux,
                                        double4* X, const XfBodyParams* bodies, float comp1_u, float a_u, float dt,
                                        float* clock) {
    XfProjectFn fn = nullptr;
    switch (energy) {
        case 0: fn = pick_aux<0>(sim, aux); break;
        case 3: fn = pick_aux<3>(sim, aux); break;
        case 4: fn = pick_aux<4>(sim, aux); break;
        case 5: fn = pick_aux<5>(sim, aux); break;
        case 6: fn = pick_aux<6>(sim, aux); break;
        case 7: fn = pick_aux<7>(sim, aux); break;
        default: return hipErrorInvalidValue;
    }
    if (count == 0) return hipSuccess;
    uint32_t grid = (count + block - 1) / block;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), 0, s, begin, count, tidx, tg0, tg1, tg2, tg3, taux, X, bodies,
                       comp1_u, a_u, dt, clock);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_predict(hipStream_t s, uint32_t n, double4* X, double4* O, const double4* V,
                                        const XfBodyParams* bodies) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_predict, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_finish(hipStream_t s, bool predict_next, uint32_t n, double4* X, double4* O,
                                       double4* V,
                                       const XfBodyParams* bodies, const float* clock) {
    if (n == 0) return hipSuccess;
    if (predict_next)
        hipLaunchKernelGGL(xf_finish<true>, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, clock);
    else
        hipLaunchKernelGGL(xf_finish<false>, dim3((n + 255) / 256), dim3(256), 0, s, n, X, O, V, bodies, clock);
    return hipGetLastError();
}

extern "C" hipError_t xf_launch_count_nonfinite(hipStream_t s, uint32_t n, const double4* X, uint32_t* count) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xf_count_nonfinite, dim3((n + 255) / 256), dim3(256), 0, s, n, X, count);
    return hipGetLastError();
}

typedef void (*XfFrameFn)(uint32_t, uint32_t, const uint32_t*, uint32_t, const uint4*, const float4*, const float4*,
                          const float4*, const float4*, const float4*, double4*, double4*, double4*,
                          const XfBodyParams*, float, float, float, float*, uint32_t*, uint32_t*);

namespace {
template <int E>
XfFrameFn pick_frame(bool sim, bool aux) {
    if (sim) return aux ? xf_frame_persistent<E, true, true> : xf_frame_persistent<E, true, false>;
    return aux ? xf_frame_persistent<E, false, true> : xf_frame_persistent<E, false, false>;
}
XfFrameFn frame_fn(int energy, bool sim, bool aux) {
    switch (energy) {
        case 0: return pick_frame<0>(sim, aux);
        case 3: return pick_frame<3>(sim, aux);
        case 4: return pick_frame<4>(sim, aux);
        case 5: return pick_frame<5>(sim, aux);
        case 6: return pick_frame<6>(sim, aux);
        case 7: return pick_frame<7>(sim, aux);
        default: return nullptr;
    }
}
}  // namespace

// Largest co-resident grid of the persistent kernel at 256 threads per block.
extern "C" hipError_t xf_frame_max_blocks(int energy, bool sim, bool aux, int* max_blocks) {
    XfFrameFn fn = frame_fn(energy, sim, aux);
    if (!fn) return hipErrorInvalidValue;
    int dev = 0, per_cu = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, 256, 0);
    if (e == hipSuccess) *max_blocks = per_cu * cus;
    return e;
}

extern "C" hipError_t xf_launch_frame(hipStream_t s, int energy, bool sim, bool aux, uint32_t blocks,
                                      uint32_t substeps, uint32_t ncol, const uint32_t* colour_begin,
                                      uint32_t n_nodes, const uint4* tidx, const float4* tg0, const float4* tg1,
                                      const float4* tg2, const float4* tg3, const float4* taux, double4* X,
                                      double4* O, double4* V, const XfBodyParams* bodies, float comp1_u, float a_u,
                                      float dt, float* clock, uint32_t* arrive, uint32_t* err) {
    XfFrameFn fn = frame_fn(energy, sim, aux);
    if (!fn) return hipErrorInvalidValue;
    void* args[] = {&substeps, &ncol,  &colour_begin, &n_nodes, &tidx,   &tg0,   &tg1, &tg2,   &tg3, &taux,
                    &X,        &O,     &V,            &bodies,  &comp1_u, &a_u,  &dt,  &clock, &arrive, &err};
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
