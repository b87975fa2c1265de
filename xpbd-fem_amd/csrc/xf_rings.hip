// xf_rings.hip -- gfx950 kernels of the vertex-ring ("continuous") energies:
// Energy_ContinuousPixar (8), Energy_ContinuousMixed (9) and
// Energy_ContinuousSkin (10). For these GeoLinear3d::Constrain sweeps vertex
// rings instead of tets (Geo.cpp:783-786): ring v gathers a volume-weighted
// I1 and J over the tets around v and their gradients on every ring node
// (SolveRing, Geo.cpp:148-188; CalculateElementEnergyAndGradients,
// Fem.cpp:941-973), then solves one or two coupled XPBD constraints over all
// of them (SolveSimplex, Geo.cpp:13-64; Xpbd.h:87-263).
//
// One ring per lane. Rings of one colour share no node (the host colours the
// ring conflict graph), so a colour is one launch with plain gathers and
// scatters, and colours in order equal the reference's serial sweep over the
// colour-major ring order. Each lane keeps its ring's FP64 positions, FP32
// inverse masses and the two FP32 gradient fields in LDS, laid out
// [field][node][lane] so that lanes of a wave touch consecutive words. Sums
// run in the reference's order (ring tets in OneRing order, ring nodes in
// OneRing order), compiled with -ffp-contract=off: a serial CPU sweep of the
// same order reproduces the result bit for bit.
#include <hip/hip_runtime.h>

#include "xf_device.h"
#include "xf_sched.h"

namespace {

// Bytes of LDS per ring node and lane: 3 doubles (position), 6 floats
// (two gradients), 1 float (inverse mass).
constexpr uint32_t kRingLdsBytesPerNode = 3 * 8 + 6 * 4 + 4;

struct RingLds {
    double* xs;  // [3][nmax][B]
    float* gs;   // [2][3][nmax][B]
    float* ws;   // [nmax][B]
    uint32_t nmax, B, tid;
    __device__ double& x(uint32_t c, uint32_t n) const { return xs[(c * nmax + n) * B + tid]; }
    __device__ float& g(uint32_t k, uint32_t c, uint32_t n) const { return gs[((k * 3 + c) * nmax + n) * B + tid]; }
    __device__ float& w(uint32_t n) const { return ws[n * B + tid]; }
};

__device__ __forceinline__ float dot3(const float* a, const float* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ float gdot(const RingLds& L, uint32_t ka, uint32_t kb, uint32_t n) {
    return L.g(ka, 0, n) * L.g(kb, 0, n) + L.g(ka, 1, n) * L.g(kb, 1, n) + L.g(ka, 2, n) * L.g(kb, 2, n);
}

// adjugate(mat3) and determinantFromAdjugate (Fem.cpp:284-306); F[c][r] is
// column c, row r.
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

// Gathers ring energies and gradients (SolveRing's loop, Geo.cpp:163-184):
// returns the ring volume, U[0] = sum w_t I1_t, U[1] = sum w_t J_t, and
// the gradients in L.g(0, ...) and L.g(1, ...).
__device__ float ring_gather(const RingLds& L, uint32_t nn, uint32_t t0, uint32_t nt,
                             const float4* __restrict__ rtet, float U[2]) {
    float vsum = 0.0f;
    for (uint32_t k = 0; k < nt; k++) vsum = vsum + 0.25f * rtet[2 * (t0 + k) + 1].z;
    for (uint32_t n = 0; n < nn; n++)
        for (uint32_t c = 0; c < 3; c++) {
            L.g(0, c, n) = 0.0f;
            L.g(1, c, n) = 0.0f;
        }
    U[0] = 0.0f;
    U[1] = 0.0f;
    for (uint32_t k = 0; k < nt; k++) {
        float4 a = rtet[2 * (t0 + k)], b = rtet[2 * (t0 + k) + 1];
        uint32_t bits = __float_as_uint(b.w);
        uint32_t li[4] = {bits & 255u, (bits >> 8) & 255u, (bits >> 16) & 255u, bits >> 24};
        float P[3][3];
        for (int i = 0; i < 3; i++)
            for (uint32_t r = 0; r < 3; r++) P[i][r] = (float)(L.x(r, li[i]) - L.x(r, li[3]));
        const float QQ[3] = {a.x, a.y, a.z}, QR[3] = {a.w, b.x, b.y};
        float vt = b.z;
        float ew = (vt / 4.0f) / vsum;
        // CalculateIncompressibleNeoHookeanEnergyAndGradients (Fem.cpp:164-190)
        float I1 = 0.0f, gI[4][3];
        for (int i = 0; i < 3; i++) {
            I1 = I1 + QQ[i] * dot3(P[i], P[i]);
            for (int r = 0; r < 3; r++) gI[i][r] = (2.0f * QQ[i]) * P[i][r];
        }
        int q = 0;
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++) {
                I1 = I1 + QR[q] * dot3(P[i], P[j]);
                for (int r = 0; r < 3; r++) gI[i][r] = gI[i][r] + QR[q] * P[j][r];
                for (int r = 0; r < 3; r++) gI[j][r] = gI[j][r] + QR[q] * P[i][r];
                ++q;
            }
        for (int r = 0; r < 3; r++) gI[3][r] = ((0.0f - gI[0][r]) - gI[1][r]) - gI[2][r];
        U[0] = U[0] + ew * I1;
        for (int n = 0; n < 4; n++)
            for (uint32_t r = 0; r < 3; r++) L.g(0, r, li[n]) = L.g(0, r, li[n]) + ew * gI[n][r];
        // J of the tet from mat3 PP with rows P_i (Fem.cpp:960-970)
        float M[3][3], adj[3][3];
        for (int j = 0; j < 3; j++)
            for (int r = 0; r < 3; r++) M[j][r] = P[r][j];
        float det = adjdet(M, adj);
        float J = ((1.0f / vt) * (1.0f / 6.0f)) * det;
        U[1] = U[1] + ew * J;
        float sJ = (ew / vt) * (1.0f / 6.0f);
        for (int n = 0; n < 3; n++)
            for (uint32_t r = 0; r < 3; r++) L.g(1, r, li[n]) = L.g(1, r, li[n]) + sJ * adj[n][r];
        for (uint32_t r = 0; r < 3; r++) {
            float a3 = ((-adj[0][r]) - adj[1][r]) - adj[2][r];
            L.g(1, r, li[3]) = L.g(1, r, li[3]) + sJ * a3;
        }
    }
    return vsum;
}

// EnergyXpbdConstrain over the ring (Xpbd.h:87-120), Rayleigh Paper / Limit
// inside the projection when damp > 0 and rtype < 2 (reads O).
__device__ void ring_constrain1(const RingLds& L, uint32_t nn, const uint32_t* __restrict__ nodes,
                                const double4* __restrict__ O, float U, float dt, float comp, float damp,
                                uint32_t rtype) {
    float alpha = comp / (dt * dt);
    float lp;
    if (damp > 0.0f && rtype < 2u) {
        float gamma = damp / dt;
        float wgg = 1.0e-22f, gV = 0.0f;
        for (uint32_t n = 0; n < nn; n++) {
            double4 o = O[nodes[n]];
            float v[3] = {(float)(L.x(0, n) - o.x), (float)(L.x(1, n) - o.y), (float)(L.x(2, n) - o.z)};
            float g[3] = {L.g(0, 0, n), L.g(0, 1, n), L.g(0, 2, n)};
            wgg = wgg + L.w(n) * dot3(g, g);
            gV = gV + dot3(v, g);
        }
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
    } else {
        float wgg = 1.0e-22f;
        for (uint32_t n = 0; n < nn; n++) wgg = wgg + L.w(n) * gdot(L, 0, 0, n);
        lp = (-2.0f * U) / (wgg + (2.0f * U * alpha));
    }
    for (uint32_t n = 0; n < nn; n++) {
        float s = L.w(n) * lp;
        for (uint32_t r = 0; r < 3; r++) L.x(r, n) = L.x(r, n) + (double)(s * L.g(0, r, n));
    }
}

// EnergyXpbdConstrainSimultaneous<2> over the ring (Xpbd.h:123-214).
__device__ void ring_constrain2(const RingLds& L, uint32_t nn, const uint32_t* __restrict__ nodes,
                                const double4* __restrict__ O, const float U[2], float dt, const float comp[2],
                                float damp, uint32_t rtype) {
    float alpha[2] = {comp[0] / (dt * dt), comp[1] / (dt * dt)};
    bool damped = damp > 0.0f && rtype < 2u;
    float gamma = damp / dt;
    float gV[2] = {0.0f, 0.0f};
    if (damped)
        for (uint32_t n = 0; n < nn; n++) {
            double4 o = O[nodes[n]];
            float v[3] = {(float)(L.x(0, n) - o.x), (float)(L.x(1, n) - o.y), (float)(L.x(2, n) - o.z)};
            float g0[3] = {L.g(0, 0, n), L.g(0, 1, n), L.g(0, 2, n)};
            float g1[3] = {L.g(1, 0, n), L.g(1, 1, n), L.g(1, 2, n)};
            gV[0] = gV[0] + dot3(g0, v);
            gV[1] = gV[1] + dot3(g1, v);
        }
    float wgg[3] = {1.0e-22f, 1.0e-22f, 1.0e-22f};
    for (uint32_t n = 0; n < nn; n++) wgg[0] = wgg[0] + L.w(n) * gdot(L, 0, 0, n);
    for (uint32_t n = 0; n < nn; n++) wgg[1] = wgg[1] + L.w(n) * gdot(L, 1, 0, n);
    for (uint32_t n = 0; n < nn; n++) wgg[2] = wgg[2] + L.w(n) * gdot(L, 1, 1, n);
    float A[3] = {wgg[0], wgg[1], wgg[2]}, b[2];
    for (int i = 0; i < 2; i++) {
        int d = 2 * i;
        if (damped && rtype == 0u && gamma > 0.0f) {
            A[d] = A[d] + (2.0f * U[i] * alpha[i]) / (1.0f + gamma);
            b[i] = (-2.0f * U[i] - gamma * gV[i]) / (1.0f + gamma);
        } else {
            A[d] = A[d] + 2.0f * U[i] * alpha[i];
            b[i] = -2.0f * U[i];
        }
    }
    // Cramer's rule (Xpbd.h:173-179)
    float invA00 = 1.0f / A[0];
    float invA11 = 1.0f / A[2];
    float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A[1] * invA00) * (A[1] * invA11));
    float l0 = invDet * ((b[0] * invA00) - (A[1] * invA00) * (b[1] * invA11));
    float l1 = invDet * ((b[1] * invA11) - (b[0] * invA00) * (A[1] * invA11));
    if (damped && rtype == 1u && gamma > 0.0f) {
        float lam[2] = {l0, l1};
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
    for (uint32_t n = 0; n < nn; n++) {
        float wn = L.w(n);
        for (uint32_t r = 0; r < 3; r++) {
            float acc = 0.0f + l0 * L.g(0, r, n);
            acc = acc + l1 * L.g(1, r, n);
            L.x(r, n) = L.x(r, n) + (double)(wn * acc);
        }
    }
}

// RayleighDamp over the ring, one constraint (Xpbd.h:216-230).
__device__ void ring_rayleigh1(const RingLds& L, uint32_t nn, const uint32_t* __restrict__ nodes,
                               double4* __restrict__ V, float U, float dt, float comp, float damp) {
    float invBeta = comp / (dt * damp);
    float wgg = 1.0e-22f, gV = 0.0f;
    for (uint32_t n = 0; n < nn; n++) {
        double4 vd = V[nodes[n]];
        float v[3] = {(float)vd.x, (float)vd.y, (float)vd.z};
        float g[3] = {L.g(0, 0, n), L.g(0, 1, n), L.g(0, 2, n)};
        wgg = wgg + L.w(n) * dot3(g, g);
        gV = gV + dot3(g, v);
    }
    float lambda = -gV / (2.0f * U * invBeta + wgg);
    for (uint32_t n = 0; n < nn; n++) {
        float s = L.w(n) * lambda;
        double4 vd = V[nodes[n]];
        vd.x = vd.x + (double)(s * L.g(0, 0, n));
        vd.y = vd.y + (double)(s * L.g(0, 1, n));
        vd.z = vd.z + (double)(s * L.g(0, 2, n));
        V[nodes[n]] = vd;
    }
}

// RayleighDamp over the ring, two coupled constraints (Xpbd.h:232-263).
__device__ void ring_rayleigh2(const RingLds& L, uint32_t nn, const uint32_t* __restrict__ nodes,
                               double4* __restrict__ V, const float U[2], float dt, const float comp[2], float damp) {
    float invBeta0 = comp[0] / (dt * damp), invBeta1 = comp[1] / (dt * damp);
    float A0 = 1.0e-22f, A1 = 1.0e-22f, A2 = 1.0e-22f, b0 = 0.0f, b1 = 0.0f;
    for (uint32_t n = 0; n < nn; n++) {
        float wn = L.w(n);
        A0 = A0 + wn * gdot(L, 0, 0, n);
        A1 = A1 + wn * gdot(L, 0, 1, n);
        A2 = A2 + wn * gdot(L, 1, 1, n);
        double4 vd = V[nodes[n]];
        float v[3] = {(float)vd.x, (float)vd.y, (float)vd.z};
        float g0[3] = {L.g(0, 0, n), L.g(0, 1, n), L.g(0, 2, n)};
        float g1[3] = {L.g(1, 0, n), L.g(1, 1, n), L.g(1, 2, n)};
        b0 = b0 - dot3(g0, v);
        b1 = b1 - dot3(g1, v);
    }
    A0 = A0 + 2.0f * U[0] * invBeta0;
    A2 = A2 + 2.0f * U[1] * invBeta1;
    float invA00 = 1.0f / A0;
    float invA11 = 1.0f / A2;
    float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A1 * invA00) * (A1 * invA11));
    float l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
    float l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
    for (uint32_t n = 0; n < nn; n++) {
        float wn = L.w(n);
        double4 vd = V[nodes[n]];
        vd.x = vd.x + (double)(wn * (l0 * L.g(0, 0, n) + l1 * L.g(1, 0, n)));
        vd.y = vd.y + (double)(wn * (l0 * L.g(0, 1, n) + l1 * L.g(1, 1, n)));
        vd.z = vd.z + (double)(wn * (l0 * L.g(0, 2, n) + l1 * L.g(1, 2, n)));
        V[nodes[n]] = vd;
    }
}

// One colour of the ring sweep. POST = false: project positions
// (GeoLinear3d::Constrain), the first colour of a substep advancing the
// clock (clock != nullptr). POST = true: the Rayleigh post pass on V
// (GeoLinear3d::Damp, Geo.cpp:802-804) of bodies with type Post or
// PostAmortized and damping > 0, over the amortized eighth of the body's
// ring order when PostAmortized (damping x 8).
template <bool POST>
__global__ __launch_bounds__(256) void xf_ring_colour(
    uint32_t begin, uint32_t count, uint32_t slabs, uint32_t nmax, const uint4* __restrict__ rmeta,
    const uint2* __restrict__ rinfo, const uint32_t* __restrict__ rnodes, const float4* __restrict__ rtet,
    double4* __restrict__ X, const double4* __restrict__ O, double4* __restrict__ V,
    const XfBodyParams* __restrict__ bodies, float dt, float* __restrict__ clock) {
    extern __shared__ double xf_ring_lds[];
    if (!POST && clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    uint32_t i = xf::colour_slot(begin, count, slabs);
    if (i == xf::kNoSlot) return;
    uint4 m = rmeta[i];
    uint2 info = rinfo[i];
    const XfBodyParams& bp = bodies[info.x];
    uint32_t energy = bp.energy, rtype = bp.rtype;
    float damp = bp.damping;
    if (POST) {
        if (rtype < 2u || !(damp > 0.0f)) return;
        if (rtype == 3u) {
            uint32_t k8 = (reinterpret_cast<const uint32_t*>(clock)[1] - 1u) % 8u;
            uint32_t lo = bp.n_rings * k8 / 8u, hi = bp.n_rings * (k8 + 1u) / 8u;
            if (info.y < lo || info.y >= hi) return;
            damp = damp * 8.0f;
        }
    }
    const uint32_t B = blockDim.x, tid = threadIdx.x;
    RingLds L;
    L.nmax = nmax;
    L.B = B;
    L.tid = tid;
    L.xs = xf_ring_lds;
    L.gs = reinterpret_cast<float*>(xf_ring_lds + 3 * nmax * B);
    L.ws = L.gs + 6 * nmax * B;
    const uint32_t nn = m.y;
    const uint32_t* nodes = rnodes + m.x;
    for (uint32_t n = 0; n < nn; n++) {
        double4 x = X[nodes[n]];
        L.x(0, n) = x.x;
        L.x(1, n) = x.y;
        L.x(2, n) = x.z;
        L.w(n) = (float)x.w;
    }
    float U[2];
    float volume = ring_gather(L, nn, m.z, m.w, rtet, U);
    float I1 = U[0], J = U[1];
    float nu = bp.poisson, compliance = bp.compliance;
    if (energy == 8u) {
        // ContinuousPixar: one constraint (Geo.cpp:21-35)
        float mu = 2.0f;
        float lambda = mu / (1.0f - 2.0f * nu);
        float a = 1.0f + mu / lambda;
        float comp = compliance / volume;
        float U0 = fmaxf(0.0000001f, (mu / 2.0f) * I1 + (lambda / 2.0f) * (J - a) * (J - a));
        for (uint32_t n = 0; n < nn; n++)
            for (uint32_t r = 0; r < 3; r++)
                L.g(0, r, n) = (mu / 2.0f) * L.g(0, r, n) + (lambda / 2.0f) * 2.0f * (J - a) * L.g(1, r, n);
        if (POST) ring_rayleigh1(L, nn, nodes, V, U0, dt, comp, damp);
        else ring_constrain1(L, nn, nodes, O, U0, dt, comp, damp, rtype);
    } else {
        // ContinuousMixed / ContinuousSkin: two coupled constraints (Geo.cpp:36-63)
        float mu = 1.0f / compliance;
        float lambda = (2.0f * mu * nu) / (1.0f - 2.0f * nu);
        float a = 1.0f + mu / lambda;
        float comp[2] = {1.0f / (mu * volume), 1.0f / (lambda * volume)};
        if (energy == 10u) {
            float C0 = bp.yeoh[0], C1 = bp.yeoh[1], C2 = bp.yeoh[2];
            float IM = I1 - 3.0f;
            U[0] = fmaxf(0.0001f, C0 * IM + C1 * IM * IM + C2 * IM * IM * IM);
            float sc = C0 + 2.0f * C1 * IM + 3.0f * C2 * IM * IM;
            for (uint32_t n = 0; n < nn; n++)
                for (uint32_t r = 0; r < 3; r++) L.g(0, r, n) = L.g(0, r, n) * sc;
        }
        U[1] = (J - a) * (J - a);
        float s1 = 2.0f * (J - a);
        for (uint32_t n = 0; n < nn; n++)
            for (uint32_t r = 0; r < 3; r++) L.g(1, r, n) = L.g(1, r, n) * s1;
        if (POST) ring_rayleigh2(L, nn, nodes, V, U, dt, comp, damp);
        else ring_constrain2(L, nn, nodes, O, U, dt, comp, damp, rtype);
    }
    if (!POST)
        for (uint32_t n = 0; n < nn; n++) {
            double4 x = X[nodes[n]];
            x.x = L.x(0, n);
            x.y = L.x(1, n);
            x.z = L.x(2, n);
            X[nodes[n]] = x;
        }
}

}  // namespace

// LDS bytes of a ring launch: nmax nodes per lane, block lanes.
extern "C" size_t xf_ring_lds_bytes(uint32_t nmax, uint32_t block) {
    return (size_t)kRingLdsBytesPerNode * nmax * block;
}

// One colour of the ring sweep (post = false) or of the ring damping pass
// (post = true). clock: advanced by the first colour of a substep (pass it
// for that launch only); read by the damping pass for the amortized range.
extern "C" hipError_t xf_launch_rings(hipStream_t s, bool post, uint32_t block, uint32_t slabs, uint32_t begin,
                                      uint32_t count, uint32_t nmax, const uint4* rmeta, const uint2* rinfo,
                                      const uint32_t* rnodes, const float4* rtet, double4* X, const double4* O,
                                      double4* V, const XfBodyParams* bodies, float dt, float* clock) {
    if (count == 0) return hipSuccess;
    size_t lds = xf_ring_lds_bytes(nmax, block);
    dim3 grid(xf::colour_grid(count, block, slabs));
    if (post)
        hipLaunchKernelGGL(xf_ring_colour<true>, grid, dim3(block), lds, s, begin, count, slabs, nmax, rmeta, rinfo,
                           rnodes, rtet, X, O, V, bodies, dt, clock);
    else
        hipLaunchKernelGGL(xf_ring_colour<false>, grid, dim3(block), lds, s, begin, count, slabs, nmax, rmeta,
                           rinfo, rnodes, rtet, X, O, V, bodies, dt, clock);
    return hipGetLastError();
}
