// xf_rings.hip -- gfx950 kernels of the vertex-ring ("continuous") energies:
// Energy_ContinuousPixar (8), Energy_ContinuousMixed (9) and
// Energy_ContinuousSkin (10). For these GeoLinear3d::Constrain sweeps vertex
// rings instead of tets (Geo.cpp:783-786): ring v gathers a volume-weighted
// I1 and J over the tets around v and their gradients on every ring node
// (SolveRing, Geo.cpp:148-188; CalculateElementEnergyAndGradients,
// Fem.cpp:941-973), then solves one or two coupled XPBD constraints over all
// of them (SolveSimplex, Geo.cpp:13-64; Xpbd.h:87-263).
//
// Rings of one colour share no node (the host colours the ring conflict
// graph), so a colour is one launch with plain gathers and scatters, and
// colours in order equal the reference's serial sweep over the colour-major
// ring order. Inside a launch a team of 16 or 32 lanes (a launch option)
// works on one ring, in LDS, in phases separated by workgroup barriers:
//   0  load the ring's FP64 node positions and its incidence lists (the ring
//      volume, a sum of rest volumes, comes precomputed in the same order);
//   1  per ring tet (lanes in parallel): I1 and J of the tet and their
//      weighted gradient contributions at its 4 corners;
//   2  per ring node (lanes in parallel): the contributions of the node's
//      tets added in ring-tet order (host-built incidence lists; phase 1
//      wrote them in that order); two lanes add the two energy terms in
//      ring-tet order;
//   3  per ring node: SolveSimplex's scaling; then the terms of the
//      constraint sums (all lanes), one lane per sum adding them in ring-node
//      order, and one lane the Lagrange multipliers;
//   4  per ring node: the position (or velocity) update.
// Every sum keeps the reference's order (ring tets in OneRing order, ring
// nodes in OneRing order), compiled with -ffp-contract=off: a serial CPU
// sweep of the same ring order reproduces the result bit for bit.
#include <hip/hip_runtime.h>

#include "xf_device.h"
#include "xf_element.h"
#include "xf_sched.h"

namespace {

// Floats per ring tet in LDS: weighted I1 and J, then the I1 and J gradient
// contributions at its 4 corners (4 x 6), stored at the corners' entries of
// the incidence lists.
constexpr uint32_t kTetFloats = 2 + 12 + 12;
constexpr uint32_t kScalars = 16;
// Incidence entries per lane whose loads phase 0 issues up front.
constexpr uint32_t kIncLoads = 4;

// LDS of one ring team: positions [3][nmax] double, inverse masses [nmax],
// gradients [6][nmax] (g0 xyz, g1 xyz), per-node velocity-like vectors
// [3][nmax] (X - O, or V), per-tet contributions [tmax * kTetFloats]
// (energy terms [tmax][2], then gradient contributions [4 tmax][6] in
// incidence-entry order; reused for the constraint-sum terms), scalars
// [kScalars], incidence lists: per-node starts [nmax + 1], entries [4 tmax]
// and their inverse [4 tmax] (uint16).
__host__ __device__ inline size_t ring_team_bytes(uint32_t nmax, uint32_t tmax) {
    size_t b = 8 * 3 * (size_t)nmax + 4 * ((size_t)nmax * (1 + 6 + 3) + (size_t)tmax * kTetFloats + kScalars) +
               2 * ((size_t)nmax + 1 + 8 * (size_t)tmax);
    return (b + 7) & ~(size_t)7;
}

struct Team {
    double* xs;
    float *ws, *g, *v, *c, *sc;
    uint16_t *incb, *inc, *inv;
    uint32_t nmax;
    __device__ double& x(uint32_t r, uint32_t n) const { return xs[r * nmax + n]; }
    __device__ float& gr(uint32_t k, uint32_t r, uint32_t n) const { return g[(k * 3 + r) * nmax + n]; }
    __device__ float& vr(uint32_t r, uint32_t n) const { return v[r * nmax + n]; }
};

using xf::adjdet;
using xf::dot3;

// Phase 1 for ring tet k (CalculateElementEnergyAndGradients for a linear
// tet, Fem.cpp:941-973): c[2 k] = w_t I1, c[2 k + 1] = w_t J and, for corner
// j at incidence entry e = inv[4 k + j], c[2 tmax + 6 e + r] = w_t dI1/dx_j
// and c[2 tmax + 6 e + 3 + r] = w_t dJ/dx_j, with w_t = (V_t / 4) / V. The
// node sums of phase 2 then read each node's entries contiguously.
__device__ void ring_tet(const Team& T, float4 a, float4 b, float vsum, uint32_t k, uint32_t tmax) {
    uint32_t bits = __float_as_uint(b.w);
    uint32_t li[4] = {bits & 255u, (bits >> 8) & 255u, (bits >> 16) & 255u, bits >> 24};
    float P[3][3];
    for (int i = 0; i < 3; i++)
        for (uint32_t r = 0; r < 3; r++) P[i][r] = (float)(T.x(r, li[i]) - T.x(r, li[3]));
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
    // energy terms at c[2 k + q]; corner j's gradients at the entry the
    // incidence lists give (tet k, corner j): c[2 tmax + 6 e + 0..5]
    float* slot[4];
    for (int n = 0; n < 4; n++) slot[n] = T.c + 2 * tmax + 6 * (uint32_t)T.inv[4 * k + n];
    T.c[2 * k] = ew * I1;
    for (int n = 0; n < 4; n++)
        for (int r = 0; r < 3; r++) slot[n][r] = ew * gI[n][r];
    // J of the tet from mat3 PP with rows P_i (Fem.cpp:960-970)
    float M[3][3], adj[3][3];
    for (int j = 0; j < 3; j++)
        for (int r = 0; r < 3; r++) M[j][r] = P[r][j];
    float det = adjdet(M, adj);
    float J = ((1.0f / vt) * (1.0f / 6.0f)) * det;
    T.c[2 * k + 1] = ew * J;
    float sJ = (ew / vt) * (1.0f / 6.0f);
    for (int n = 0; n < 3; n++)
        for (int r = 0; r < 3; r++) slot[n][3 + r] = sJ * adj[n][r];
    for (int r = 0; r < 3; r++) slot[3][3 + r] = sJ * (((-adj[0][r]) - adj[1][r]) - adj[2][r]);
}

// One colour of the ring sweep. POST = false: project positions
// (GeoLinear3d::Constrain), the first colour of a substep advancing the
// clock (clock != nullptr). POST = true: the Rayleigh post pass on V
// (GeoLinear3d::Damp, Geo.cpp:802-804) of bodies with type Post or
// PostAmortized and damping > 0, over the amortized eighth of the body's
// ring order when PostAmortized (damping x 8).
template <bool POST, uint32_t LANES>
__global__ __launch_bounds__(256) void xf_ring_colour(
    uint32_t begin, uint32_t count, uint32_t slabs, uint32_t nmax, uint32_t tmax, const uint4* __restrict__ rmeta,
    const uint4* __restrict__ rinfo, const uint32_t* __restrict__ rnodes, const uint16_t* __restrict__ rincb,
    const uint16_t* __restrict__ rinc, const float4* __restrict__ rtet, double4* __restrict__ X,
    const double4* __restrict__ O, double4* __restrict__ V, const XfBodyParams* __restrict__ bodies, float dt,
    float* __restrict__ clock) {
    extern __shared__ double xf_ring_lds[];
    if (!POST && clock != nullptr && blockIdx.x == 0 && threadIdx.x == 0) {
        clock[0] = clock[0] + dt;
        uint32_t* tick = reinterpret_cast<uint32_t*>(clock) + 1;
        tick[0] = tick[0] + 1u;
    }
    const uint32_t lane = threadIdx.x % LANES, team = threadIdx.x / LANES;
    Team T;
    T.nmax = nmax;
    T.xs = reinterpret_cast<double*>(reinterpret_cast<char*>(xf_ring_lds) + team * ring_team_bytes(nmax, tmax));
    T.ws = reinterpret_cast<float*>(T.xs + 3 * nmax);
    T.g = T.ws + nmax;
    T.v = T.g + 6 * nmax;
    T.c = T.v + 3 * nmax;
    T.sc = T.c + tmax * kTetFloats;
    T.incb = reinterpret_cast<uint16_t*>(T.sc + kScalars);
    T.inc = T.incb + nmax + 1;
    T.inv = T.inc + 4 * tmax;

    // Every lane reaches every barrier; a team without a ring (colour tail,
    // or a body the damping pass skips) has nn = nt = 0.
    uint32_t i = xf::colour_slot(begin, count, slabs, LANES);
    uint4 m = make_uint4(0, 0, 0, 0);
    float volume = 0.0f;
    uint32_t energy = 0, rtype = 0;
    float damp = 0.0f, nu = 0.0f, compliance = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    if (i != xf::kNoSlot) {
        m = rmeta[i];
        uint4 info = rinfo[i];
        volume = __uint_as_float(info.z);
        const XfBodyParams& bp = bodies[info.x];
        energy = bp.energy;
        rtype = bp.rtype;
        damp = bp.damping;
        nu = bp.poisson;
        compliance = bp.compliance;
        C0 = bp.yeoh[0];
        C1 = bp.yeoh[1];
        C2 = bp.yeoh[2];
        if (POST) {
            bool on = rtype >= 2u && damp > 0.0f;
            if (on && rtype == 3u) {
                uint32_t k8 = (reinterpret_cast<const uint32_t*>(clock)[1] - 1u) % 8u;
                uint32_t lo = bp.n_rings * k8 / 8u, hi = bp.n_rings * (k8 + 1u) / 8u;
                on = info.y >= lo && info.y < hi;
                damp = damp * 8.0f;
            }
            if (!on) m = make_uint4(0, 0, 0, 0);
        }
    }
    const uint32_t nn = m.y, nt = m.w;
    const uint32_t* nodes = rnodes + m.x;

    // Phase 0: positions, inverse masses and incidence lists. The loads of a
    // lane's first node, its first kIncLoads incidence entries and its first
    // ring tet (for phase 1) are issued before any of them is used, at
    // clamped in-range positions (a team without a ring reads entry 0), so
    // the phase waits for two dependent round trips (node id, position), not
    // one per loop iteration; rings beyond those counts loop as before.
    const uint32_t ne = 4u * nt;
    const uint32_t n0 = min(lane, nn > 0u ? nn - 1u : 0u);
    const uint32_t g0 = nodes[n0];
    const uint16_t b0 = rincb[m.x + n0];
    uint16_t code[kIncLoads];
#pragma unroll
    for (uint32_t u = 0; u < kIncLoads; u++) code[u] = rinc[4 * m.z + min(lane + u * LANES, ne > 0u ? ne - 1u : 0u)];
    const uint32_t k0 = min(lane, nt > 0u ? nt - 1u : 0u);
    float4 ta = rtet[2 * (m.z + k0)], tb = rtet[2 * (m.z + k0) + 1];
    const double4 x0 = X[g0];
    if (lane < nn) {
        T.x(0, lane) = x0.x;
        T.x(1, lane) = x0.y;
        T.x(2, lane) = x0.z;
        T.ws[lane] = (float)x0.w;
        T.incb[lane] = b0;
    }
    for (uint32_t n = lane + LANES; n < nn; n += LANES) {
        double4 x = X[nodes[n]];
        T.x(0, n) = x.x;
        T.x(1, n) = x.y;
        T.x(2, n) = x.z;
        T.ws[n] = (float)x.w;
        T.incb[n] = rincb[m.x + n];
    }
    if (lane == 0) T.incb[nn] = (uint16_t)ne;
#pragma unroll
    for (uint32_t u = 0; u < kIncLoads; u++) {
        const uint32_t e = lane + u * LANES;
        if (e < ne) {
            T.inc[e] = code[u];
            T.inv[code[u]] = (uint16_t)e;  // entry e of the lists holds (tet, corner) = code
        }
    }
    for (uint32_t e = lane + kIncLoads * LANES; e < ne; e += LANES) {
        const uint16_t c = rinc[4 * m.z + e];
        T.inc[e] = c;
        T.inv[c] = (uint16_t)e;
    }
    __syncthreads();

    // Phase 1: per ring tet (the lane's first from the phase-0 loads).
    for (uint32_t k = lane; k < nt; k += LANES) {
        if (k != lane) {
            ta = rtet[2 * (m.z + k)];
            tb = rtet[2 * (m.z + k) + 1];
        }
        ring_tet(T, ta, tb, volume, k, tmax);
    }
    __syncthreads();

    // Phase 2: ring-node gradients in ring-tet order (a node's entries are
    // contiguous: no dependent index load); energy terms.
    for (uint32_t n = lane; n < nn; n += LANES) {
        const uint32_t e0 = T.incb[n], e1 = T.incb[n + 1];
        const float* sl = T.c + 2 * tmax;
        float a[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
        for (uint32_t e = e0; e < e1; e++)
#pragma unroll
            for (uint32_t r = 0; r < 6; r++) a[r] = a[r] + sl[6 * e + r];
        for (uint32_t r = 0; r < 3; r++) {
            T.gr(0, r, n) = a[r];
            T.gr(1, r, n) = a[3 + r];
        }
    }
    if (lane >= LANES - 2 && nt > 0) {
        uint32_t q = lane - (LANES - 2);  // 0: I1 term, 1: J term
        float u = 0.0f;
#pragma unroll 4
        for (uint32_t k = 0; k < nt; k++) u = u + T.c[2 * k + q];
        T.sc[1 + q] = u;
    }
    __syncthreads();

    // Phase 3: SolveSimplex (Geo.cpp:13-64): constants, gradient scaling and
    // the per-node vectors of the damping terms.
    float I1 = 0.0f, J = 0.0f, U0 = 0.0f, U1 = 0.0f, comp0 = 0.0f, comp1 = 0.0f;
    bool two = energy != 8u;
    bool damped_proj = !POST && damp > 0.0f && rtype < 2u;
    if (nt > 0) {
        I1 = T.sc[1];
        J = T.sc[2];
        if (!two) {
            float mu = 2.0f;
            float lambda = mu / (1.0f - 2.0f * nu);
            float a = 1.0f + mu / lambda;
            comp0 = compliance / volume;
            U0 = fmaxf(0.0000001f, (mu / 2.0f) * I1 + (lambda / 2.0f) * (J - a) * (J - a));
            for (uint32_t n = lane; n < nn; n += LANES)
                for (uint32_t r = 0; r < 3; r++)
                    T.gr(0, r, n) = (mu / 2.0f) * T.gr(0, r, n) + (lambda / 2.0f) * 2.0f * (J - a) * T.gr(1, r, n);
        } else {
            float mu = 1.0f / compliance;
            float lambda = (2.0f * mu * nu) / (1.0f - 2.0f * nu);
            float a = 1.0f + mu / lambda;
            comp0 = 1.0f / (mu * volume);
            comp1 = 1.0f / (lambda * volume);
            U0 = I1;
            float sc = 1.0f;
            bool skin = energy == 10u;
            if (skin) {
                float IM = I1 - 3.0f;
                U0 = fmaxf(0.0001f, C0 * IM + C1 * IM * IM + C2 * IM * IM * IM);
                sc = C0 + 2.0f * C1 * IM + 3.0f * C2 * IM * IM;
            }
            U1 = (J - a) * (J - a);
            float s1 = 2.0f * (J - a);
            for (uint32_t n = lane; n < nn; n += LANES)
                for (uint32_t r = 0; r < 3; r++) {
                    if (skin) T.gr(0, r, n) = T.gr(0, r, n) * sc;
                    T.gr(1, r, n) = T.gr(1, r, n) * s1;
                }
        }
        if (POST || damped_proj)
            for (uint32_t n = lane; n < nn; n += LANES) {
                if (POST) {
                    double4 vd = V[nodes[n]];
                    T.vr(0, n) = (float)vd.x;
                    T.vr(1, n) = (float)vd.y;
                    T.vr(2, n) = (float)vd.z;
                } else {
                    double4 o = O[nodes[n]];
                    for (uint32_t r = 0; r < 3; r++) T.vr(r, n) = (float)(T.x(r, n) - (r == 0 ? o.x : r == 1 ? o.y : o.z));
                }
            }
    }
    __syncthreads();

    // Phase 3b: the constraint sums, each in ring-node order (the sums of
    // Xpbd.h:87-263):
    //   sc[5] = 1e-22 + sum w g0.g0     sc[6] = 1e-22 + sum w g1.g0
    //   sc[7] = 1e-22 + sum w g1.g1     sc[8] = sum +-g0.v, sc[9] = sum +-g1.v
    // with v = X - O (projection) or V (damping pass, where the two-constraint
    // RayleighDamp sums -g.v). Their terms are formed by all lanes (one
    // (sum, node) pair at a time) into the per-tet area, free after phase 2;
    // then one lane per sum adds them in order. Multiplying by 1 and negating
    // are exact, so each sum rounds as the reference's does.
    const uint32_t nsum = POST || damped_proj ? 5u : 3u;
    float* terms = T.c;  // nsum x nn floats (nn <= 3 nt + 1 fits the 26 nt floats)
    if (nt > 0)
        for (uint32_t it = lane; it < nsum * nn; it += LANES) {
            const uint32_t q = it / nn, n = it - q * nn;
            const float* a = T.g + (q == 1 || q == 2 || q == 4 ? 3 * nmax : 0);
            const float* b = q <= 1 ? T.g : q == 2 ? T.g + 3 * nmax : T.v;
            float d = a[n] * b[n] + a[nmax + n] * b[nmax + n] + a[2 * nmax + n] * b[2 * nmax + n];
            float term = q <= 2 ? T.ws[n] * d : d;
            terms[it] = q >= 3 && POST && two ? -1.0f * term : term;
        }
    __syncthreads();
    if (nt > 0 && lane < nsum) {
        const float* tq = terms + lane * nn;
        float acc = lane <= 2 ? 1.0e-22f : 0.0f;
#pragma unroll 8
        for (uint32_t n = 0; n < nn; n++) acc = acc + tq[n];
        T.sc[5 + lane] = acc;
    }
    __syncthreads();

    // Phase 3c: one lane forms the Lagrange multipliers.
    if (lane == 0 && nt > 0) {
        float l0 = 0.0f, l1 = 0.0f;
        float wgg[3] = {T.sc[5], T.sc[6], T.sc[7]};
        float gV[2] = {T.sc[8], T.sc[9]};
        if (!two && !POST) {
            // EnergyXpbdConstrain (Xpbd.h:87-120)
            float alpha = comp0 / (dt * dt);
            if (damped_proj) {
                float gamma = damp / dt;
                if (rtype == 0u) {
                    float A = wgg[0] * (1.0f + gamma) + (2.0f * U0 * alpha);
                    l0 = (-2.0f * U0 - gamma * gV[0]) / A;
                } else {
                    l0 = (-2.0f * U0) / (wgg[0] + (2.0f * U0 * alpha));
                    float invBeta = comp0 / (dt * damp);
                    float A = wgg[0] + 2.0f * U0 * invBeta;
                    float b = -gV[0] - wgg[0] * l0;
                    b = b * (A / fmaxf(A, 4.0f * wgg[0]));
                    l0 = l0 + b / A;
                }
            } else {
                l0 = (-2.0f * U0) / (wgg[0] + (2.0f * U0 * alpha));
            }
        } else if (!POST) {
            // EnergyXpbdConstrainSimultaneous<2> (Xpbd.h:123-214)
            float U[2] = {U0, U1}, comp[2] = {comp0, comp1};
            float alpha[2] = {comp0 / (dt * dt), comp1 / (dt * dt)};
            float gamma = damp / dt;
            float A[3] = {wgg[0], wgg[1], wgg[2]}, b[2];
            for (int k = 0; k < 2; k++) {
                int d = 2 * k;
                if (damped_proj && rtype == 0u && gamma > 0.0f) {
                    A[d] = A[d] + (2.0f * U[k] * alpha[k]) / (1.0f + gamma);
                    b[k] = (-2.0f * U[k] - gamma * gV[k]) / (1.0f + gamma);
                } else {
                    A[d] = A[d] + 2.0f * U[k] * alpha[k];
                    b[k] = -2.0f * U[k];
                }
            }
            // Cramer's rule (Xpbd.h:173-179)
            float invA00 = 1.0f / A[0];
            float invA11 = 1.0f / A[2];
            float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A[1] * invA00) * (A[1] * invA11));
            l0 = invDet * ((b[0] * invA00) - (A[1] * invA00) * (b[1] * invA11));
            l1 = invDet * ((b[1] * invA11) - (b[0] * invA00) * (A[1] * invA11));
            if (damped_proj && rtype == 1u && gamma > 0.0f) {
                float lam[2] = {l0, l1};
                for (int k = 0; k < 2; k++) {
                    int d = 2 * k;
                    float invBeta = comp[k] / (dt * damp);
                    A[d] = wgg[d] + 2.0f * U[k] * invBeta;
                    b[k] = b[k] * (-1.0f - alpha[k] * lam[k]);
                    b[k] = b[k] + -gV[k];
                    b[k] = b[k] * (A[d] / fmaxf(A[d], 8.0f * wgg[d]));
                }
                invA00 = 1.0f / A[0];
                invA11 = 1.0f / A[2];
                invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A[1] * invA00) * (A[1] * invA11));
                l0 = l0 + invDet * ((b[0] * invA00) - (A[1] * invA00) * (b[1] * invA11));
                l1 = l1 + invDet * ((b[1] * invA11) - (b[0] * invA00) * (A[1] * invA11));
            }
        } else if (!two) {
            // RayleighDamp, one constraint (Xpbd.h:216-230)
            float invBeta = comp0 / (dt * damp);
            l0 = -gV[0] / (2.0f * U0 * invBeta + wgg[0]);
        } else {
            // RayleighDamp, two coupled constraints (Xpbd.h:232-263); gV holds b
            float invBeta0 = comp0 / (dt * damp), invBeta1 = comp1 / (dt * damp);
            float A0 = wgg[0] + 2.0f * U0 * invBeta0;
            float A1 = wgg[1];
            float A2 = wgg[2] + 2.0f * U1 * invBeta1;
            float b0 = gV[0], b1 = gV[1];
            float invA00 = 1.0f / A0;
            float invA11 = 1.0f / A2;
            float invDet = 1.0f / fmaxf(0.00000001f, 1.0f - (A1 * invA00) * (A1 * invA11));
            l0 = invDet * ((b0 * invA00) - (A1 * invA00) * (b1 * invA11));
            l1 = invDet * ((b1 * invA11) - (b0 * invA00) * (A1 * invA11));
        }
        T.sc[3] = l0;
        T.sc[4] = l1;
    }
    __syncthreads();

    // Phase 4: node updates.
    if (nt > 0) {
        float l0 = T.sc[3], l1 = T.sc[4];
        for (uint32_t n = lane; n < nn; n += LANES) {
            float wn = T.ws[n];
            if (!POST) {
                double xr[3];
                for (uint32_t r = 0; r < 3; r++) {
                    if (!two) {
                        float s = wn * l0;
                        xr[r] = T.x(r, n) + (double)(s * T.gr(0, r, n));
                    } else {
                        float acc = 0.0f + l0 * T.gr(0, r, n);
                        acc = acc + l1 * T.gr(1, r, n);
                        xr[r] = T.x(r, n) + (double)(wn * acc);
                    }
                }
                // x, y, z only: w (the inverse mass) stays as stored
                double* xn = reinterpret_cast<double*>(X + nodes[n]);
                xn[0] = xr[0];
                xn[1] = xr[1];
                xn[2] = xr[2];
            } else {
                double4 vd = V[nodes[n]];
                double* vr[3] = {&vd.x, &vd.y, &vd.z};
                for (uint32_t r = 0; r < 3; r++) {
                    if (!two) *vr[r] = *vr[r] + (double)((wn * l0) * T.gr(0, r, n));
                    else *vr[r] = *vr[r] + (double)(wn * (l0 * T.gr(0, r, n) + l1 * T.gr(1, r, n)));
                }
                V[nodes[n]] = vd;
            }
        }
    }
}

}  // namespace

// LDS bytes of one ring team (nmax ring nodes, tmax ring tets at most).
extern "C" size_t xf_ring_team_bytes(uint32_t nmax, uint32_t tmax) { return ring_team_bytes(nmax, tmax); }

// One colour of the ring sweep (post = false) or of the ring damping pass
// (post = true), `teams` rings per workgroup and `lanes` (16 or 32) lanes per
// ring. clock: advanced by the first colour of a substep (pass it for that
// launch only); read by the damping pass for the amortized range.
extern "C" hipError_t xf_launch_rings(hipStream_t s, bool post, uint32_t lanes, uint32_t teams, uint32_t slabs,
                                      uint32_t begin, uint32_t count, uint32_t nmax, uint32_t tmax,
                                      const uint4* rmeta, const uint4* rinfo, const uint32_t* rnodes,
                                      const uint16_t* rincb, const uint16_t* rinc, const float4* rtet, double4* X,
                                      const double4* O, double4* V, const XfBodyParams* bodies, float dt,
                                      float* clock) {
    if (count == 0) return hipSuccess;
    if ((lanes != 16 && lanes != 32) || teams == 0 || teams * lanes > 256) return hipErrorInvalidValue;
    size_t lds = (size_t)teams * ring_team_bytes(nmax, tmax);
    dim3 grid(xf::colour_grid(count, teams, slabs)), block(teams * lanes);
    auto k = post ? (lanes == 16 ? xf_ring_colour<true, 16> : xf_ring_colour<true, 32>)
                  : (lanes == 16 ? xf_ring_colour<false, 16> : xf_ring_colour<false, 32>);
    hipLaunchKernelGGL(k, grid, block, lds, s, begin, count, slabs, nmax, tmax, rmeta, rinfo, rnodes, rincb, rinc,
                       rtet, X, O, V, bodies, dt, clock);
    return hipGetLastError();
}
