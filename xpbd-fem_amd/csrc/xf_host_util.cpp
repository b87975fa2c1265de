// xf_host_util.cpp -- host-side mesh, element-constant and colouring code.
#include "xf_host_util.h"

#include <algorithm>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>

namespace xf {

void TetBlock(uint32_t W, uint32_t H, uint32_t D, const float scale[3], const float origin[3], bool mirrored,
              float* xyz, uint32_t* tets) {
    // Node placement and numbering of GenerateHexBlock (MeshGen.cpp:171-185)
    // with an explicit origin instead of the centred one.
    for (uint32_t z = 0; z <= D; z++)
        for (uint32_t y = 0; y <= H; y++)
            for (uint32_t x = 0; x <= W; x++) {
                size_t i = x + (size_t)y * (W + 1) + (size_t)z * (W + 1) * (H + 1);
                xyz[3 * i + 0] = origin[0] + scale[0] * (float)x;
                xyz[3 * i + 1] = origin[1] + scale[1] * (float)y;
                xyz[3 * i + 2] = origin[2] + scale[2] * (float)z;
            }
    // Hex corner numbering (MeshGen.cpp:189-208) and the 6-tet split around
    // the 0-7 diagonal (MeshGen.cpp:235-240).
    static const uint8_t kSplit[6][4] = {{0, 1, 5, 7}, {0, 7, 3, 1}, {0, 2, 3, 7},
                                         {0, 7, 6, 2}, {0, 4, 6, 7}, {0, 7, 5, 4}};
    size_t t = 0;
    for (uint32_t z = 0; z < D; z++)
        for (uint32_t y = 0; y < H; y++)
            for (uint32_t x = 0; x < W; x++) {
                uint32_t corner[8];
                for (uint32_t c = 0; c < 8; c++) {
                    uint32_t i = c & 1, j = (c >> 1) & 1, k = c >> 2;
                    uint32_t slot = c, ox = i, oy = j, oz = k;
                    if (mirrored) {
                        ox = (x & 1) ? 1 - i : i;
                        oy = (y & 1) ? 1 - j : j;
                        oz = (z & 1) ? 1 - k : k;
                        if ((x + y + z) & 1) slot = 7 - c;
                    }
                    corner[slot] = (x + ox) + (y + oy) * (W + 1) + (z + oz) * (W + 1) * (H + 1);
                }
                for (int s = 0; s < 6; s++, t++)
                    for (int c = 0; c < 4; c++) tets[4 * t + c] = corner[kSplit[s][c]];
            }
}

void MinstdOrder(uint32_t n, uint32_t* order) {
    // GeoLinear3d::Init sweep order (Geo.cpp:759-771).
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    uint32_t state = 1;
    for (uint32_t i = 0; i < n; i++) {
        state = (uint32_t)(((uint64_t)state * 48271u) % 0x7fffffffu);
        uint32_t j = state % n;
        uint32_t tmp = order[i];
        order[i] = order[j];
        order[j] = tmp;
    }
}

void NormalFill(uint64_t seed, float sigma, uint32_t n, float* out) {
    // SplitMix64 stream, 53-bit uniforms, Box-Muller pairs (cos branch first).
    uint64_t s = seed;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const double kTwoPi = 6.283185307179586476925286766559;
    for (uint32_t i = 0; i < n; i += 2) {
        double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0, 1]
        double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
        double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = (float)(sigma * r * std::cos(kTwoPi * u2));
        if (i + 1 < n) out[i + 1] = (float)(sigma * r * std::sin(kTwoPi * u2));
    }
}

namespace {
// inverse(mat3) of vectormath.cpp:41-60: adjugate in double, scale in float.
void Inverse3(const float m[3][3], float out[3][3]) {
    double a00 = m[0][0], a01 = m[0][1], a02 = m[0][2];
    double a10 = m[1][0], a11 = m[1][1], a12 = m[1][2];
    double a20 = m[2][0], a21 = m[2][1], a22 = m[2][2];
    float adj[3][3];
    adj[0][0] = (float)+(a11 * a22 - a12 * a21);
    adj[0][1] = (float)-(a01 * a22 - a02 * a21);
    adj[0][2] = (float)+(a01 * a12 - a02 * a11);
    adj[1][0] = (float)-(a10 * a22 - a12 * a20);
    adj[1][1] = (float)+(a00 * a22 - a02 * a20);
    adj[1][2] = (float)-(a00 * a12 - a02 * a10);
    adj[2][0] = (float)+(a10 * a21 - a11 * a20);
    adj[2][1] = (float)-(a00 * a21 - a01 * a20);
    adj[2][2] = (float)+(a00 * a11 - a01 * a10);
    float det = m[0][0] * adj[0][0] + m[0][1] * adj[1][0] + m[0][2] * adj[2][0];
    float s = 1.0f / det;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) out[c][r] = adj[c][r] * s;
}

// determinant(mat3) of vectormath.cpp:34-39.
float Det3(const float m[3][3]) {
    float a = m[0][0], b = m[1][0], c = m[2][0];
    float d = m[0][1], e = m[1][1], f = m[2][1];
    float g = m[0][2], h = m[1][2], i = m[2][2];
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

float Dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
}  // namespace

namespace {
template <typename T>
bool InitElementsT(const T* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                   ElementConstants* out) {
    out->qi.assign((size_t)9 * n_tets, 0.0f);
    out->qq.assign((size_t)3 * n_tets, 0.0f);
    out->qr.assign((size_t)3 * n_tets, 0.0f);
    out->vol.assign(n_tets, 0.0f);
    out->area.assign(n_tets, 0.0f);
    out->w.assign(n_nodes, 0.0f);
    const float kVolMult = 1.0f / 6.0f;  // VolumeMult<Tet> (Fem.cpp:80)
    const float kQuadW = 0.25f;          // QuadratureWeights<Tet, 4> (Fem.cpp:11)
    for (uint32_t t = 0; t < n_tets; t++) {
        const uint32_t* is = tets + 4 * (size_t)t;
        for (int c = 0; c < 4; c++)
            if (is[c] >= n_nodes) return false;
        // InitTriTetFiniteElement (Fem.cpp:197-224): P in float from double positions.
        float P[3][3];
        for (int nd = 0; nd < 3; nd++)
            for (int r = 0; r < 3; r++)
                P[nd][r] = (float)((double)xyz[3 * (size_t)is[nd] + r] - (double)xyz[3 * (size_t)is[3] + r]);
        float J[3][3];  // MatMul(dN, P): column k holds (P0[k], P1[k], P2[k])
        for (int k = 0; k < 3; k++)
            for (int nd = 0; nd < 3; nd++) J[k][nd] = P[nd][k];
        float Qi[3][3];
        Inverse3(J, Qi);
        // GeneratePrefactoredCoefficients (Fem.cpp:132-161), 4 points.
        float QQ[3] = {0.0f, 0.0f, 0.0f}, QR[3] = {0.0f, 0.0f, 0.0f}, volume = 0.0f;
        for (int p = 0; p < 4; p++) {
            float Ji[3][3];
            Inverse3(J, Ji);
            float wq = Det3(J) * kVolMult * kQuadW;
            volume += wq;
            for (int i = 0; i < 3; i++) QQ[i] += wq * Dot3(Ji[i], Ji[i]);
            int head = 0;
            for (int i = 0; i < 3; i++)
                for (int j = i + 1; j < 3; j++) QR[head++] += wq * 2.0f * Dot3(Ji[i], Ji[j]);
        }
        for (int i = 0; i < 3; i++) {
            QQ[i] /= volume;
            QR[i] /= volume;
        }
        float v = kVolMult * Det3(J);
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) out->qi[9 * (size_t)t + 3 * c + r] = Qi[c][r];
        for (int i = 0; i < 3; i++) {
            out->qq[3 * (size_t)t + i] = QQ[i];
            out->qr[3 * (size_t)t + i] = QR[i];
        }
        out->vol[t] = v;
        // surfaceArea (Fem.cpp:211-214): half cross-product norms of the faces
        // (0,1,2), (1,3,2), (0,2,3), (3,1,0), with P3 = 0.
        {
            static const int kFaces[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {3, 1, 0}};
            float Q[4][3];
            for (int nd = 0; nd < 3; nd++)
                for (int r = 0; r < 3; r++) Q[nd][r] = P[nd][r];
            for (int r = 0; r < 3; r++) Q[3][r] = (float)((double)xyz[3 * (size_t)is[3] + r] - (double)xyz[3 * (size_t)is[3] + r]);
            float sum = 0.0f;
            for (int f = 0; f < 4; f++) {
                const float* a = Q[kFaces[f][0]];
                const float* b = Q[kFaces[f][1]];
                const float* c = Q[kFaces[f][2]];
                float e0[3] = {a[0] - c[0], a[1] - c[1], a[2] - c[2]};
                float e1[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]};
                float n[3] = {e0[1] * e1[2] - e1[1] * e0[2], e0[2] * e1[0] - e1[2] * e0[0],
                              e0[0] * e1[1] - e1[0] * e0[1]};
                float fa = 0.5f * std::sqrt(Dot3(n, n));
                sum = f == 0 ? fa : sum + fa;
            }
            out->area[t] = sum;
        }
        // Lumped mass, LumpsTet4 = 1/4 each (Fem.cpp:220-223).
        for (int c = 0; c < 4; c++) out->w[is[c]] += 0.25f * density * v;
    }
    for (uint32_t i = 0; i < n_nodes; i++) out->w[i] = 1.0f / out->w[i];  // Geo.cpp:744
    return true;
}
}  // namespace

bool InitElements(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                  ElementConstants* out) {
    return InitElementsT(xyz, n_nodes, tets, n_tets, density, out);
}

bool InitElements(const double* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                  ElementConstants* out) {
    return InitElementsT(xyz, n_nodes, tets, n_tets, density, out);
}

uint32_t GreedyColour(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, const uint32_t* visit,
                      std::vector<uint32_t>* colour) {
    // Each tet, in visit order, takes the lowest colour none of its 4 nodes
    // has seen; node masks are kWords x 64 bits.
    const int kWords = 4;
    std::vector<uint64_t> mask((size_t)n_nodes * kWords, 0);
    colour->assign(n_tets, 0);
    uint32_t ncol = 0;
    for (uint32_t k = 0; k < n_tets; k++) {
        uint32_t t = visit[k];
        const uint32_t* is = tets + 4 * (size_t)t;
        uint32_t c = kMaxColours;
        for (int word = 0; word < kWords && c == kMaxColours; word++) {
            uint64_t used = mask[(size_t)is[0] * kWords + word] | mask[(size_t)is[1] * kWords + word] |
                            mask[(size_t)is[2] * kWords + word] | mask[(size_t)is[3] * kWords + word];
            if (used != ~0ull) c = 64 * word + (uint32_t)__builtin_ctzll(~used);
        }
        if (c == kMaxColours) return 0;
        (*colour)[t] = c;
        for (int i = 0; i < 4; i++) mask[(size_t)is[i] * kWords + c / 64] |= 1ull << (c % 64);
        if (c + 1 > ncol) ncol = c + 1;
    }
    return ncol;
}

bool BuildRings(const uint32_t* tets, uint32_t n_tets, Rings* out) {
    uint32_t nv = 0;
    for (size_t i = 0; i < 4 * (size_t)n_tets; i++) nv = std::max(nv, tets[i] + 1);
    // tets of each vertex in increasing tet index (Connectivity.cpp:300-312, 436-440)
    std::vector<uint32_t> first(nv + 1, 0), list(4 * (size_t)n_tets);
    for (size_t i = 0; i < 4 * (size_t)n_tets; i++) first[tets[i] + 1]++;
    for (uint32_t v = 0; v < nv; v++) first[v + 1] += first[v];
    for (uint32_t v = 0; v < nv; v++)
        if (first[v + 1] - first[v] > kMaxRingTets) return false;
    std::vector<uint32_t> fill(first.begin(), first.end() - 1);
    for (uint32_t t = 0; t < n_tets; t++)
        for (int j = 0; j < 4; j++) list[fill[tets[4 * (size_t)t + j]]++] = t;
    Rings& r = *out;
    r.n = nv;
    r.largest = 0;
    r.node_off.assign(nv + 1, 0);
    r.tet_off = first;
    r.tets = list;
    r.idx.assign(16 * (size_t)n_tets, 0);  // 4 ring-local indices per ring tet, 4 ring tets per tet
    r.nodes.clear();
    std::vector<uint32_t> verts;
    for (uint32_t v = 0; v < nv; v++) {
        verts.assign(1, v);
        for (uint32_t k = first[v]; k < first[v + 1]; k++) {
            const uint32_t* is = tets + 4 * (size_t)list[k];
            for (int j = 0; j < 4; j++) {
                uint32_t li = (uint32_t)(std::find(verts.begin(), verts.end(), is[j]) - verts.begin());
                if (li == verts.size()) {
                    if (verts.size() == kMaxRingNodes) return false;
                    verts.push_back(is[j]);
                }
                r.idx[4 * (size_t)k + j] = (uint8_t)li;
            }
        }
        r.node_off[v] = (uint32_t)r.nodes.size();
        r.nodes.insert(r.nodes.end(), verts.begin(), verts.end());
        r.largest = std::max(r.largest, (uint32_t)verts.size());
    }
    r.node_off[nv] = (uint32_t)r.nodes.size();
    return true;
}

void MinstdRingOrder(uint32_t n_tets, uint32_t n_rings, uint32_t* order) {
    uint32_t state = 1;
    for (uint32_t i = 0; i < n_tets; i++) state = (uint32_t)(((uint64_t)state * 48271u) % 0x7fffffffu);
    for (uint32_t i = 0; i < n_rings; i++) order[i] = i;
    for (uint32_t i = 0; i < n_rings; i++) {
        state = (uint32_t)(((uint64_t)state * 48271u) % 0x7fffffffu);
        uint32_t j = state % n_rings;
        std::swap(order[i], order[j]);
    }
}

uint32_t RingColourSweepOrder(const Rings& r, uint32_t n_nodes, const uint32_t* visit,
                              std::vector<uint32_t>* colour, std::vector<uint32_t>* order) {
    const uint32_t n = r.n;
    std::vector<uint32_t> rank(n, 0);
    for (uint32_t k = 0; k < n; k++) rank[visit[k]] = k;
    // node -> rings holding it
    std::vector<uint32_t> first(n_nodes + 1, 0);
    for (uint32_t v = 0; v < n; v++)
        for (uint32_t q = r.node_off[v]; q < r.node_off[v + 1]; q++) first[r.nodes[q] + 1]++;
    for (uint32_t u = 0; u < n_nodes; u++) first[u + 1] += first[u];
    std::vector<uint32_t> inc(first[n_nodes]);
    {
        std::vector<uint32_t> fill(first.begin(), first.end() - 1);
        for (uint32_t v = 0; v < n; v++)
            for (uint32_t q = r.node_off[v]; q < r.node_off[v + 1]; q++) inc[fill[r.nodes[q]]++] = v;
    }
    // conflicting rings (sharing a node), distinct
    std::vector<uint64_t> noff(n + 1, 0);
    std::vector<uint32_t> nbr, stamp(n, ~0u);
    for (uint32_t v = 0; v < n; v++) {
        for (uint32_t q = r.node_off[v]; q < r.node_off[v + 1]; q++) {
            uint32_t u = r.nodes[q];
            for (uint32_t k = first[u]; k < first[u + 1]; k++) {
                uint32_t w = inc[k];
                if (w != v && stamp[w] != v) {
                    stamp[w] = v;
                    nbr.push_back(w);
                }
            }
        }
        noff[v + 1] = nbr.size();
    }
    // DSatur: the uncoloured ring with the most distinct colours among its
    // conflicting rings, then the most conflicting rings, then the earliest
    // in visit order, takes the lowest colour none of them has.
    const int kWords = 4;
    std::vector<uint64_t> sat((size_t)n * kWords, 0);
    std::vector<uint32_t> nsat(n, 0);
    struct Item {
        uint32_t sat, deg, rank, ring;
        bool operator<(const Item& o) const {
            if (sat != o.sat) return sat < o.sat;
            if (deg != o.deg) return deg < o.deg;
            return rank > o.rank;
        }
    };
    std::vector<Item> heap;
    heap.reserve(n);
    for (uint32_t v = 0; v < n; v++) heap.push_back({0u, (uint32_t)(noff[v + 1] - noff[v]), rank[v], v});
    std::make_heap(heap.begin(), heap.end());
    colour->assign(n, kMaxColours);
    uint32_t ncol = 0;
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end());
        Item it = heap.back();
        heap.pop_back();
        uint32_t v = it.ring;
        if ((*colour)[v] != kMaxColours || it.sat != nsat[v]) continue;
        uint32_t c = kMaxColours;
        for (int word = 0; word < kWords && c == kMaxColours; word++) {
            uint64_t used = sat[(size_t)v * kWords + word];
            if (used != ~0ull) c = 64 * word + (uint32_t)__builtin_ctzll(~used);
        }
        if (c == kMaxColours) return 0;
        (*colour)[v] = c;
        ncol = std::max(ncol, c + 1);
        for (uint64_t k = noff[v]; k < noff[v + 1]; k++) {
            uint32_t w = nbr[k];
            uint64_t& m = sat[(size_t)w * kWords + c / 64];
            if ((*colour)[w] != kMaxColours || (m >> (c % 64) & 1ull)) continue;
            m |= 1ull << (c % 64);
            nsat[w]++;
            heap.push_back({nsat[w], (uint32_t)(noff[w + 1] - noff[w]), rank[w], w});
            std::push_heap(heap.begin(), heap.end());
        }
    }
    std::vector<uint32_t> start(ncol + 1, 0);
    for (uint32_t v = 0; v < r.n; v++) start[(*colour)[v] + 1]++;
    for (uint32_t c = 0; c < ncol; c++) start[c + 1] += start[c];
    order->assign(r.n, 0);
    for (uint32_t k = 0; k < r.n; k++) {
        uint32_t v = visit[k];
        (*order)[start[(*colour)[v]]++] = v;
    }
    return ncol;
}

uint32_t ColourSweepOrder(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, bool minstd_visit,
                          std::vector<uint32_t>* colour, std::vector<uint32_t>* order) {
    std::vector<uint32_t> visit(n_tets);
    if (minstd_visit) MinstdOrder(n_tets, visit.data());
    else
        for (uint32_t t = 0; t < n_tets; t++) visit[t] = t;
    uint32_t ncol = GreedyColour(n_nodes, tets, n_tets, visit.data(), colour);
    if (ncol == 0) return 0;
    // counting sort by colour, stable in visit order
    std::vector<uint32_t> start(ncol + 1, 0);
    for (uint32_t t = 0; t < n_tets; t++) start[(*colour)[t] + 1]++;
    for (uint32_t c = 0; c < ncol; c++) start[c + 1] += start[c];
    order->assign(n_tets, 0);
    for (uint32_t k = 0; k < n_tets; k++) {
        uint32_t t = visit[k];
        (*order)[start[(*colour)[t]]++] = t;
    }
    return ncol;
}

namespace {

// Breadth-first levels over tet edges (the nodes of one tet are adjacent)
// from the used nodes whose coordinate v = d . p lies within 1e-4 of its
// range above the least; components the search did not reach restart at
// level 0 from their node of least v (ties by index). Unused nodes keep
// level -1.
void DirectionLevels(const float* xyz, uint32_t n_nodes, const uint32_t* tets, const std::vector<uint32_t>& first,
                     const std::vector<uint32_t>& inc, const int d[3], std::vector<int32_t>* level) {
    std::vector<int32_t>& lev = *level;
    lev.assign(n_nodes, -1);
    std::vector<float> v(n_nodes, 0.0f);
    float lo = 0.0f, hi = 0.0f;
    bool any = false;
    for (uint32_t n = 0; n < n_nodes; n++) {
        if (first[n + 1] == first[n]) continue;
        const float* p = xyz + 3 * (size_t)n;
        v[n] = ((float)d[0] * p[0] + (float)d[1] * p[1]) + (float)d[2] * p[2];
        if (!any || v[n] < lo) lo = v[n];
        if (!any || v[n] > hi) hi = v[n];
        any = true;
    }
    if (!any) return;
    float tol = lo + 1e-4f * (hi - lo);
    std::vector<uint32_t> queue;
    queue.reserve(n_nodes);
    auto run = [&](size_t head) {
        for (; head < queue.size(); head++) {
            uint32_t u = queue[head];
            for (uint32_t k = first[u]; k < first[u + 1]; k++) {
                const uint32_t* is = tets + 4 * (size_t)inc[k];
                for (int i = 0; i < 4; i++)
                    if (lev[is[i]] < 0) {
                        lev[is[i]] = lev[u] + 1;
                        queue.push_back(is[i]);
                    }
            }
        }
    };
    for (uint32_t n = 0; n < n_nodes; n++)
        if (first[n + 1] > first[n] && v[n] <= tol) {
            lev[n] = 0;
            queue.push_back(n);
        }
    run(0);
    std::vector<uint32_t> rest;
    for (uint32_t n = 0; n < n_nodes; n++)
        if (first[n + 1] > first[n] && lev[n] < 0) rest.push_back(n);
    std::stable_sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) { return v[a] < v[b]; });
    for (uint32_t n : rest)
        if (lev[n] < 0) {
            size_t head = queue.size();
            lev[n] = 0;
            queue.push_back(n);
            run(head);
        }
}

// Level directions tried for the block partition: x with y, y - z, y + z,
// z; y with z, x - z, x + z; z with x - y, x + y.
constexpr int kDirs[9][3] = {{1, 0, 0}, {0, 1, 0},  {0, 0, 1},  {0, 1, -1}, {0, 1, 1},
                             {1, 0, -1}, {1, 0, 1}, {1, -1, 0}, {1, 1, 0}};
constexpr int kPairs[9][2] = {{0, 1}, {0, 3}, {0, 4}, {0, 2}, {1, 2}, {1, 5}, {1, 6}, {2, 7}, {2, 8}};

uint64_t BlockKey(const std::vector<int32_t>& l1, const std::vector<int32_t>& l2, const uint32_t* is) {
    int32_t k1 = std::min(std::min(l1[is[0]], l1[is[1]]), std::min(l1[is[2]], l1[is[3]]));
    int32_t k2 = std::min(std::min(l2[is[0]], l2[is[1]]), std::min(l2[is[2]], l2[is[3]]));
    return ((uint64_t)(uint32_t)k1 << 32) | (uint32_t)k2;
}

}  // namespace

bool BuildSweepBlocks(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets,
                      const uint32_t* visit, SweepBlocks* out, std::string* err) {
    // node -> tets incidence
    std::vector<uint32_t> first(n_nodes + 1, 0), inc(4 * (size_t)n_tets);
    for (size_t i = 0; i < 4 * (size_t)n_tets; i++) {
        if (tets[i] >= n_nodes) {
            if (err) *err = "tet index out of range";
            return false;
        }
        first[tets[i] + 1]++;
    }
    for (uint32_t v = 0; v < n_nodes; v++) first[v + 1] += first[v];
    {
        std::vector<uint32_t> fill(first.begin(), first.end() - 1);
        for (uint32_t t = 0; t < n_tets; t++)
            for (int j = 0; j < 4; j++) inc[fill[tets[4 * (size_t)t + j]]++] = t;
    }
    // Levels of each direction (computed when a pair needs them); the pair
    // whose blocks put the fewest tets on one node (a lower bound on a
    // block's phases) wins, the first on ties. A pair reaching a quarter of
    // the largest node valence (the least any 4-colour block scheme can
    // reach) ends the search.
    std::vector<int32_t> lv[9];
    bool have[9] = {};
    uint32_t valence = 0;
    for (uint32_t v = 0; v < n_nodes; v++) valence = std::max(valence, first[v + 1] - first[v]);
    const uint32_t floor_bound = (valence + 3) / 4;
    std::vector<uint64_t> key(n_tets);
    int best = 0;
    uint32_t best_bound = ~0u;
    for (int c = 0; c < 9 && best_bound > floor_bound; c++) {
        for (int a = 0; a < 2; a++)
            if (!have[kPairs[c][a]]) {
                DirectionLevels(xyz, n_nodes, tets, first, inc, kDirs[kPairs[c][a]], &lv[kPairs[c][a]]);
                have[kPairs[c][a]] = true;
            }
        const std::vector<int32_t>&l1 = lv[kPairs[c][0]], &l2 = lv[kPairs[c][1]];
        for (uint32_t t = 0; t < n_tets; t++) key[t] = BlockKey(l1, l2, tets + 4 * (size_t)t);
        uint32_t bound = 0;
        std::vector<uint64_t> ks;
        for (uint32_t n = 0; n < n_nodes && bound < best_bound; n++) {
            ks.clear();
            for (uint32_t k = first[n]; k < first[n + 1]; k++) ks.push_back(key[inc[k]]);
            std::sort(ks.begin(), ks.end());
            for (size_t i = 0, j; i < ks.size(); i = j) {
                for (j = i; j < ks.size() && ks[j] == ks[i]; j++) {
                }
                bound = std::max(bound, (uint32_t)(j - i));
            }
        }
        if (bound < best_bound) {
            best_bound = bound;
            best = c;
        }
    }
    const std::vector<int32_t>&l1 = lv[kPairs[best][0]], &l2 = lv[kPairs[best][1]];
    for (uint32_t t = 0; t < n_tets; t++) key[t] = BlockKey(l1, l2, tets + 4 * (size_t)t);

    auto launch_of = [](uint64_t k) -> uint32_t { return (uint32_t)((k >> 32) & 1u) | (uint32_t)((k & 1u) << 1); };
    auto before = [&](uint64_t a, uint64_t b) {
        uint32_t qa = launch_of(a), qb = launch_of(b);
        return qa != qb ? qa < qb : a < b;
    };
    SweepBlocks& r = *out;
    r.axes[0] = kPairs[best][0];
    r.axes[1] = kPairs[best][1];
    r.block.assign(n_tets, 0);
    std::vector<uint64_t> keys;
    uint32_t m1 = 0, m2 = 0;
    for (uint32_t t = 0; t < n_tets; t++) {
        m1 = std::max(m1, (uint32_t)(key[t] >> 32) + 1);
        m2 = std::max(m2, (uint32_t)key[t] + 1);
    }
    if ((uint64_t)m1 * m2 <= 4 * (uint64_t)n_tets + 1024) {
        // dense (k1, k2) grid: blocks numbered by launch colour, then k1, k2
        std::vector<uint32_t> id((size_t)m1 * m2, ~0u);
        for (uint32_t t = 0; t < n_tets; t++) id[(size_t)(key[t] >> 32) * m2 + (uint32_t)key[t]] = 0;
        for (uint32_t q = 0; q < 4; q++)
            for (uint32_t k1 = q & 1u; k1 < m1; k1 += 2)
                for (uint32_t k2 = q >> 1; k2 < m2; k2 += 2)
                    if (id[(size_t)k1 * m2 + k2] == 0) {
                        id[(size_t)k1 * m2 + k2] = (uint32_t)keys.size() + 1;
                        keys.push_back(((uint64_t)k1 << 32) | k2);
                    }
        for (uint32_t t = 0; t < n_tets; t++) r.block[t] = id[(size_t)(key[t] >> 32) * m2 + (uint32_t)key[t]] - 1;
    } else {
        keys = key;
        std::sort(keys.begin(), keys.end(), before);
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        for (uint32_t t = 0; t < n_tets; t++)
            r.block[t] = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), key[t], before) - keys.begin());
    }
    r.n_blocks = (uint32_t)keys.size();
    r.launch.assign(r.n_blocks, 0);
    for (uint32_t b = 0; b < r.n_blocks; b++) r.launch[b] = (uint8_t)launch_of(keys[b]);
    // tets of each block in visit order
    std::vector<uint32_t> bstart(r.n_blocks + 1, 0), bl(n_tets);
    for (uint32_t t = 0; t < n_tets; t++) bstart[r.block[t] + 1]++;
    for (uint32_t b = 0; b < r.n_blocks; b++) bstart[b + 1] += bstart[b];
    {
        std::vector<uint32_t> fill(bstart.begin(), bstart.end() - 1);
        for (uint32_t k = 0; k < n_tets; k++) bl[fill[r.block[visit[k]]]++] = visit[k];
    }
    // Phases inside each block: DSatur (the uncoloured tet with the most
    // distinct phases among its neighbours -- tets of the block sharing a
    // node --, then the most neighbours, then the earliest in visit order
    // takes the lowest phase none of its neighbours has).
    std::vector<uint64_t> mask(n_nodes, 0);
    r.phase.assign(n_tets, 0);
    r.phases.assign(r.n_blocks, 0);
    constexpr uint32_t kNoLocal = ~0u;
    std::vector<uint32_t> slot_of(n_nodes, kNoLocal), nodes_seen, goff, gfill, gtet, noff, nbr, stamp, at;
    std::vector<uint64_t> sat, keyv;
    for (uint32_t b = 0; b < r.n_blocks; b++) {
        const uint32_t* L = bl.data() + bstart[b];
        uint32_t n = bstart[b + 1] - bstart[b];
        // neighbour lists (distinct): the block's tets grouped by node
        // (counting sort over block-local node slots), deduplicated by stamp
        nodes_seen.clear();
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) {
                uint32_t g = tets[4 * (size_t)L[i] + c];
                if (slot_of[g] == kNoLocal) {
                    slot_of[g] = (uint32_t)nodes_seen.size();
                    nodes_seen.push_back(g);
                }
            }
        uint32_t nn = (uint32_t)nodes_seen.size();
        goff.assign(nn + 1, 0);
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) goff[slot_of[tets[4 * (size_t)L[i] + c]] + 1]++;
        for (uint32_t k = 0; k < nn; k++) goff[k + 1] += goff[k];
        gtet.resize(4 * (size_t)n);
        gfill.assign(goff.begin(), goff.end() - 1);
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) gtet[gfill[slot_of[tets[4 * (size_t)L[i] + c]]]++] = i;
        noff.assign(n + 1, 0);
        nbr.clear();
        stamp.assign(n, ~0u);
        for (uint32_t i = 0; i < n; i++) {
            for (int c = 0; c < 4; c++) {
                uint32_t sl = slot_of[tets[4 * (size_t)L[i] + c]];
                for (uint32_t k = goff[sl]; k < goff[sl + 1]; k++) {
                    uint32_t j = gtet[k];
                    if (j != i && stamp[j] != i) {
                        stamp[j] = i;
                        nbr.push_back(j);
                    }
                }
            }
            noff[i + 1] = (uint32_t)nbr.size();
        }
        for (uint32_t g : nodes_seen) slot_of[g] = kNoLocal;
        sat.assign(n, 0);
        // selection key: saturation << 48 | degree << 32 | ~position, kept
        // for the uncoloured tets only (at[] locates each tet's key)
        keyv.resize(n);
        at.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            keyv[i] = ((uint64_t)(noff[i + 1] - noff[i]) << 32) | (uint64_t)(0xFFFFFFFFu - i);
            at[i] = i;
        }
        uint32_t np = 0;
        for (uint32_t live = n; live > 0; live--) {
            uint32_t top = 0;
            for (uint32_t k = 1; k < live; k++)
                if (keyv[k] > keyv[top]) top = k;
            uint32_t i = 0xFFFFFFFFu - (uint32_t)keyv[top];
            if (sat[i] == ~0ull) {
                if (err) *err = "a sweep block needs more than 64 phases";
                return false;
            }
            uint32_t c = (uint32_t)__builtin_ctzll(~sat[i]);
            r.phase[L[i]] = c;
            np = std::max(np, c + 1);
            // remove i's key (the last live key takes its place)
            keyv[top] = keyv[live - 1];
            at[0xFFFFFFFFu - (uint32_t)keyv[top]] = top;
            at[i] = ~0u;
            for (uint32_t k = noff[i]; k < noff[i + 1]; k++) {
                uint32_t j = nbr[k];
                if (at[j] == ~0u || (sat[j] >> c & 1ull)) continue;
                sat[j] |= 1ull << c;
                keyv[at[j]] += 1ull << 48;
            }
        }
        // Balance: tets of phases above ceil(n / np), in visit order, move to
        // the first phase below that size that none of their nodes touches
        // (a phase is node-disjoint, so the tet is the only one of its phase
        // on each of its nodes).
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t* is = tets + 4 * (size_t)L[i];
            for (int c = 0; c < 4; c++) mask[is[c]] |= 1ull << r.phase[L[i]];
        }
        uint32_t cap = (n + np - 1) / np;
        uint32_t size[kMaxBlockPhases] = {};
        for (uint32_t i = 0; i < n; i++) size[r.phase[L[i]]]++;
        for (uint32_t i = 0; i < n; i++) {
            uint32_t t = L[i], c = r.phase[t];
            if (size[c] <= cap) continue;
            const uint32_t* is = tets + 4 * (size_t)t;
            uint64_t used = mask[is[0]] | mask[is[1]] | mask[is[2]] | mask[is[3]];
            for (uint32_t d = 0; d < np; d++) {
                if (size[d] >= cap || (used >> d & 1ull)) continue;
                for (int k = 0; k < 4; k++) mask[is[k]] = (mask[is[k]] & ~(1ull << c)) | (1ull << d);
                size[c]--;
                size[d]++;
                r.phase[t] = d;
                break;
            }
        }
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t* is = tets + 4 * (size_t)L[i];
            for (int c = 0; c < 4; c++) mask[is[c]] = 0;
        }
        r.phases[b] = np;
    }
    return true;
}

int LdltSolve(uint32_t n, const float* A, const float* b, float* x) {
    // LDLTDecomposition / LDLTSolve (Ldlt.h:35-86) on packed lower-triangular A.
    if (n == 0 || n > 25) return -1;
    uint32_t rowA[26], rowL[26];
    for (uint32_t i = 0; i <= n; i++) {
        rowA[i] = i * (i + 1) / 2;  // packed row start including the diagonal
        rowL[i] = i * (i - 1) / 2;  // strictly lower part
    }
    float L[325], D[25], invD[25];
    for (uint32_t j = 0; j < n; j++) {
        D[j] = A[rowA[j] + j];
        for (uint32_t k = 0; k < j; k++) D[j] -= L[rowL[j] + k] * L[rowL[j] + k] * D[k];
        for (uint32_t i = j + 1; i < n; i++) {
            L[rowL[i] + j] = A[rowA[i] + j];
            for (uint32_t k = 0; k < j; k++) L[rowL[i] + j] -= L[rowL[i] + k] * L[rowL[j] + k] * D[k];
            L[rowL[i] + j] /= D[j];
        }
    }
    for (uint32_t j = 0; j < n; j++) invD[j] = 1.0f / D[j];
    for (uint32_t i = 0; i < n; i++) {
        x[i] = b[i];
        for (uint32_t j = 0; j < i; j++) x[i] -= L[rowL[i] + j] * x[j];
    }
    for (uint32_t i = 0; i < n; i++) x[i] *= invD[i];
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < i; j++) x[n - 1 - i] -= L[rowL[n - 1 - j] + n - 1 - i] * x[n - 1 - j];
    return 0;
}

namespace {

// Numeric rows of a tetgen file, comments ('#' to end of line) dropped.
bool ReadRows(const char* path, std::vector<std::vector<double>>* rows, std::string* err) {
    FILE* f = fopen(path, "r");
    if (!f) {
        *err = std::string("cannot open ") + path;
        return false;
    }
    std::string line;
    int ch;
    bool ok = true;
    auto flush = [&]() {
        size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::vector<double> row;
        const char* p = line.c_str();
        while (*p && ok) {
            if (*p == ' ' || *p == '\t' || *p == '\r' || *p == ',') {
                ++p;
                continue;
            }
            char* end = nullptr;
            double v = strtod(p, &end);
            if (end == p) {
                *err = std::string("unparsable text in ") + path + ": " + line;
                ok = false;
                break;
            }
            row.push_back(v);
            p = end;
        }
        if (ok && !row.empty()) rows->push_back(row);
        line.clear();
    };
    while ((ch = fgetc(f)) != EOF && ok) {
        if (ch == '\n') flush();
        else line.push_back((char)ch);
    }
    if (ok) flush();
    fclose(f);
    return ok;
}

// A header count: finite, integral and in [0, 2^31) (casts of anything else
// to an integer are undefined, and counts near 2^64 would wrap in n + 1).
bool CountOk(double v) {
    return std::isfinite(v) && v >= 0.0 && v < 2147483648.0 && v == std::floor(v);
}

}  // namespace

bool ReadTetgen(const char* node_path, const char* ele_path, bool orient, std::vector<float>* xyz,
                std::vector<uint32_t>* tets, std::string* err) {
    std::vector<std::vector<double>> nr, er;
    if (!ReadRows(node_path, &nr, err) || !ReadRows(ele_path, &er, err)) return false;
    if (nr.empty() || nr[0].size() < 2 || nr[0][1] != 3.0) {
        *err = "node file: header must read <count> 3 [attributes] [markers]";
        return false;
    }
    if (!CountOk(nr[0][0])) {
        *err = "node file: the node count must be a whole number in [0, 2^31)";
        return false;
    }
    size_t n = (size_t)nr[0][0];
    if (nr.size() < n + 1) {
        *err = "node file: fewer node lines than the header count";
        return false;
    }
    double base = n ? nr[1][0] : 0.0;
    if (base != 0.0 && base != 1.0) {
        *err = "node file: numbering must start at 0 or 1";
        return false;
    }
    xyz->assign(3 * n, 0.0f);
    for (size_t i = 0; i < n; i++) {
        const std::vector<double>& r = nr[i + 1];
        if (r.size() < 4 || r[0] != base + (double)i) {
            *err = "node file: node lines must be <index> <x> <y> <z> in order";
            return false;
        }
        for (int c = 0; c < 3; c++) (*xyz)[3 * i + c] = (float)r[1 + c];
    }
    if (er.empty() || er[0].size() < 2) {
        *err = "ele file: header must read <count> <nodes per tet> [attributes]";
        return false;
    }
    if (er[0][1] != 4.0) {
        *err = "ele file: only linear tets (4 nodes per tet) are supported";
        return false;
    }
    if (!CountOk(er[0][0])) {
        *err = "ele file: the tet count must be a whole number in [0, 2^31)";
        return false;
    }
    size_t t = (size_t)er[0][0];
    if (er.size() < t + 1) {
        *err = "ele file: fewer tet lines than the header count";
        return false;
    }
    tets->assign(4 * t, 0u);
    for (size_t k = 0; k < t; k++) {
        const std::vector<double>& r = er[k + 1];
        if (r.size() < 5) {
            *err = "ele file: tet lines must be <index> <n0> <n1> <n2> <n3>";
            return false;
        }
        for (int c = 0; c < 4; c++) {
            double v = r[1 + c] - base;
            if (!CountOk(v) || v >= (double)n) {
                *err = "ele file: node index out of range";
                return false;
            }
            (*tets)[4 * k + c] = (uint32_t)v;
        }
        if (orient) {
            uint32_t* is = &(*tets)[4 * k];
            double P[3][3];
            for (int a = 0; a < 3; a++)
                for (int r2 = 0; r2 < 3; r2++)
                    P[a][r2] = (double)(*xyz)[3 * is[a] + r2] - (double)(*xyz)[3 * is[3] + r2];
            double det = P[0][0] * (P[1][1] * P[2][2] - P[1][2] * P[2][1]) -
                         P[0][1] * (P[1][0] * P[2][2] - P[1][2] * P[2][0]) +
                         P[0][2] * (P[1][0] * P[2][1] - P[1][1] * P[2][0]);
            if (det < 0.0) std::swap(is[0], is[1]);
        }
    }
    return true;
}

bool TetsFromStream(const uint32_t* idx, uint32_t count, std::vector<uint32_t>* tets) {
    tets->clear();
    for (uint32_t i = 0; i < count;) {
        if (idx[i] != 4 || i + 4 >= count) return false;
        tets->insert(tets->end(), idx + i + 1, idx + i + 5);
        i += 5;
    }
    return true;
}

}  // namespace xf
