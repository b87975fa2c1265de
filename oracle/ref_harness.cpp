// ref_harness.cpp -- extern "C" glue over the unmodified reference sources
// (XPBDFEM/*.cpp compiled from /root/reference by oracle/Makefile into
// oracle/_ref/libxfref.so).
//
// TEST INFRASTRUCTURE ONLY: it lets the tests and the cpu_baseline leg of
// bench.py drive the reference's own GeoLinear3d on the same mesh and the
// same sweep order as the GPU path. It calls only the reference's public
// functions and writes only the public members the reference exposes
// (tOrder, vOrder, flags, w, X, V; Geo.h:61-66, 161-162).
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "Geo.h"
#include "Ldlt.h"
#include "MeshGen.h"

// Defined in MeshGen.cpp:223 (not declared in MeshGen.h).
void GenerateTetBlock(Allocator* alloc, uint32_t width, uint32_t height, uint32_t depth, vec3 scale,
                      uint32_t elementPattern, float wonkiness, const float** outNodeData,
                      uint32_t* outNodeDataCount, const uint32_t** outIdxData, uint32_t* outIdxDataCount);

namespace {
struct RefBody {
    std::unique_ptr<uint8_t[]> mem;  // the bump allocator's buffer (pages touched only as used)
    Allocator alloc;
    GeoLinear3d* geo = nullptr;
};

struct xr_settings {
    float compliance;
    float poissons_ratio;
    float gravity[2];
    float drag;
    uint32_t energy;
    uint32_t simultaneous;
    uint32_t volume_passes;       // Settings::volumePasses
    float damping;                // Settings::damping (Rayleigh gamma)
    uint32_t rayleigh_type;       // Settings flag bits 20-21 (Rayleigh_Paper .. Rayleigh_PostAmortized)
    float pbd_damping;            // Settings::pbdDamping (only its sign gates the pass, Geo.cpp:807)
    float volume_pbd;             // Settings::volumeAndTimeCorrectedPbdDamping
    float amortized_volume_pbd;   // Settings::amortizedVolumeAndTimeCorrectedPbdDamping
    uint32_t tick_id;             // Settings::tickId of the first substep; +1 per substep
};

Settings MakeSettings(const xr_settings* s) {
    Settings st{};
    st.compliance = s->compliance;
    st.poissonsRatio = s->poissons_ratio;
    st.gravity = vec2(s->gravity[0], s->gravity[1]);
    st.damping = s->damping;
    st.pbdDamping = s->pbd_damping;
    st.volumeAndTimeCorrectedPbdDamping = s->volume_pbd;
    st.amortizedVolumeAndTimeCorrectedPbdDamping = s->amortized_volume_pbd;
    st.drag = 0.0f;
    st.timeCorrectedDrag = s->drag;
    st.volumePasses = s->volume_passes;
    st.tickId = s->tick_id;
    st.flags = (s->energy << Settings_EnergyBit) | (s->simultaneous ? (1u << Settings_XpbdSolveBit) : 0u) |
               ((s->rayleigh_type & Settings_RayleighTypeMask) << Settings_RayleighTypeBit) | Settings_LockLeft;
    return st;
}
}  // namespace

extern "C" {

// Builds a GeoLinear3d from xyz[n_nodes*3] and tets[n_tets*4] through the
// reference's own ingestion path (index stream [4, i0, i1, i2, i3]*), with
// Init's autoResize when auto_resize != 0.
void* xr_create_ex(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                   int auto_resize) {
    RefBody* b = new RefBody();
    size_t bytes = (size_t)4096 * n_tets + (size_t)4096 * n_nodes + ((size_t)4 << 20);
    b->mem.reset(new uint8_t[bytes]);
    b->alloc.Initialize(b->mem.get(), bytes);
    std::vector<uint32_t> idx((size_t)5 * n_tets);
    for (uint32_t t = 0; t < n_tets; t++) {
        idx[5 * t] = 4;
        for (int c = 0; c < 4; c++) idx[5 * t + 1 + c] = tets[4 * t + c];
    }
    b->geo = b->alloc.New<GeoLinear3d>();
    b->geo->Init(&b->alloc, density, xyz, 3 * n_nodes, idx.data(), (uint32_t)idx.size(), auto_resize != 0);
    return b;
}

void* xr_create(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density) {
    return xr_create_ex(xyz, n_nodes, tets, n_tets, density, 0);
}

void xr_destroy(void* h) { delete (RefBody*)h; }

uint32_t xr_vert_count(void* h) { return ((RefBody*)h)->geo->VertCount(); }
uint32_t xr_element_count(void* h) { return ((RefBody*)h)->geo->ElementCount(); }

void xr_get_order(void* h, uint32_t* order) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    memcpy(order, g->tOrder, sizeof(uint32_t) * g->con.tetCount);
}

void xr_set_order(void* h, const uint32_t* order) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    memcpy(g->tOrder, order, sizeof(uint32_t) * g->con.tetCount);
}

// Pins exactly the given nodes (Left flag + w = 0 from the first substep).
void xr_set_pins(void* h, const uint8_t* pinned) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t i = 0; i < g->vertCount; i++) {
        g->flags[i] &= (uint8_t)~(Geo::Left | Geo::Right);
        if (pinned && pinned[i]) {
            g->flags[i] |= Geo::Left;
            g->w[i] = 0.0f;
        }
    }
}

void xr_get_elements(void* h, float* qi9, float* qq3, float* qr3, float* vol, float* w) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t t = 0; t < g->con.tetCount; t++) {
        const ElementParamsLinearTet& ep = g->t[t].ep;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) qi9[9 * t + 3 * c + r] = ep.Qi[c][r];
        for (int i = 0; i < 3; i++) { qq3[3 * t + i] = ep.ic.QQ[i]; qr3[3 * t + i] = ep.ic.QR[i]; }
        vol[t] = ep.volume;
    }
    for (uint32_t i = 0; i < g->vertCount; i++) w[i] = g->w[i];
}

// The reference's compiled-in Armadillo tet mesh (GenerateArmadillo,
// MeshGen.cpp:54-81): xyz node data and the [4, i0..i3]* index stream.
void xr_armadillo_t4(const float** nodes, uint32_t* node_count, const uint32_t** idx, uint32_t* idx_count) {
    GenerateArmadillo(Element_T4, nodes, node_count, idx, idx_count);
}

// ElementParamsLinearTet::surfaceArea per tet (used by the PBD damping pass).
void xr_get_area(void* h, float* area) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t t = 0; t < g->con.tetCount; t++) area[t] = g->t[t].ep.surfaceArea;
}

void xr_get_tets(void* h, uint32_t* tets) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t t = 0; t < g->con.tetCount; t++)
        for (int c = 0; c < 4; c++) tets[4 * t + c] = g->t[t].i[c];
}

void xr_get_state(void* h, double* X, double* V) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t i = 0; i < g->vertCount; i++)
        for (int r = 0; r < 3; r++) {
            if (X) X[3 * i + r] = g->X[i][r];
            if (V) V[3 * i + r] = g->V[i][r];
        }
}

void xr_set_state(void* h, const double* X, const double* V) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    for (uint32_t i = 0; i < g->vertCount; i++)
        for (int r = 0; r < 3; r++) {
            if (X) g->X[i][r] = g->O[i][r] = X[3 * i + r];
            if (V) g->V[i][r] = V[3 * i + r];
        }
}

// Runs `count` calls of Geo3d::Substep (Geo.cpp:305-356).
void xr_substeps(void* h, const xr_settings* s, float dt, uint32_t count) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    Settings st = MakeSettings(s);
    Manipulator manip;
    manip.pickedGeo = nullptr;
    for (uint32_t k = 0; k < count; k++) {
        g->Substep(st, manip, dt);
        ++st.tickId;  // Sim::Update advances tickId once per substep (Demo.cpp:81-89)
    }
}

// `count` substeps of each of n bodies, bodies dealt to `threads` host
// threads in turn (each body is the reference's single-threaded Substep;
// bodies share nothing). settings: one per body.
void xr_substeps_many(void** hs, const xr_settings* ss, uint32_t n, float dt, uint32_t count, uint32_t threads) {
    std::atomic<uint32_t> next(0);
    auto work = [&]() {
        for (uint32_t i = next++; i < n; i = next++) xr_substeps(hs[i], ss + i, dt, count);
    };
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < threads; k++) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
}

// One SolveElement call (Fem.cpp:892) on a 4-node local state.
void xr_solve_tet(const xr_settings* s, float dt, double* X4, const float* w4, const float* qi9, const float* qq3,
                  const float* qr3, float volume) {
    Settings st = MakeSettings(s);
    ElementParamsLinearTet ep;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) ep.Qi[c][r] = qi9[3 * c + r];
    for (int i = 0; i < 3; i++) { ep.ic.QQ[i] = qq3[i]; ep.ic.QR[i] = qr3[i]; }
    ep.volume = volume;
    ep.surfaceArea = 0.0f;
    dvec3 X[4], O[4];
    for (int n = 0; n < 4; n++) {
        X[n] = dvec3(X4[3 * n], X4[3 * n + 1], X4[3 * n + 2]);
        O[n] = X[n];
    }
    const uint32_t is[4] = {0, 1, 2, 3};
    SolveElement(dt, X, O, w4, is, ep, st);
    for (int n = 0; n < 4; n++)
        for (int r = 0; r < 3; r++) X4[3 * n + r] = X[n][r];
}

// GenerateTetBlock (MeshGen.cpp:223-244) with the [4, ...] headers removed.
// Returns the node count; xyz and tets may be null to query sizes.
uint32_t xr_gen_tet_block(uint32_t W, uint32_t H, uint32_t D, float sx, float sy, float sz, uint32_t pattern,
                          float* xyz, uint32_t* tets) {
    size_t bytes = (size_t)64 * (W + 1) * (H + 1) * (D + 1) + (size_t)256 * W * H * D + (1 << 16);
    std::vector<uint8_t> mem(bytes);
    Allocator alloc;
    alloc.Initialize(mem.data(), bytes);
    const float* nodes = nullptr;
    const uint32_t* idx = nullptr;
    uint32_t nodeCount = 0, idxCount = 0;
    GenerateTetBlock(&alloc, W, H, D, vec3(sx, sy, sz), pattern, 0.0f, &nodes, &nodeCount, &idx, &idxCount);
    if (xyz) memcpy(xyz, nodes, sizeof(float) * nodeCount);
    if (tets)
        for (uint32_t t = 0; t < idxCount / 5; t++)
            for (int c = 0; c < 4; c++) tets[4 * t + c] = idx[5 * t + 1 + c];
    return nodeCount / 3;
}

// LDLTDecomposeAndSolve (Ldlt.h:88-95).
void xr_ldlt(uint32_t n, const float* A, const float* b, float* x) {
    float Ac[325], bc[25];
    memcpy(Ac, A, sizeof(float) * (n * (n + 1) / 2));
    memcpy(bc, b, sizeof(float) * n);
    LDLTDecomposeAndSolve(n, Ac, bc, x);
}

// GeoLinear3d::CalculateVolume (Geo.cpp:827-832)
float xr_volume(void* h) { return ((RefBody*)h)->geo->CalculateVolume(); }

// Vertex rings (GeoLinear3d::rings and vOrder, Geo.h:158-162; one per
// Connectivity3d vertex).
uint32_t xr_ring_count(void* h) { return ((RefBody*)h)->geo->con.vertCount; }

void xr_get_vorder(void* h, uint32_t* order) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    memcpy(order, g->vOrder, sizeof(uint32_t) * g->con.vertCount);
}

void xr_set_vorder(void* h, const uint32_t* order) {
    GeoLinear3d* g = ((RefBody*)h)->geo;
    memcpy(g->vOrder, order, sizeof(uint32_t) * g->con.vertCount);
}

// OneRing of vertex v: counts always; nodes (Verts), tets (Polys) and the
// ring-local node indices (Idxs, 4 per tet) where the pointers are set.
void xr_get_ring(void* h, uint32_t v, uint32_t* n_nodes, uint32_t* n_tets, uint32_t* nodes, uint32_t* tets,
                 uint16_t* idx) {
    const OneRing* r = ((RefBody*)h)->geo->rings[v];
    *n_nodes = r->vertCount;
    *n_tets = r->polyCount;
    if (nodes) memcpy(nodes, r->Verts(), sizeof(uint32_t) * r->vertCount);
    if (tets) memcpy(tets, r->Polys(), sizeof(uint32_t) * r->polyCount);
    if (idx) memcpy(idx, r->Idxs(), sizeof(uint16_t) * r->idxCount);
}

}  // extern "C"
