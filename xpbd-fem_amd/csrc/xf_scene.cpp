// xf_scene.cpp -- C ABI of the gfx950 XPBD substep (include/xpbdfem.h).
//
// A scene holds any number of bodies in one device layout: their nodes are
// concatenated, every body is coloured on its own, and colour c of the
// scene sweeps colour c of all bodies in one launch. One body is the plain
// method; many bodies are the batched one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xpbdfem.h"
#include "xf_device.h"
#include "xf_host_util.h"
#include "xf_sched.h"

extern "C" hipError_t xf_launch_project(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t block,
                                        uint32_t slabs, uint32_t begin, uint32_t count, const uint4* tidx,
                                        const float4* tg0,
                                        const float4* tg1, const float4* tg2, const float4* tg3, const float4* taux,
                                        double4* X, const double4* O, const XfBodyParams* bodies, float comp1_u,
                                        float a_u, float dt, float* clock);
extern "C" hipError_t xf_launch_damp(hipStream_t s, int energy, bool aux, uint32_t block, uint32_t slabs,
                                     uint32_t begin, uint32_t count, const uint4* tidx, const float4* tg0, const float4* tg1,
                                     const float4* tg2, const float4* tg3, const float4* taux, const float2* tdmp,
                                     const double4* X, double4* V, const XfBodyParams* bodies, float comp1_u,
                                     float a_u, float dt, const float* clock);
extern "C" hipError_t xf_launch_pbd(hipStream_t s, bool aux, uint32_t block, uint32_t slabs, uint32_t begin,
                                    uint32_t count,
                                    const uint4* tidx, const float4* taux, const float2* tdmp, const double4* X,
                                    double4* V, const XfBodyParams* bodies, const float* clock);
extern "C" hipError_t xf_launch_volume(hipStream_t s, uint32_t block, uint32_t slabs, uint32_t begin, uint32_t count,
                                       uint32_t pass,
                                       const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                       const float4* taux, double4* X, const XfBodyParams* bodies, float dt);
extern "C" hipError_t xf_launch_rings(hipStream_t s, bool post, uint32_t lanes, uint32_t teams, uint32_t slabs,
                                      uint32_t begin, uint32_t count, uint32_t nmax, uint32_t tmax,
                                      const uint4* rmeta, const uint4* rinfo, const uint32_t* rnodes,
                                      const uint16_t* rincb, const uint16_t* rinc, const float4* rtet, double4* X,
                                      const double4* O, double4* V, const XfBodyParams* bodies, float dt,
                                      float* clock);
extern "C" size_t xf_ring_team_bytes(uint32_t nmax, uint32_t tmax);
extern "C" size_t xf_block_lds_bytes(uint32_t nmax, uint32_t ns);
extern "C" size_t xf_block_fin_lds_bytes(uint32_t nmax, uint32_t fmax, uint32_t ns);
extern "C" hipError_t xf_launch_blocks(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t threads,
                                       uint32_t nslot, uint32_t nph, uint32_t nmax, uint32_t ntet, uint32_t pmax,
                                       const uint32_t* rec,
                                       const uint32_t* bnode, const uint2* tloc, const uint4* tidx,
                                       const float4* tg0, const float4* tg1, const float4* tg2, const float4* tg3,
                                       const float4* taux, double4* X, double4* O,
                                       const XfBodyParams* bodies, float comp1_u, float a_u, float dt,
                                       float* clock, bool fin, const uint32_t* fextra, uint32_t fmax,
                                       const float* tend, bool single_body, bool quad, size_t lds_limit,
                                       bool wide, uint32_t* nstride);
// the same launch with the kernels compiled with multiply-add contraction
// (xf_blocks_contract.hip)
extern "C" hipError_t xf_launch_blocks_contract(
    hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t threads, uint32_t nslot, uint32_t nph,
    uint32_t nmax, uint32_t ntet, uint32_t pmax, const uint32_t* rec, const uint32_t* bnode, const uint2* tloc,
    const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2, const float4* tg3, const float4* taux,
    double4* X, double4* O, const XfBodyParams* bodies, float comp1_u, float a_u, float dt, float* clock, bool fin,
    const uint32_t* fextra, uint32_t fmax, const float* tend, bool single_body, bool quad, size_t lds_limit,
    bool wide, uint32_t* nstride);
extern "C" hipError_t xf_launch_predict(hipStream_t s, uint32_t n, double4* X, double4* O, const double4* V,
                                        const XfBodyParams* bodies);
extern "C" hipError_t xf_launch_finish(hipStream_t s, bool predict_next, bool single_body, uint32_t n, double4* X,
                                       double4* O, double4* V, const XfBodyParams* bodies, const float* clock);
extern "C" hipError_t xf_launch_count_nonfinite(hipStream_t s, uint32_t n, const double4* X, uint32_t* count);
extern "C" hipError_t xf_launch_kinematic(hipStream_t s, uint32_t n, const uint32_t* idx, const double4* target,
                                          double4* X, double4* O);
extern "C" hipError_t xf_launch_manip(hipStream_t s, const float4* target, float dt, double4* X);
extern "C" hipError_t xf_frame_max_blocks(int energy, bool sim, bool aux, bool damp, int* max_blocks);
extern "C" hipError_t xf_launch_project_record(hipStream_t s, int energy, bool aux, uint32_t begin, uint32_t count,
                                               const uint4* tidx, const float4* tg0, const float4* tg1,
                                               const float4* tg2, const float4* tg3, const float4* taux, double4* X,
                                               double* tape, const XfBodyParams* bodies, float comp1_u, float a_u,
                                               float dt, float* clock);
extern "C" hipError_t xf_launch_project_reverse(hipStream_t s, bool aux, bool fp64, uint32_t begin, uint32_t count,
                                                const uint4* tidx, const float4* tg0, const float4* tg1,
                                                const float4* tg2, const float4* tg3, const float4* taux,
                                                const double* tape, const double4* W, double4* Xb, void* gacc,
                                                const XfBodyParams* bodies, float comp1_u, float a_u, float dt);
extern "C" hipError_t xf_launch_adj_pair(hipStream_t s, bool aux, bool fp64, uint32_t rbegin, uint32_t rcount,
                                         double4* X, double* tape_rec, uint32_t vbegin, uint32_t vcount,
                                         const double* tape_rev, const double4* W, double4* Xb, void* gacc,
                                         const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                         const float4* tg3, const float4* taux, const XfBodyParams* bodies,
                                         float comp1_u, float a_u, float dt);
extern "C" hipError_t xf_launch_adj_nodes(hipStream_t s, bool fp64, bool start, bool end, uint32_t n,
                                          const double4* O, const double4* xs, const double4* o, double4* Xb,
                                          double4* Vb, double4* Ob, const XfBodyParams* bodies,
                                          const double4* xloss, const double4* target, const double* lw);
extern "C" hipError_t xf_launch_finish_snap(hipStream_t s, bool predict_next, bool single_body, uint32_t n,
                                            double4* X, double4* O, double4* V, const XfBodyParams* bodies,
                                            const float* clock, double4* xs_out, double4* oo_out);
extern "C" hipError_t xf_launch_adj_flush(hipStream_t s, bool fp64, uint32_t t, void* gacc, double* glog);
extern "C" hipError_t xf_launch_predict64(hipStream_t s, uint32_t n, double4* X, double4* O, const double4* V,
                                          const XfBodyParams* bodies, float dt, float* clock);
extern "C" hipError_t xf_launch_project64(hipStream_t s, bool aux, bool rec, uint32_t begin, uint32_t count,
                                          const uint4* tidx, const float4* tg0, const float4* tg1, const float4* tg2,
                                          const float4* tg3, const float4* taux, double4* X, double* tape,
                                          const XfBodyParams* bodies, float comp1_u, float a_u, float dt);
extern "C" hipError_t xf_launch_finish64(hipStream_t s, uint32_t n, double4* X, const double4* O, double4* V,
                                         const XfBodyParams* bodies, double4* xs_out, double4* oo_out);
extern "C" hipError_t xf_launch_zero_d4(hipStream_t s, uint32_t n, double4* a);
extern "C" hipError_t xf_launch_zero_u32(hipStream_t s, uint32_t n, uint32_t* a);
extern "C" hipError_t xf_launch_expand_targets(hipStream_t s, uint32_t n, const double* raw, double4* out);
extern "C" hipError_t xf_launch_loss_weights(hipStream_t s, uint32_t n, const double4* O, const XfBodyParams* bodies,
                                             double* lw);
extern "C" hipError_t xf_launch_loss(hipStream_t s, uint32_t n_loss, uint32_t nb, uint32_t n, const double4* xloss,
                                     const double4* tgt, const XfBodyParams* bodies, double* e, double* out);
extern "C" hipError_t xf_launch_grad_out(hipStream_t s, uint32_t t, const double* glog, const uint32_t* tout,
                                         double* g);
extern "C" hipError_t xf_launch_copy2_d4(hipStream_t s, uint32_t n, double4* d0, const double4* s0, double4* d1,
                                         const double4* s1);
extern "C" hipError_t xf_launch_frame(hipStream_t s, int energy, bool sim, bool aux, bool damp, uint32_t blocks,
                                      uint32_t substeps, uint32_t ncol, uint32_t vpasses, uint32_t post,
                                      const float2* tdmp, const uint32_t* colour_begin,
                                      uint32_t n_nodes, const uint4* tidx, const float4* tg0, const float4* tg1,
                                      const float4* tg2, const float4* tg3, const float4* taux, double4* X,
                                      double4* O, double4* V, const XfBodyParams* bodies, float comp1_u, float a_u,
                                      float dt, float* clock, uint32_t* arrive, uint32_t* err);

struct xf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned char* scratch = nullptr;  // this device's copy of xf_adjoint_scratch (device address)
    std::string err;
    std::vector<xf_scene*> scenes;  // open scenes, destroyed with the context
};

// Device scratch of xf_scene_adjoint_noalloc: checkpoints, the window's
// sweep tape, adjoint state, loss staging. Each device holds its own copy
// for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress reports for the device (the name itself is a
// host-side stand-in). Calls that overlap on one device share it and race.
constexpr size_t kAdjointScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char xf_adjoint_scratch[kAdjointScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Resolves this device's scratch address once per device (the device must
// be current).
hipError_t ResolveScratch(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(xf_adjoint_scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

int Fail(xf_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define XF_HIP(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return Fail((ctx), XF_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

struct Body {
    uint32_t n = 0, t = 0;
    float density = 0.0f;
    std::vector<float> xyz, vel, tet_mu;
    std::vector<double> xyz64;  // FP64 positions when the caller gave them (xyz then holds their float values)
    std::vector<uint32_t> tets;
    std::vector<uint32_t> user_colour;  // caller-supplied colouring (empty: greedy)
    std::vector<uint8_t> pinned;
    bool pin_first_substep = false;  // pins keep 1/m until their first locked substep (Geo.cpp:318-322)
    // kinematic targets (body-local node, FP64 position) and manipulator
    std::vector<uint32_t> kin_nodes;
    std::vector<double> kin_xyz;
    xf_settings settings{};
    // finalize products
    xf::ElementConstants ec;
    std::vector<uint32_t> colour;     // per local tet
    std::vector<uint32_t> order;      // local tets in sweep order
    std::vector<uint32_t> sched_pos;  // scene schedule position of each local tet
    uint32_t ncol = 0;
    uint32_t node_base = 0;
    // vertex rings (built on first use of a vertex-ring energy)
    xf::Rings rings;
    std::vector<uint32_t> ring_colour;  // per local ring
    std::vector<uint32_t> ring_order;   // local rings in sweep order
};

template <typename T>
int DevAlloc(xf_ctx* ctx, T** p, size_t count, uint64_t* total) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) return Fail(ctx, XF_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", sizeof(T) * count,
                                     hipGetErrorString(e));
    *total += sizeof(T) * count;
    return XF_OK;
}

// Settings energy field: 5 bits (Settings_EnergyMask), every value dispatched
// as TSolveElement does for linear tets (Fem.cpp:870-889).
bool EnergySupported(uint32_t e) { return e <= 31; }

// Kernel variant of a tet energy: 0/1 Pixar / PixarReduced (one quadrature
// point on a linear tet: the same code) -> 20, PixarSel -> 22, CubeNeo /
// CubeSkin (quads and hexes only: no constraint on tets) -> 30, 3..7 as
// they are, anything above 12 the switch's default MixedSelective -> 4.
uint32_t KernelEnergy(uint32_t e) {
    if (e <= 1) return 20;
    if (e == 2) return 22;
    if (e == 11 || e == 12) return 30;
    if (e > 12) return 4;
    return e;
}

bool PixarEnergy(uint32_t e) { return e <= 2; }

// Energy_ContinuousPixar / ContinuousMixed / ContinuousSkin sweep vertex rings
// instead of tets (GeoLinear3d::Constrain, Geo.cpp:774-788).
bool RingEnergy(uint32_t e) { return e >= XF_ENERGY_CONTINUOUS_PIXAR && e <= XF_ENERGY_CONTINUOUS_SKIN; }

// Scene-wide mode value of ResolveMode when every body uses a ring energy.
constexpr int kRingMode = XF_ENERGY_CONTINUOUS_PIXAR;

// LDS budget of one ring workgroup.
constexpr size_t kRingLdsLimit = 65536;

// XCD-slab order of a colour-major sweep order: inside each colour the items
// are sorted by x (ties keep their order) and cut into the eighths the
// per-colour launch maps to XCDs (positions [n r / 8, n (r + 1) / 8)); inside
// an eighth they are then sorted by their smallest node index, so that
// neighbouring lanes read neighbouring node records.
void SlabOrder(std::vector<uint32_t>* order, const std::vector<uint32_t>& colour, const std::vector<float>& x,
               const std::vector<uint32_t>& first_node) {
    std::vector<uint32_t>& o = *order;
    std::stable_sort(o.begin(), o.end(), [&](uint32_t a, uint32_t c) {
        return colour[a] != colour[c] ? colour[a] < colour[c] : x[a] < x[c];
    });
    for (size_t b = 0; b < o.size();) {
        size_t e = b;
        while (e < o.size() && colour[o[e]] == colour[o[b]]) e++;
        uint64_t n = e - b;
        for (uint32_t r = 0; r < 8; r++) {
            size_t lo = b + (size_t)((n * r) >> 3), hi = b + (size_t)((n * (r + 1)) >> 3);
            std::stable_sort(o.begin() + lo, o.begin() + hi,
                             [&](uint32_t a, uint32_t c) { return first_node[a] < first_node[c]; });
        }
        b = e;
    }
}

// Volume passes per substep accepted by xf_scene_set_settings.
constexpr uint32_t kMaxVolumePasses = 64;

}  // namespace

struct xf_scene {
    xf_ctx* ctx = nullptr;
    std::vector<Body> bodies;
    bool finalized = false;
    xf_scene_opts opts{};
    uint32_t N = 0, T = 0, ncol = 0;
    std::vector<uint32_t> colour_begin;  // ncol + 1 offsets into the schedule
    // device
    double4 *X = nullptr, *O = nullptr, *V = nullptr;
    uint4* tidx = nullptr;
    float4* tg[4] = {nullptr, nullptr, nullptr, nullptr};
    float4* taux = nullptr;
    float2* tdmp = nullptr;  // per tet (surface area, position in the body's sweep order)
    XfBodyParams* bp = nullptr;
    float* clock = nullptr;  // kClockWords: time, tick
    uint32_t* counter = nullptr;
    uint32_t* colour_begin_d = nullptr;  // ncol + 1 offsets (persistent launch)
    uint32_t* arrive = nullptr;          // grid-barrier arrivals (persistent launch)
    uint32_t* barrier_err = nullptr;     // set when a grid barrier timed out
    // vertex rings in sweep order (colour-major, body-major inside a colour)
    bool rings_host = false, rings_dev = false;
    uint32_t R = 0, rcol = 0, rnmax = 0, rtmax = 0;
    std::vector<uint32_t> ring_colour_begin;
    uint4* rmeta = nullptr;
    uint4* rinfo = nullptr;  // body id, position in the body's ring order, ring volume bits, 0
    uint32_t* rnodes = nullptr;
    uint16_t* rincb = nullptr;  // per ring node: first entry of its incidence list (ring-local)
    uint16_t* rinc = nullptr;   // per ring: (ring tet << 2 | corner) of each node, node-major, tets ascending
    float4* rtet = nullptr;
    // sweep blocks (opts.sweep_blocks): per launch colour q, nslot blocks
    // (body-major, then block number) and nph phases; blk_rec + rec_off holds
    // per block nph (first schedule position, count) pairs, blk_node +
    // node_off per block blk_nmax global node ids (xf::kNoNode padded); tloc
    // holds four block-local node indices per schedule position
    struct BlockLaunch {
        uint32_t nslot = 0, nph = 0, rec_off = 0, node_off = 0;
    };
    BlockLaunch blk[4];
    uint32_t n_blocks = 0, blk_nmax = 0, blk_nodes = 0, blk_pmax = 0, blk_lanes = 0;
    uint32_t* blk_rec = nullptr;
    uint32_t* blk_node = nullptr;
    size_t blk_rec_n = 0, blk_node_n = 0, blk_fin_n = 0;  // their allocated lengths
    size_t lds_limit = 65536;  // LDS bytes one workgroup may use on the device
    // the blocks of every body and the first colour of each launch colour,
    // kept for repacking the slots (xf_scene_autotune)
    std::vector<xf::SweepBlocks> sweep;
    uint32_t blk_base[4] = {0, 0, 0, 0};
    // fused node phase: the last launch colour with blocks (blk_last; -1 with
    // fewer than two such colours, more than two of its slots per compute
    // unit, or when its nodes with their O and ids do not fit LDS) ends the
    // substep; blk_fin holds per slot of that launch its
    // blk_nmax block nodes and blk_fmax nodes that none of the colour's
    // blocks touches
    int blk_last = -1;
    uint32_t blk_fmax = 0;
    // 48 bytes per slot node in LDS allowed (fewer bank conflicts): every
    // block launch has at most two workgroups per compute unit, so LDS is not
    // what limits them; blk_nstride: the bytes per node of the last launch
    bool blk_wide = false;
    uint32_t blk_nstride = 32;
    uint32_t* blk_fin = nullptr;
    uint2* tloc = nullptr;
    uint32_t* tout = nullptr;  // per schedule position: the tet's slot in body-concatenated, body-local order
    // kinematic targets of all bodies (global node ids, FP64 positions) and
    // the manipulator (target xyz, node id bits in w; kNoNode: none)
    uint32_t nkin = 0, kin_cap = 0;
    uint32_t* kin_idx = nullptr;
    double4* kin_pos = nullptr;
    float4* manip = nullptr;
    bool manip_on = false;
    uint64_t device_bytes = 0;
    int n_cu = 256;
    // host mirrors
    std::vector<float4> h_tg3, h_taux;
    std::vector<double> h_w;       // per node inverse mass
    std::vector<double> h_oflags;  // per node O.w: (body << 1) | pinned
    std::vector<XfBodyParams> h_bp;  // body table staging (sized at finalize)
    bool material_dirty = true, params_dirty = true;
    bool aux = false;
    uint32_t vpasses = 0;  // max volume passes over bodies (taux is then always present)
    bool dproj = false;    // some body damps inside the projection (Rayleigh Paper / Limit)
    uint32_t post = 0;     // damping passes after the velocity update: bit 0 Rayleigh post, bit 1 PBD
    float comp1_u = 0.0f, a_u = 1.0f;
    float params_dt = -1.0f;
    float time_host = 0.0f;
    uint32_t tick_host = 0;
    // adjoint buffers (grown on demand, freed with the scene)
    struct AdjBuf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    AdjBuf adj_ck_x, adj_ck_v, adj_tape, adj_xs, adj_o, adj_xb, adj_vb, adj_ob, adj_gacc, adj_glog, adj_xloss,
        adj_tgt, adj_lw, adj_save_x, adj_save_v, adj_traw, adj_eterm, adj_lout, adj_gout;
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_substeps = 0;
    float graph_dt = 0.0f;
    int graph_energy = -1;
    bool graph_sim = false, graph_aux = false, graph_dproj = false;
    uint32_t graph_vpasses = 0, graph_post = 0, graph_nkin = 0;
    bool graph_manip = false;
    // graph of the adjoint's reverse pass (window replays and reverse
    // sweeps), kept for the call shape in adj_key (sizes, loss substeps,
    // precision, scene constants baked into the launches, buffer addresses)
    hipGraphExec_t adj_exec = nullptr;
    std::vector<uint64_t> adj_key;
};

namespace {

// Called when the step call shape changes: a launch of the old executable
// graph may still be running on the stream (xf_scene_step returns once the
// work is enqueued), so the stream drains before the graph is destroyed.
void DropAdjGraph(xf_scene* s) {
    if (s->adj_exec) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
        (void)hipGraphExecDestroy(s->adj_exec);
    }
    s->adj_exec = nullptr;
    s->adj_key.clear();
}

void DropGraph(xf_scene* s) {
    if (s->graph_exec) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
        (void)hipGraphExecDestroy(s->graph_exec);
    }
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    s->graph_substeps = 0;
}

// Scene-wide projection variant: one energy for all bodies, or per-body
// MixedSel / YeohSkinFast (energy 0). Returns XF_OK or an error.
int ResolveMode(xf_scene* s, int* energy, bool* sim) {
    size_t n_ring = 0;
    for (const Body& b : s->bodies) n_ring += RingEnergy(b.settings.energy) ? 1 : 0;
    if (n_ring == s->bodies.size()) {
        *energy = kRingMode;
        *sim = false;
        return XF_OK;
    }
    if (n_ring > 0)
        return Fail(s->ctx, XF_E_UNSUPPORTED,
                    "bodies of one scene must all use vertex-ring energies (8..10) or all tet energies (3..7)");
    uint32_t e0 = KernelEnergy(s->bodies[0].settings.energy);
    uint32_t s0 = s->bodies[0].settings.simultaneous ? 1u : 0u;
    bool same = true, all47 = true;
    for (const Body& b : s->bodies) {
        uint32_t e = KernelEnergy(b.settings.energy);
        uint32_t si = b.settings.simultaneous ? 1u : 0u;
        if (e != e0 || si != s0) same = false;
        if (!((e == 4 || e == 7) && si == 1)) all47 = false;
    }
    if (same) {
        *energy = (int)e0;
        *sim = s0 != 0;
        return XF_OK;
    }
    if (all47) {
        *energy = 0;
        *sim = true;
        return XF_OK;
    }
    return Fail(s->ctx, XF_E_UNSUPPORTED,
                "bodies of one scene must share energy and solve mode, or all use simultaneous MixedSel/YeohSkinFast");
}

// Per-tet material constants of SolveElementMixed (Fem.cpp:445-449).
void ComputeMaterial(xf_scene* s) {
    bool compressible = false;
    for (Body& b : s->bodies) {
        float nu = b.settings.poissons_ratio;
        bool pixar = PixarEnergy(b.settings.energy);
        for (uint32_t t = 0; t < b.t; t++) {
            float mu = b.tet_mu.empty() ? 1.0f / b.settings.compliance : b.tet_mu[t];
            float lambda = (2.0f * mu * nu) / (1.0f - 2.0f * nu);
            float a = 1.0f + mu / lambda;
            float comp0 = 1.0f / mu / b.ec.vol[t];
            float comp1 = 1.0f / lambda / b.ec.vol[t];
            if (pixar) {
                // SolveElementPixar (Fem.cpp:367-370): mu = 2; the slots hold
                // compliance / V, lambda and a
                float pmu = 2.0f;
                float plambda = pmu / (1.0f - 2.0f * nu);
                comp0 = b.settings.compliance / b.ec.vol[t];
                comp1 = plambda;
                a = 1.0f + pmu / plambda;
            }
            uint32_t pos = b.sched_pos[t];
            s->h_tg3[pos].w = comp0;
            s->h_taux[pos].x = comp1;
            s->h_taux[pos].y = a;
            s->h_taux[pos].w = b.settings.compliance * b.ec.vol[t];  // SolveVolumeOnly (Fem.cpp:845)
            if (comp1 != 0.0f || a != 1.0f) compressible = true;
        }
    }
    s->aux = compressible || s->bodies.size() > 1;
    s->vpasses = 0;
    s->dproj = false;
    s->post = 0;
    for (const Body& b : s->bodies) {
        const xf_settings& st = b.settings;
        s->vpasses = std::max(s->vpasses, st.volume_passes);
        if (st.damping > 0.0f && st.rayleigh_type < XF_RAYLEIGH_POST) s->dproj = true;
        if (st.damping > 0.0f && st.rayleigh_type >= XF_RAYLEIGH_POST) s->post |= 1u;
        if (st.pbd_damping > 0.0f) s->post |= 2u;
    }
}

int UploadMaterial(xf_scene* s) {
    xf_ctx* ctx = s->ctx;
    bool had_aux = s->aux;
    ComputeMaterial(s);
    if (s->aux != had_aux) DropGraph(s);
    bool need_taux = s->aux || s->vpasses > 0;
    if (need_taux && !s->taux) {
        int rc = DevAlloc(ctx, &s->taux, s->T, &s->device_bytes);
        if (rc) return rc;
    }
    XF_HIP(ctx, hipMemcpyAsync(s->tg[3], s->h_tg3.data(), sizeof(float4) * s->T, hipMemcpyHostToDevice, ctx->stream));
    if (need_taux)
        XF_HIP(ctx, hipMemcpyAsync(s->taux, s->h_taux.data(), sizeof(float4) * s->T, hipMemcpyHostToDevice,
                                   ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->material_dirty = false;
    return XF_OK;
}

int UploadParams(xf_scene* s, float dt) {
    xf_ctx* ctx = s->ctx;
    std::vector<XfBodyParams>& hp = s->h_bp;
    hp.resize(s->bodies.size());
    for (size_t i = 0; i < s->bodies.size(); i++) {
        const xf_settings& st = s->bodies[i].settings;
        XfBodyParams& p = hp[i];
        memset(&p, 0, sizeof(p));
        p.gdx = st.gravity[0] * dt;
        p.gdy = st.gravity[1] * dt;
        p.dragf = 1.0f - st.drag;
        p.dt = dt;
        p.invdt = 1.0f / dt;
        p.collide = st.collide;
        p.ground_y = st.ground_y;
        p.friction = st.friction;
        for (int r = 0; r < 3; r++) {
            p.sc0[r] = st.sphere_c0[r];
            p.sv[r] = st.sphere_v[r];
            p.yeoh[r] = st.yeoh[r];
        }
        p.sr = st.sphere_r;
        p.energy = st.energy;
        p.volume_passes = st.volume_passes;
        p.damping = st.damping;
        p.rtype = st.rayleigh_type;
        p.pbd = st.pbd_damping;
        p.pbd_amortized = st.pbd_damping_amortized;
        p.n_tets = s->bodies[i].t;
        p.simultaneous = st.simultaneous ? 1u : 0u;
        p.n_rings = s->bodies[i].rings.n;
        p.compliance = st.compliance;
        p.poisson = st.poissons_ratio;
        p.node_base = s->bodies[i].node_base;
        p.n_nodes = s->bodies[i].n;
        p.unlock = st.unlock_pins;
        p.gravity[0] = st.gravity[0];
        p.gravity[1] = st.gravity[1];
        p.drag = st.drag;
    }
    XF_HIP(ctx, hipMemcpyAsync(s->bp, hp.data(), sizeof(XfBodyParams) * hp.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->params_dt = dt;
    s->params_dirty = false;
    return XF_OK;
}

// Host side of the vertex rings: per body the reference's OneRings, a greedy
// colouring of the ring conflict graph visited in vertex order (or the
// reference's vOrder with XF_COLOUR_MINSTD), and with xcd_slabs the rings of
// each colour ordered by vertex x.
int BuildRingsHost(xf_scene* s) {
    if (s->rings_host) return XF_OK;
    xf_ctx* ctx = s->ctx;
    s->rcol = 0;
    s->rnmax = 0;
    s->rtmax = 0;
    s->R = 0;
    for (Body& b : s->bodies) {
        if (!xf::BuildRings(b.tets.data(), b.t, &b.rings))
            return Fail(ctx, XF_E_UNSUPPORTED, "a vertex ring has more than %u nodes or %u tets", xf::kMaxRingNodes,
                        xf::kMaxRingTets);
        std::vector<uint32_t> visit(b.rings.n);
        if (s->opts.colour_order == XF_COLOUR_MINSTD) xf::MinstdRingOrder(b.t, b.rings.n, visit.data());
        else
            for (uint32_t v = 0; v < b.rings.n; v++) visit[v] = v;
        uint32_t nc = xf::RingColourSweepOrder(b.rings, b.n, visit.data(), &b.ring_colour, &b.ring_order);
        if (nc == 0) return Fail(ctx, XF_E_UNSUPPORTED, "mesh needs more than %u ring colours", xf::kMaxColours);
        if (s->opts.xcd_slabs) {
            std::vector<float> vx(b.rings.n);
            std::vector<uint32_t> vid(b.rings.n);
            for (uint32_t v = 0; v < b.rings.n; v++) {
                vx[v] = b.xyz[3 * (size_t)v];
                vid[v] = v;
            }
            SlabOrder(&b.ring_order, b.ring_colour, vx, vid);
        }
        s->rcol = std::max(s->rcol, nc);
        s->rnmax = std::max(s->rnmax, b.rings.largest);
        for (uint32_t v = 0; v < b.rings.n; v++) s->rtmax = std::max(s->rtmax, b.rings.tet_off[v + 1] - b.rings.tet_off[v]);
        s->R += b.rings.n;
    }
    s->rings_host = true;
    return XF_OK;
}

// Device side: ring schedule, node lists and per-ring-tet constants.
int UploadRings(xf_scene* s) {
    if (s->rings_dev) return XF_OK;
    int rc = BuildRingsHost(s);
    if (rc) return rc;
    xf_ctx* ctx = s->ctx;
    std::vector<uint4> meta(s->R);
    std::vector<uint4> info(s->R);
    std::vector<uint32_t> nodes;
    std::vector<uint16_t> incb, inc;
    std::vector<float4> rtet;
    s->ring_colour_begin.assign(s->rcol + 1, 0);
    std::vector<uint32_t> cursor(s->bodies.size(), 0);
    uint32_t pos = 0;
    for (uint32_t c = 0; c < s->rcol; c++) {
        s->ring_colour_begin[c] = pos;
        for (size_t bi = 0; bi < s->bodies.size(); bi++) {
            const Body& b = s->bodies[bi];
            const xf::Rings& r = b.rings;
            uint32_t& k = cursor[bi];
            while (k < r.n && b.ring_colour[b.ring_order[k]] == c) {
                uint32_t v = b.ring_order[k++];
                uint32_t n0 = r.node_off[v], n1 = r.node_off[v + 1], t0 = r.tet_off[v], t1 = r.tet_off[v + 1];
                meta[pos] = make_uint4((uint32_t)nodes.size(), n1 - n0, (uint32_t)(rtet.size() / 2), t1 - t0);
                // ring volume: sum of V_t / 4 in ring-tet order (SolveRing, Geo.cpp:160-163)
                float vsum = 0.0f;
                for (uint32_t q = t0; q < t1; q++) vsum = vsum + 0.25f * b.ec.vol[r.tets[q]];
                uint32_t vbits;
                memcpy(&vbits, &vsum, 4);
                info[pos] = make_uint4((uint32_t)bi, k - 1, vbits, 0u);
                for (uint32_t q = n0; q < n1; q++) nodes.push_back(b.node_base + r.nodes[q]);
                // incidence lists: for each ring node, its (tet, corner) slots in ring-tet order
                std::vector<uint16_t> cnt(n1 - n0 + 1, 0);
                for (uint32_t q = 4 * t0; q < 4 * t1; q++) cnt[r.idx[q] + 1]++;
                for (uint32_t n = 0; n < n1 - n0; n++) cnt[n + 1] += cnt[n];
                incb.insert(incb.end(), cnt.begin(), cnt.end() - 1);
                size_t base = inc.size();
                inc.resize(base + 4 * (size_t)(t1 - t0));
                for (uint32_t q = 4 * t0; q < 4 * t1; q++) inc[base + cnt[r.idx[q]]++] = (uint16_t)(q - 4 * t0);
                for (uint32_t q = t0; q < t1; q++) {
                    uint32_t t = r.tets[q];
                    const float* qq = &b.ec.qq[3 * (size_t)t];
                    const float* qr = &b.ec.qr[3 * (size_t)t];
                    const uint8_t* li = &r.idx[4 * (size_t)q];
                    uint32_t bits = (uint32_t)li[0] | ((uint32_t)li[1] << 8) | ((uint32_t)li[2] << 16) |
                                    ((uint32_t)li[3] << 24);
                    float bitsf;
                    memcpy(&bitsf, &bits, 4);
                    rtet.push_back(make_float4(qq[0], qq[1], qq[2], qr[0]));
                    rtet.push_back(make_float4(qr[1], qr[2], b.ec.vol[t], bitsf));
                }
                ++pos;
            }
        }
    }
    s->ring_colour_begin[s->rcol] = pos;
    if (pos != s->R) return Fail(ctx, XF_E_INVALID, "internal: ring schedule covers %u of %u rings", pos, s->R);
    // One padding entry past the end of each list: the ring kernel loads a
    // lane's first entries at clamped positions, which for a ring without
    // nodes or tets (or a lane without a ring) is the ring's start offset.
    nodes.push_back(0u);
    incb.push_back(0u);
    inc.insert(inc.end(), 4, (uint16_t)0u);
    rtet.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    rtet.push_back(make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    rc = DevAlloc(ctx, &s->rmeta, s->R, &s->device_bytes);
    if (!rc) rc = DevAlloc(ctx, &s->rinfo, s->R, &s->device_bytes);
    if (!rc) rc = DevAlloc(ctx, &s->rnodes, nodes.size(), &s->device_bytes);
    if (!rc) rc = DevAlloc(ctx, &s->rincb, incb.size(), &s->device_bytes);
    if (!rc) rc = DevAlloc(ctx, &s->rinc, inc.size(), &s->device_bytes);
    if (!rc) rc = DevAlloc(ctx, &s->rtet, rtet.size(), &s->device_bytes);
    if (rc) return rc;
    XF_HIP(ctx, hipMemcpyAsync(s->rincb, incb.data(), sizeof(uint16_t) * incb.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    XF_HIP(ctx, hipMemcpyAsync(s->rinc, inc.data(), sizeof(uint16_t) * inc.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    XF_HIP(ctx, hipMemcpyAsync(s->rmeta, meta.data(), sizeof(uint4) * s->R, hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, hipMemcpyAsync(s->rinfo, info.data(), sizeof(uint4) * s->R, hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, hipMemcpyAsync(s->rnodes, nodes.data(), sizeof(uint32_t) * nodes.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    XF_HIP(ctx, hipMemcpyAsync(s->rtet, rtet.data(), sizeof(float4) * rtet.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->rings_dev = true;
    s->params_dirty = true;  // n_rings per body
    return XF_OK;
}

// Enqueues `substeps` substeps on the context stream. ev (nullable) gets an
// event pair per launch for profiling: [2*k] before, [2*k+1] after.
struct LaunchTimer {
    std::vector<hipEvent_t> ev_proj, ev_node;
};

// After the sweep of a substep: nodes held at kinematic targets (LockRight,
// Geo.cpp:324-331), then the manipulation constraint (Geo.cpp:334-339). The
// lock of pinned nodes (Geo.cpp:318-322) runs in the node phase after them;
// it touches other nodes, or overrides them, as the reference's order does.
int HeldNodes(xf_scene* s, float dt) {
    xf_ctx* ctx = s->ctx;
    if (s->nkin) XF_HIP(ctx, xf_launch_kinematic(ctx->stream, s->nkin, s->kin_idx, s->kin_pos, s->X, s->O));
    if (s->manip_on) XF_HIP(ctx, xf_launch_manip(ctx->stream, s->manip, dt, s->X));
    return XF_OK;
}

int Enqueue(xf_scene* s, uint32_t substeps, float dt, int energy, bool sim, LaunchTimer* timer) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    uint32_t block = s->opts.block_size ? s->opts.block_size : 64;
    uint32_t slabs = s->opts.xcd_slabs;
    auto node = [&](int kind, bool predict_next) -> int {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timer) {
            XF_HIP(ctx, hipEventCreate(&e0));
            XF_HIP(ctx, hipEventCreate(&e1));
            XF_HIP(ctx, hipEventRecord(e0, st));
        }
        if (kind == 0) XF_HIP(ctx, xf_launch_predict(st, s->N, s->X, s->O, s->V, s->bp));
        else
            XF_HIP(ctx, xf_launch_finish(st, predict_next, s->bodies.size() == 1, s->N, s->X, s->O, s->V, s->bp,
                                         s->clock));
        if (timer) {
            XF_HIP(ctx, hipEventRecord(e1, st));
            timer->ev_node.push_back(e0);
            timer->ev_node.push_back(e1);
        }
        return XF_OK;
    };
    if (energy == kRingMode) {
        // Geo3d::Substep with a vertex-ring energy: ring colours (the first
        // advances the clock), velocity, then GeoLinear3d::Damp's ring pass
        // and PBD tet pass (Geo.cpp:790-811), then the next prediction.
        // lanes per ring: the option, else 32 when rings have more than 16
        // tets (a tet per lane in the gather); rings per workgroup: the
        // scene's workgroup size over the lanes, halved until the rings' LDS
        // fits the budget
        uint32_t lanes = s->opts.ring_lanes ? s->opts.ring_lanes : (s->rtmax > 16 ? 32u : 16u);
        uint32_t rb = std::max(1u, block / lanes);
        while (rb > 1 && rb * xf_ring_team_bytes(s->rnmax, s->rtmax) > kRingLdsLimit) rb /= 2;
        int rc = node(0, false);
        if (rc) return rc;
        for (uint32_t k = 0; k < substeps; k++) {
            for (uint32_t c = 0; c < s->rcol; c++) {
                uint32_t begin = s->ring_colour_begin[c], count = s->ring_colour_begin[c + 1] - begin;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (timer) {
                    XF_HIP(ctx, hipEventCreate(&e0));
                    XF_HIP(ctx, hipEventCreate(&e1));
                    XF_HIP(ctx, hipEventRecord(e0, st));
                }
                XF_HIP(ctx, xf_launch_rings(st, false, lanes, rb, slabs, begin, count, s->rnmax, s->rtmax, s->rmeta,
                                            s->rinfo, s->rnodes, s->rincb, s->rinc, s->rtet, s->X, s->O, s->V,
                                            s->bp, dt, c == 0 ? s->clock : nullptr));
                if (timer) {
                    XF_HIP(ctx, hipEventRecord(e1, st));
                    timer->ev_proj.push_back(e0);
                    timer->ev_proj.push_back(e1);
                }
            }
            rc = HeldNodes(s, dt);
            if (rc) return rc;
            if (s->post == 0) {
                rc = node(1, k + 1 < substeps);
                if (rc) return rc;
                continue;
            }
            rc = node(1, false);
            if (rc) return rc;
            if (s->post & 1u)
                for (uint32_t c = 0; c < s->rcol; c++) {
                    uint32_t begin = s->ring_colour_begin[c], count = s->ring_colour_begin[c + 1] - begin;
                    XF_HIP(ctx, xf_launch_rings(st, true, lanes, rb, slabs, begin, count, s->rnmax, s->rtmax, s->rmeta,
                                                s->rinfo, s->rnodes, s->rincb, s->rinc, s->rtet, s->X, s->O, s->V,
                                                s->bp, dt, s->clock));
                }
            if (s->post & 2u)
                for (uint32_t c = 0; c < s->ncol; c++) {
                    uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
                    XF_HIP(ctx, xf_launch_pbd(st, s->aux, block, slabs, begin, count, s->tidx, s->taux, s->tdmp,
                                              s->X, s->V, s->bp, s->clock));
                }
            if (k + 1 < substeps) {
                rc = node(0, false);
                if (rc) return rc;
            }
        }
        return XF_OK;
    }
    if (s->opts.launch_mode == XF_LAUNCH_PERSISTENT) {
        int maxb = 0;
        XF_HIP(ctx, xf_frame_max_blocks(energy, sim, s->aux, s->dproj, &maxb));
        uint32_t blocks = s->opts.persistent_blocks ? s->opts.persistent_blocks : (uint32_t)s->n_cu;
        if (maxb <= 0) return Fail(ctx, XF_E_UNSUPPORTED, "persistent kernel does not fit on the device");
        if (blocks > (uint32_t)maxb) blocks = (uint32_t)maxb;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timer) {
            XF_HIP(ctx, hipEventCreate(&e0));
            XF_HIP(ctx, hipEventCreate(&e1));
        }
        XF_HIP(ctx, hipMemsetAsync(s->arrive, 0, sizeof(uint32_t), st));
        if (timer) XF_HIP(ctx, hipEventRecord(e0, st));
        XF_HIP(ctx, xf_launch_frame(st, energy, sim, s->aux, s->dproj, blocks, substeps, s->ncol, s->vpasses, s->post,
                                    s->tdmp, s->colour_begin_d, s->N, s->tidx, s->tg[0], s->tg[1], s->tg[2],
                                    s->tg[3], s->taux, s->X, s->O, s->V, s->bp, s->comp1_u, s->a_u, dt, s->clock,
                                    s->arrive, s->barrier_err));
        if (timer) {
            XF_HIP(ctx, hipEventRecord(e1, st));
            timer->ev_proj.push_back(e0);
            timer->ev_proj.push_back(e1);
        }
        return XF_OK;
    }
    int rc = node(0, false);
    if (rc) return rc;
    const bool blocks = s->opts.launch_mode == XF_LAUNCH_BLOCKS;
    // the node phase rides on the substep's last block launch when nothing
    // runs between the sweep and the velocity update (no volume passes), the
    // velocities need not be stored (no damping passes, not the call's last
    // substep) and the projection leaves O unread
    const bool fuse = blocks && s->blk_last >= 0 && !s->opts.separate_finish && s->vpasses == 0 && s->post == 0 &&
                      !s->dproj && energy != 30 && s->nkin == 0 && !s->manip_on;
    for (uint32_t k = 0; k < substeps; k++) {
        const bool fin = fuse && k + 1 < substeps;
        // sweep blocks: one launch per launch colour that holds blocks; the
        // first advances the clock
        bool first = true;
        for (int q = 0; q < 4 && blocks; q++) {
            const xf_scene::BlockLaunch& L = s->blk[q];
            if (L.nslot == 0) continue;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer) {
                XF_HIP(ctx, hipEventCreate(&e0));
                XF_HIP(ctx, hipEventCreate(&e1));
                XF_HIP(ctx, hipEventRecord(e0, st));
            }
            auto launch = s->opts.contract ? xf_launch_blocks_contract : xf_launch_blocks;
            XF_HIP(ctx, launch(st, energy, sim, s->aux, s->dproj, block, L.nslot, L.nph, s->blk_nmax, s->T,
                               s->blk_pmax, s->blk_rec + L.rec_off, s->blk_node + L.node_off, s->tloc, s->tidx,
                               s->tg[0], s->tg[1], s->tg[2], s->tg[3], s->taux, s->X, s->O, s->bp, s->comp1_u,
                               s->a_u, dt, first ? s->clock : nullptr, fin && q == s->blk_last, s->blk_fin,
                               s->blk_fmax, s->clock, s->bodies.size() == 1, s->opts.tet_lanes == 4,
                               s->lds_limit, s->blk_wide, &s->blk_nstride));
            first = false;
            if (timer) {
                XF_HIP(ctx, hipEventRecord(e1, st));
                timer->ev_proj.push_back(e0);
                timer->ev_proj.push_back(e1);
            }
        }
        for (uint32_t c = 0; c < s->ncol && !blocks; c++) {
            uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timer) {
                XF_HIP(ctx, hipEventCreate(&e0));
                XF_HIP(ctx, hipEventCreate(&e1));
                XF_HIP(ctx, hipEventRecord(e0, st));
            }
            XF_HIP(ctx, xf_launch_project(st, energy, sim, s->aux, s->dproj, block, slabs, begin, count, s->tidx, s->tg[0],
                                          s->tg[1], s->tg[2], s->tg[3], s->taux, s->X, s->O, s->bp, s->comp1_u,
                                          s->a_u, dt, c == 0 ? s->clock : nullptr));
            if (timer) {
                XF_HIP(ctx, hipEventRecord(e1, st));
                timer->ev_proj.push_back(e0);
                timer->ev_proj.push_back(e1);
            }
        }
        // volumePasses x SolveElementVolume over the same colour order (Geo.cpp:779-782)
        for (uint32_t pass = 0; pass < s->vpasses; pass++)
            for (uint32_t c = 0; c < s->ncol; c++) {
                uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
                XF_HIP(ctx, xf_launch_volume(st, block, slabs, begin, count, pass, s->tidx, s->tg[0], s->tg[1], s->tg[2],
                                             s->taux, s->X, s->bp, dt));
            }
        rc = HeldNodes(s, dt);
        if (rc) return rc;
        if (s->post == 0) {
            if (fin) continue;
            rc = node(1, k + 1 < substeps);
            if (rc) return rc;
            continue;
        }
        // GeoLinear3d::Damp after the velocity update (Geo.cpp:346-355, 790-811):
        // velocities are stored, damped colour by colour, then the next prediction.
        rc = node(1, false);
        if (rc) return rc;
        if (s->post & 1u)
            for (uint32_t c = 0; c < s->ncol; c++) {
                uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
                XF_HIP(ctx, xf_launch_damp(st, energy, s->aux, block, slabs, begin, count, s->tidx, s->tg[0], s->tg[1],
                                           s->tg[2], s->tg[3], s->taux, s->tdmp, s->X, s->V, s->bp, s->comp1_u,
                                           s->a_u, dt, s->clock));
            }
        if (s->post & 2u)
            for (uint32_t c = 0; c < s->ncol; c++) {
                uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
                XF_HIP(ctx, xf_launch_pbd(st, s->aux, block, slabs, begin, count, s->tidx, s->taux, s->tdmp, s->X, s->V,
                                          s->bp, s->clock));
            }
        if (k + 1 < substeps) {
            rc = node(0, false);
            if (rc) return rc;
        }
    }
    return XF_OK;
}

int Prepare(xf_scene* s, float dt, int* energy, bool* sim) {
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return Fail(s->ctx, XF_E_INVALID, "dt must be positive and finite");
    int rc = ResolveMode(s, energy, sim);
    if (rc) return rc;
    if (*energy == 20 || *energy == 22 || *energy == 30) {
        if (s->opts.launch_mode == XF_LAUNCH_PERSISTENT)
            return Fail(s->ctx, XF_E_UNSUPPORTED, "Pixar and Cube energies run with per-colour launches only");
        for (const Body& b : s->bodies)
            if (!b.tet_mu.empty() && PixarEnergy(b.settings.energy))
                return Fail(s->ctx, XF_E_UNSUPPORTED, "Pixar energies take settings.compliance; a body has per-tet mu");
    }
    if (*energy == kRingMode) {
        if (s->opts.launch_mode != XF_LAUNCH_PER_COLOUR)
            return Fail(s->ctx, XF_E_UNSUPPORTED, "vertex-ring energies run with per-colour launches only");
        for (const Body& b : s->bodies)
            if (!b.tet_mu.empty())
                return Fail(s->ctx, XF_E_UNSUPPORTED,
                            "vertex-ring energies take settings.compliance; a body has per-tet mu");
        rc = UploadRings(s);
        if (rc) return rc;
    }
    if (s->material_dirty) {
        rc = UploadMaterial(s);
        if (rc) return rc;
    }
    if (s->opts.launch_mode == XF_LAUNCH_PERSISTENT && (s->nkin || s->manip_on))
        return Fail(s->ctx, XF_E_UNSUPPORTED, "kinematic nodes and the manipulator need per-colour or block launches");
    if (s->opts.launch_mode == XF_LAUNCH_BLOCKS && s->opts.tet_lanes == 4 &&
        (!(*energy == 0 || *energy == 4 || *energy == 7) || !*sim || s->dproj))
        return Fail(s->ctx, XF_E_UNSUPPORTED,
                    "tet_lanes 4 takes the simultaneous MixedSel / YeohSkinFast projection without damping");
    if (s->params_dirty || s->params_dt != dt) {
        rc = UploadParams(s, dt);
        if (rc) return rc;
    }
    return XF_OK;
}

// Host tables of the sweep-block launches (uploaded to blk_rec, blk_node,
// tloc and blk_fin).
struct BlockTables {
    std::vector<uint2> tloc;
    std::vector<uint32_t> rec, node, fin;
};

// Sweep-block tables for the slot cap lanes_opt (0: automatic). Blocks of
// launch colour q in layout order (body, then block number: the order of
// their tets inside each colour q, j) are merged, consecutive ones, into
// workgroup slots while every phase of the slot holds at most lanes_opt
// tets and its nodes fit LDS; a slot's
// tets of phase j are then one contiguous range. Slots are cut into
// eighths and dealt round-robin, so that under round-robin dispatch over
// the 8 XCDs each XCD sweeps one contiguous run of blocks (neighbouring
// blocks share node lines).
int PackBlocks(xf_scene* s, uint32_t lanes_opt, BlockTables* bt) {
    xf_ctx* ctx = s->ctx;
    *bt = BlockTables();
    // a slot's nodes (32 B each in LDS) within the device's per-workgroup LDS
    const uint32_t kSlotNodes = (uint32_t)std::min<size_t>(s->lds_limit / 32, 65535);
    uint32_t nblk[4] = {0, 0, 0, 0};
    std::vector<std::vector<uint32_t>> lay(s->bodies.size());
    for (size_t bi = 0; bi < s->bodies.size(); bi++) {
        const xf::SweepBlocks& sb = s->sweep[bi];
        lay[bi].resize(sb.n_blocks);
        for (uint32_t k = 0; k < sb.n_blocks; k++) lay[bi][k] = nblk[sb.launch[k]]++;
    }
    s->n_blocks = nblk[0] + nblk[1] + nblk[2] + nblk[3];
    // per block: sorted node list; per schedule position: the block's layout index
    std::vector<std::vector<uint32_t>> bnodes[4];
    for (int q = 0; q < 4; q++) bnodes[q].resize(nblk[q]);
    std::vector<uint32_t> sp(s->T, 0);
    for (size_t bi = 0; bi < s->bodies.size(); bi++) {
        const Body& b = s->bodies[bi];
        const xf::SweepBlocks& sb = s->sweep[bi];
        for (uint32_t t = 0; t < b.t; t++) {
            uint32_t k = sb.block[t], q = sb.launch[k], m = lay[bi][k];
            for (int i = 0; i < 4; i++) bnodes[q][m].push_back(b.node_base + b.tets[4 * (size_t)t + i]);
            sp[b.sched_pos[t]] = m;
        }
    }
    for (int q = 0; q < 4; q++)
        for (std::vector<uint32_t>& v : bnodes[q]) {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            if (v.size() > kSlotNodes)
                return Fail(ctx, XF_E_UNSUPPORTED, "a sweep block has %zu nodes (LDS holds %u)", v.size(),
                            kSlotNodes);
        }
    // per block and phase: first schedule position and count (walk of each colour)
    std::vector<uint32_t> bst[4], bct[4];
    for (int q = 0; q < 4; q++) {
        uint32_t nph = nblk[q] ? s->blk[q].nph : 0;
        bst[q].assign((size_t)nblk[q] * nph, 0);
        bct[q].assign((size_t)nblk[q] * nph, 0);
        for (uint32_t j = 0; j < nph; j++) {
            uint32_t c = s->blk_base[q] + j, p = s->colour_begin[c], end = s->colour_begin[c + 1];
            for (uint32_t m = 0; m < nblk[q]; m++) {
                uint32_t p0 = p;
                while (p < end && sp[p] == m) p++;
                bst[q][(size_t)m * nph + j] = p0;
                bct[q][(size_t)m * nph + j] = p - p0;
            }
            if (p != end) return Fail(ctx, XF_E_INVALID, "internal: sweep-block table of colour %u", c);
        }
    }
    // slots: runs of consecutive blocks, each phase at most `lanes_cap` tets
    std::vector<uint32_t> gfirst[4];  // per slot (in layout order): first block; sentinel at the end
    auto group = [&](uint32_t lanes_cap) {
        for (int q = 0; q < 4; q++) {
            gfirst[q].clear();
            uint32_t nph = nblk[q] ? s->blk[q].nph : 0;
            std::vector<uint32_t> lanes(nph, 0);
            uint32_t nodes = 0;
            for (uint32_t m = 0; m < nblk[q]; m++) {
                bool fits = !gfirst[q].empty() && nodes + bnodes[q][m].size() <= kSlotNodes;
                for (uint32_t j = 0; j < nph && fits; j++)
                    fits = lanes[j] + bct[q][(size_t)m * nph + j] <= lanes_cap;
                if (!fits) {
                    gfirst[q].push_back(m);
                    std::fill(lanes.begin(), lanes.end(), 0u);
                    nodes = 0;
                }
                for (uint32_t j = 0; j < nph; j++) lanes[j] += bct[q][(size_t)m * nph + j];
                nodes += (uint32_t)bnodes[q][m].size();
            }
            gfirst[q].push_back(nblk[q]);
        }
    };
    int n_cu = 256;
    XF_HIP(ctx, hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    uint32_t lanes_cap = lanes_opt;
    if (lanes_cap == 0) {
        // the largest cap that leaves >= 4 slots per compute unit in every launch colour
        for (uint32_t cap : {256u, 128u, 64u}) {
            group(cap);
            lanes_cap = cap;
            bool enough = true;
            for (int q = 0; q < 4; q++)
                if (nblk[q] && gfirst[q].size() - 1 < 4 * (size_t)n_cu) enough = false;
            if (enough) break;
        }
    } else {
        group(lanes_cap);
    }
    s->blk_lanes = lanes_cap;
    uint32_t nslot[4], slot_base[5] = {0, 0, 0, 0, 0};
    std::vector<uint32_t> deal[4];  // per launch colour: slot in layout order -> dispatched slot
    for (int q = 0; q < 4; q++) {
        uint32_t n = nslot[q] = (uint32_t)gfirst[q].size() - 1, next = 0;
        slot_base[q + 1] = slot_base[q] + n;
        deal[q].assign(n, 0);
        for (uint32_t m = 0; next < n; m++)
            for (uint32_t r = 0; r < 8; r++) {
                uint32_t lo = (uint32_t)(((uint64_t)n * r) / 8), hi = (uint32_t)(((uint64_t)n * (r + 1)) / 8);
                if (lo + m < hi) deal[q][lo + m] = next++;
            }
    }
    // slot node lists (member blocks' nodes are disjoint), records, local indices
    std::vector<std::vector<uint32_t>> snodes(slot_base[4]);
    std::vector<uint32_t> slot_of_block[4];
    s->blk_nmax = s->blk_nodes = s->blk_pmax = 0;
    for (int q = 0; q < 4; q++) {
        slot_of_block[q].assign(nblk[q], 0);
        for (uint32_t g = 0; g < nslot[q]; g++) {
            std::vector<uint32_t>& v = snodes[slot_base[q] + deal[q][g]];
            for (uint32_t m = gfirst[q][g]; m < gfirst[q][g + 1]; m++) {
                slot_of_block[q][m] = deal[q][g];
                v.insert(v.end(), bnodes[q][m].begin(), bnodes[q][m].end());
            }
            std::sort(v.begin(), v.end());
            s->blk_nmax = std::max(s->blk_nmax, (uint32_t)v.size());
            s->blk_nodes += (uint32_t)v.size();
        }
    }
    bt->tloc.assign(s->T, make_uint2(0u, 0u));
    for (size_t bi = 0; bi < s->bodies.size(); bi++) {
        const Body& b = s->bodies[bi];
        const xf::SweepBlocks& sb = s->sweep[bi];
        for (uint32_t t = 0; t < b.t; t++) {
            uint32_t k = sb.block[t], q = sb.launch[k];
            const std::vector<uint32_t>& v = snodes[slot_base[q] + slot_of_block[q][lay[bi][k]]];
            uint32_t l[4];
            for (int i = 0; i < 4; i++)
                l[i] = (uint32_t)(std::lower_bound(v.begin(), v.end(), b.node_base + b.tets[4 * (size_t)t + i]) -
                                  v.begin());
            bt->tloc[b.sched_pos[t]] = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
        }
    }
    bt->node.assign((size_t)slot_base[4] * s->blk_nmax, xf::kNoNode);
    for (uint32_t g = 0; g < slot_base[4]; g++)
        std::copy(snodes[g].begin(), snodes[g].end(), bt->node.begin() + (size_t)g * s->blk_nmax);
    // fused node phase: per slot of the last launch colour, its node row
    // followed by its share of the nodes outside that colour's slots (in
    // ascending order, dealt in equal runs)
    int launches = 0;
    s->blk_last = -1;
    for (int q = 0; q < 4; q++)
        if (nslot[q]) {
            launches++;
            s->blk_last = q;
        }
    s->blk_fmax = 0;
    bt->fin.clear();
    std::vector<uint32_t> extra;
    int cus = 256;
    XF_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    // fused only where the last launch is latency-bound (at most two
    // workgroups per compute unit): with more, the longer write-back and
    // the LDS of O cost more than the node launch they save
    if (launches < 2 || nslot[s->blk_last] > 2 * (uint32_t)cus) {
        s->blk_last = -1;
    } else {
        const int q = s->blk_last;
        std::vector<uint8_t> covered(s->N, 0);
        for (uint32_t g = slot_base[q]; g < slot_base[q + 1]; g++)
            for (uint32_t v : snodes[g]) covered[v] = 1;
        for (uint32_t v = 0; v < s->N; v++)
            if (!covered[v]) extra.push_back(v);
        s->blk_fmax = (uint32_t)((extra.size() + nslot[q] - 1) / nslot[q]);
        // staged with O and ids in LDS; else the node phase stays separate
        if (xf_block_fin_lds_bytes(s->blk_nmax, s->blk_fmax, 32) > s->lds_limit) {
            s->blk_last = -1;
            s->blk_fmax = 0;
        }
    }
    if (s->blk_last >= 0) {
        const int q = s->blk_last;
        const uint32_t fmax = s->blk_fmax;
        const size_t row = (size_t)s->blk_nmax + fmax;
        bt->fin.assign((size_t)nslot[q] * row, xf::kNoNode);
        for (uint32_t b = 0; b < nslot[q]; b++) {
            const uint32_t* bn = bt->node.data() + (size_t)(slot_base[q] + b) * s->blk_nmax;
            std::copy(bn, bn + s->blk_nmax, bt->fin.begin() + b * row);
            for (uint32_t j = 0; j < fmax && (size_t)b * fmax + j < extra.size(); j++)
                bt->fin[b * row + s->blk_nmax + j] = extra[(size_t)b * fmax + j];
        }
    }
    s->blk_wide = true;
    for (int q = 0; q < 4; q++)
        if (nslot[q] > 2 * (uint32_t)cus) s->blk_wide = false;
    for (int q = 0; q < 4; q++) {
        xf_scene::BlockLaunch& L = s->blk[q];
        L.nslot = nslot[q];
        if (L.nslot == 0) L.nph = 0;
        L.rec_off = (uint32_t)bt->rec.size();
        L.node_off = slot_base[q] * s->blk_nmax;
        bt->rec.resize(bt->rec.size() + 2 * (size_t)L.nslot * L.nph, 0u);
        for (uint32_t g = 0; g < L.nslot; g++)
            for (uint32_t j = 0; j < L.nph; j++) {
                uint32_t first = gfirst[q][g], count = 0;
                for (uint32_t m = first; m < gfirst[q][g + 1]; m++) count += bct[q][(size_t)m * L.nph + j];
                size_t at = L.rec_off + 2 * ((size_t)deal[q][g] * L.nph + j);
                bt->rec[at] = bst[q][(size_t)first * L.nph + j];
                bt->rec[at + 1] = count;
                s->blk_pmax = std::max(s->blk_pmax, count);
            }
    }
    return XF_OK;
}

// Uploads sweep-block tables, growing the device arrays when a packing needs
// more room than they hold. The stream is idle on return.
int UploadBlocks(xf_scene* s, const BlockTables& bt) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    XF_HIP(ctx, hipStreamSynchronize(st));
    auto fit = [&](uint32_t** p, size_t* have, size_t need) -> int {
        if (*p && *have >= need) return XF_OK;
        if (*p) {
            (void)hipFree(*p);
            s->device_bytes -= sizeof(uint32_t) * std::max<size_t>(*have, 1);
            *p = nullptr;
        }
        *have = need;
        return DevAlloc(ctx, p, need, &s->device_bytes);
    };
    int rc;
    if ((rc = fit(&s->blk_rec, &s->blk_rec_n, bt.rec.size())) || (rc = fit(&s->blk_node, &s->blk_node_n, bt.node.size())) ||
        (!bt.fin.empty() && (rc = fit(&s->blk_fin, &s->blk_fin_n, bt.fin.size()))))
        return rc;
    if (!s->tloc && (rc = DevAlloc(ctx, &s->tloc, s->T, &s->device_bytes))) return rc;
    if (!bt.rec.empty())
        XF_HIP(ctx, hipMemcpyAsync(s->blk_rec, bt.rec.data(), sizeof(uint32_t) * bt.rec.size(), hipMemcpyHostToDevice, st));
    if (!bt.node.empty())
        XF_HIP(ctx, hipMemcpyAsync(s->blk_node, bt.node.data(), sizeof(uint32_t) * bt.node.size(), hipMemcpyHostToDevice,
                                   st));
    XF_HIP(ctx, hipMemcpyAsync(s->tloc, bt.tloc.data(), sizeof(uint2) * s->T, hipMemcpyHostToDevice, st));
    if (!bt.fin.empty())
        XF_HIP(ctx, hipMemcpyAsync(s->blk_fin, bt.fin.data(), sizeof(uint32_t) * bt.fin.size(), hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipStreamSynchronize(st));
    return XF_OK;
}

int CheckSettings(xf_ctx* ctx, const xf_settings* st) {
    if (!EnergySupported(st->energy))
        return Fail(ctx, XF_E_INVALID, "energy %u out of range (the Settings field has 5 bits)", st->energy);
    if (!(st->damping >= 0.0f) || !std::isfinite(st->damping) || !(st->pbd_damping >= 0.0f) ||
        !std::isfinite(st->pbd_damping) || !(st->pbd_damping_amortized >= 0.0f) ||
        !std::isfinite(st->pbd_damping_amortized))
        return Fail(ctx, XF_E_INVALID, "damping, pbd_damping and pbd_damping_amortized must be finite and >= 0");
    if (st->rayleigh_type > XF_RAYLEIGH_POST_AMORTIZED)
        return Fail(ctx, XF_E_INVALID, "rayleigh_type %u unknown (0..3)", st->rayleigh_type);
    if (st->volume_passes > kMaxVolumePasses)
        return Fail(ctx, XF_E_INVALID, "volume_passes %u above the limit %u", st->volume_passes, kMaxVolumePasses);
    if (!(st->compliance > 0.0f)) return Fail(ctx, XF_E_INVALID, "compliance must be positive");
    if (!(st->poissons_ratio > 0.0f && st->poissons_ratio <= 0.5f))
        return Fail(ctx, XF_E_INVALID, "poissons_ratio must be in (0, 0.5]");
    if (st->collide & ~3u) return Fail(ctx, XF_E_INVALID, "unknown collide bits");
    if ((st->collide & XF_COLLIDE_SPHERE) && !(st->sphere_r > 0.0f))
        return Fail(ctx, XF_E_INVALID, "sphere radius must be positive");
    if (st->unlock_pins > 1) return Fail(ctx, XF_E_INVALID, "unlock_pins must be 0 or 1");
    return XF_OK;
}

}  // namespace

extern "C" {

XF_API const char* xf_version(void) { return "xpbdfem-gfx950 0.1"; }

XF_API void xf_settings_default(xf_settings* s) {
    memset(s, 0, sizeof(*s));
    s->energy = XF_ENERGY_MIXED_SEL;
    s->simultaneous = 1;
    s->compliance = 0.01f;
    s->poissons_ratio = 0.45f;
    s->gravity[0] = 0.0f;
    s->gravity[1] = -9.81f;
    s->yeoh[0] = 0.1095f;
    s->yeoh[1] = 14.95f;
    s->yeoh[2] = 4.595f;
}

XF_API int xf_ctx_create(int device, xf_ctx** out) {
    if (!out) return XF_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return XF_E_HIP;
    if (device < 0 || device >= count) return XF_E_INVALID;
    xf_ctx* ctx = new xf_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
                                                  hipSuccess) {
        delete ctx;
        return XF_E_HIP;
    }
    if (ResolveScratch(device, &ctx->scratch) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return XF_E_HIP;
    }
    *out = ctx;
    return XF_OK;
}

XF_API void xf_ctx_destroy(xf_ctx* ctx) {
    if (!ctx) return;
    while (!ctx->scenes.empty()) xf_scene_destroy(ctx->scenes.back());
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

XF_API const char* xf_last_error(const xf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

XF_API int xf_scene_create(xf_ctx* ctx, xf_scene** out) {
    if (!ctx || !out) return XF_E_INVALID;
    xf_scene* s = new xf_scene();
    s->ctx = ctx;
    ctx->scenes.push_back(s);
    s->opts.colour_order = XF_COLOUR_NATURAL;
    s->opts.block_size = 64;
    s->opts.use_graph = 1;
    s->opts.launch_mode = XF_LAUNCH_PER_COLOUR;
    s->opts.persistent_blocks = 0;
    *out = s;
    return XF_OK;
}

XF_API void xf_scene_destroy(xf_scene* s) {
    if (!s) return;
    xf_ctx* ctx = s->ctx;
    ctx->scenes.erase(std::remove(ctx->scenes.begin(), ctx->scenes.end(), s), ctx->scenes.end());
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DropGraph(s);
    DropAdjGraph(s);
    void* ptrs[] = {s->X,  s->O,     s->V,     s->tidx,  s->tg[0],         s->tg[1],  s->tg[2],
                    s->tg[3], s->taux, s->bp, s->clock, s->counter, s->colour_begin_d, s->arrive,
                    s->barrier_err, s->tdmp, s->rmeta, s->rinfo, s->rnodes, s->rincb, s->rinc, s->rtet,
                    s->blk_rec, s->blk_node, s->tloc, s->blk_fin, s->tout, s->kin_idx, s->kin_pos, s->manip};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (xf_scene::AdjBuf* b : {&s->adj_ck_x, &s->adj_ck_v, &s->adj_tape, &s->adj_xs, &s->adj_o, &s->adj_xb,
                                &s->adj_vb, &s->adj_ob, &s->adj_gacc, &s->adj_glog, &s->adj_xloss, &s->adj_tgt,
                                &s->adj_lw, &s->adj_save_x, &s->adj_save_v, &s->adj_traw, &s->adj_eterm,
                                &s->adj_lout, &s->adj_gout})
        if (b->p) (void)hipFree(b->p);
    delete s;
}

XF_API int xf_scene_add_body(xf_scene* s, const xf_body_desc* d, uint32_t* body_id) {
    if (!s || !d) return XF_E_INVALID;
    xf_ctx* ctx = s->ctx;
    if (s->finalized) return Fail(ctx, XF_E_STATE, "scene already finalized");
    if ((!d->xyz && !d->xyz64) || !d->tets || d->n_nodes < 4 || d->n_tets < 1)
        return Fail(ctx, XF_E_INVALID, "body needs xyz (or xyz64), tets, >= 4 nodes and >= 1 tet");
    if (!(d->density > 0.0f)) return Fail(ctx, XF_E_INVALID, "density must be positive");
    if ((uint64_t)s->N + d->n_nodes >= (1ull << 31) || (uint64_t)s->T + d->n_tets >= (1ull << 31))
        return Fail(ctx, XF_E_INVALID, "scene too large");
    if (s->bodies.size() >= (1u << 30)) return Fail(ctx, XF_E_INVALID, "too many bodies");
    Body b;
    b.n = d->n_nodes;
    b.t = d->n_tets;
    b.density = d->density;
    if (d->xyz64) {
        b.xyz64.assign(d->xyz64, d->xyz64 + 3 * (size_t)b.n);
        b.xyz.resize(3 * (size_t)b.n);
        for (size_t i = 0; i < b.xyz.size(); i++) {
            if (!std::isfinite(b.xyz64[i])) return Fail(ctx, XF_E_INVALID, "non-finite node position");
            b.xyz[i] = (float)b.xyz64[i];  // ordering only (slab schedule)
        }
    } else {
        b.xyz.assign(d->xyz, d->xyz + 3 * (size_t)b.n);
    }
    b.tets.assign(d->tets, d->tets + 4 * (size_t)b.t);
    for (uint32_t v : b.tets)
        if (v >= b.n) return Fail(ctx, XF_E_INVALID, "tet index %u out of range (n_nodes %u)", v, b.n);
    for (uint32_t t = 0; t < b.t; t++) {
        const uint32_t* is = &b.tets[4 * (size_t)t];
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++)
                if (is[i] == is[j]) return Fail(ctx, XF_E_INVALID, "tet %u repeats node %u", t, is[i]);
    }
    for (float v : b.xyz)
        if (!std::isfinite(v)) return Fail(ctx, XF_E_INVALID, "non-finite node position");
    b.vel.assign(3 * (size_t)b.n, 0.0f);
    if (d->velocity) b.vel.assign(d->velocity, d->velocity + 3 * (size_t)b.n);
    b.pinned.assign(b.n, 0);
    if (d->pinned)
        for (uint32_t i = 0; i < b.n; i++) b.pinned[i] = d->pinned[i] ? 1 : 0;
    if (d->pin_first_substep > 1) return Fail(ctx, XF_E_INVALID, "pin_first_substep must be 0 or 1");
    b.pin_first_substep = d->pin_first_substep != 0;
    if (d->tet_mu) {
        b.tet_mu.assign(d->tet_mu, d->tet_mu + b.t);
        for (float m : b.tet_mu)
            if (!(m > 0.0f) || !std::isfinite(m)) return Fail(ctx, XF_E_INVALID, "tet_mu must be positive");
    }
    if (d->tet_colour) {
        b.user_colour.assign(d->tet_colour, d->tet_colour + b.t);
        // each colour class must be node-disjoint
        std::vector<uint32_t> seen(b.n, ~0u);
        std::vector<uint32_t> order(b.t);
        for (uint32_t t = 0; t < b.t; t++) {
            if (b.user_colour[t] >= xf::kMaxColours)
                return Fail(ctx, XF_E_INVALID, "tet %u: colour %u >= %u", t, b.user_colour[t], xf::kMaxColours);
            order[t] = t;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t c) { return b.user_colour[a] < b.user_colour[c]; });
        for (uint32_t t : order)
            for (int i = 0; i < 4; i++) {
                uint32_t v = b.tets[4 * (size_t)t + i];
                if (seen[v] == b.user_colour[t])
                    return Fail(ctx, XF_E_INVALID, "colour %u has two tets sharing node %u", b.user_colour[t], v);
                seen[v] = b.user_colour[t];
            }
    }
    xf_settings_default(&b.settings);
    if (!(b.xyz64.empty() ? xf::InitElements(b.xyz.data(), b.n, b.tets.data(), b.t, b.density, &b.ec)
                          : xf::InitElements(b.xyz64.data(), b.n, b.tets.data(), b.t, b.density, &b.ec)))
        return Fail(ctx, XF_E_INVALID, "bad tet indices");
    for (uint32_t t = 0; t < b.t; t++)
        if (!(b.ec.vol[t] > 0.0f)) return Fail(ctx, XF_E_INVALID, "tet %u has non-positive rest volume", t);
    for (uint32_t i = 0; i < b.n; i++) {
        if (!std::isfinite(b.ec.w[i])) return Fail(ctx, XF_E_INVALID, "node %u belongs to no tet", i);
        if (b.pinned[i] && !b.pin_first_substep) b.ec.w[i] = 0.0f;
    }
    b.node_base = s->N;
    s->N += b.n;
    s->T += b.t;
    s->bodies.push_back(std::move(b));
    if (body_id) *body_id = (uint32_t)s->bodies.size() - 1;
    return XF_OK;
}

XF_API int xf_scene_finalize(xf_scene* s, const xf_scene_opts* opts) {
    if (!s) return XF_E_INVALID;
    xf_ctx* ctx = s->ctx;
    if (s->finalized) return Fail(ctx, XF_E_STATE, "scene already finalized");
    if (s->bodies.empty()) return Fail(ctx, XF_E_STATE, "scene has no bodies");
    if (opts) s->opts = *opts;
    if (s->opts.block_size == 0) s->opts.block_size = 64;
    if (s->opts.block_size % 64 || s->opts.block_size > 256)
        return Fail(ctx, XF_E_INVALID, "block_size must be 64, 128, 192 or 256");
    if (s->opts.colour_order > XF_COLOUR_MINSTD) return Fail(ctx, XF_E_INVALID, "unknown colour order");
    if (s->opts.launch_mode > XF_LAUNCH_BLOCKS) return Fail(ctx, XF_E_INVALID, "unknown launch mode");
    if (s->opts.xcd_slabs > 1) return Fail(ctx, XF_E_INVALID, "xcd_slabs must be 0 or 1");
    if (s->opts.ring_lanes != 0 && s->opts.ring_lanes != 16 && s->opts.ring_lanes != 32)
        return Fail(ctx, XF_E_INVALID, "ring_lanes must be 0, 16 or 32");
    if (s->opts.sweep_blocks > 1) return Fail(ctx, XF_E_INVALID, "sweep_blocks must be 0 or 1");
    if (s->opts.separate_finish > 1) return Fail(ctx, XF_E_INVALID, "separate_finish must be 0 or 1");
    if (s->opts.tet_lanes == 0) s->opts.tet_lanes = 1;
    if (s->opts.tet_lanes != 1 && s->opts.tet_lanes != 4) return Fail(ctx, XF_E_INVALID, "tet_lanes must be 0, 1 or 4");
    if (s->opts.block_lanes > 256) return Fail(ctx, XF_E_INVALID, "block_lanes must be at most 256");
    if (s->opts.contract > 1) return Fail(ctx, XF_E_INVALID, "contract must be 0 or 1");
    if (s->opts.sweep_blocks && s->opts.xcd_slabs)
        return Fail(ctx, XF_E_INVALID, "sweep_blocks and xcd_slabs are exclusive");
    if (s->opts.launch_mode == XF_LAUNCH_BLOCKS && !s->opts.sweep_blocks)
        return Fail(ctx, XF_E_INVALID, "XF_LAUNCH_BLOCKS needs sweep_blocks");
    if (s->opts.sweep_blocks)
        for (const Body& b : s->bodies)
            if (!b.user_colour.empty())
                return Fail(ctx, XF_E_INVALID, "sweep_blocks colours the tets itself; a body has tet_colour");
    XF_HIP(ctx, hipSetDevice(ctx->device));

    // Sweep blocks: colour (q, j) = phase j of the blocks of launch colour q,
    // numbered over the launch colours that hold blocks.
    std::vector<xf::SweepBlocks>& sweep = s->sweep;
    sweep.assign(s->opts.sweep_blocks ? s->bodies.size() : 0, xf::SweepBlocks());
    uint32_t* blk_base = s->blk_base;
    if (s->opts.sweep_blocks) {
        uint32_t nph[4] = {0, 0, 0, 0};
        for (size_t bi = 0; bi < s->bodies.size(); bi++) {
            Body& b = s->bodies[bi];
            std::vector<uint32_t> visit(b.t);
            if (s->opts.colour_order == XF_COLOUR_MINSTD) xf::MinstdOrder(b.t, visit.data());
            else
                for (uint32_t t = 0; t < b.t; t++) visit[t] = t;
            std::string err;
            if (!xf::BuildSweepBlocks(b.xyz.data(), b.n, b.tets.data(), b.t, visit.data(), &sweep[bi], &err))
                return Fail(ctx, XF_E_UNSUPPORTED, "sweep blocks of body %zu: %s", bi, err.c_str());
            const xf::SweepBlocks& sb = sweep[bi];
            for (uint32_t k = 0; k < sb.n_blocks; k++) nph[sb.launch[k]] = std::max(nph[sb.launch[k]], sb.phases[k]);
        }
        s->ncol = 0;
        for (int q = 0; q < 4; q++) {
            blk_base[q] = s->ncol;
            s->blk[q].nph = nph[q];
            s->ncol += nph[q];
        }
        if (s->ncol > xf::kMaxColours)
            return Fail(ctx, XF_E_UNSUPPORTED, "sweep blocks need %u colours (limit %u)", s->ncol, xf::kMaxColours);
        for (size_t bi = 0; bi < s->bodies.size(); bi++) {
            Body& b = s->bodies[bi];
            const xf::SweepBlocks& sb = sweep[bi];
            std::vector<uint32_t> visit(b.t), rank(b.t);
            if (s->opts.colour_order == XF_COLOUR_MINSTD) xf::MinstdOrder(b.t, visit.data());
            else
                for (uint32_t t = 0; t < b.t; t++) visit[t] = t;
            for (uint32_t k = 0; k < b.t; k++) rank[visit[k]] = k;
            b.colour.resize(b.t);
            b.order.resize(b.t);
            for (uint32_t t = 0; t < b.t; t++) {
                b.colour[t] = blk_base[sb.launch[sb.block[t]]] + sb.phase[t];
                b.order[t] = t;
            }
            b.ncol = 1 + *std::max_element(b.colour.begin(), b.colour.end());
            std::sort(b.order.begin(), b.order.end(), [&](uint32_t a, uint32_t c) {
                if (b.colour[a] != b.colour[c]) return b.colour[a] < b.colour[c];
                if (sb.block[a] != sb.block[c]) return sb.block[a] < sb.block[c];
                return rank[a] < rank[c];
            });
        }
    }

    // Colour each body; the scene runs max(ncol) phases.
    if (!s->opts.sweep_blocks) s->ncol = 0;
    for (Body& b : s->bodies) {
        if (s->opts.sweep_blocks) break;
        if (!b.user_colour.empty()) {
            b.colour = b.user_colour;
            b.ncol = 1 + *std::max_element(b.colour.begin(), b.colour.end());
            b.order.resize(b.t);
            for (uint32_t t = 0; t < b.t; t++) b.order[t] = t;
            std::stable_sort(b.order.begin(), b.order.end(),
                             [&](uint32_t a, uint32_t c) { return b.colour[a] < b.colour[c]; });
            s->ncol = std::max(s->ncol, b.ncol);
            continue;
        }
        b.ncol = xf::ColourSweepOrder(b.n, b.tets.data(), b.t, s->opts.colour_order == XF_COLOUR_MINSTD, &b.colour,
                                      &b.order);
        if (b.ncol == 0) return Fail(ctx, XF_E_UNSUPPORTED, "mesh needs more than %u colours", xf::kMaxColours);
        s->ncol = std::max(s->ncol, b.ncol);
    }
    // XCD slabs: eighths of each colour are spatial slabs (by centroid x),
    // node-index ordered inside.
    if (s->opts.xcd_slabs)
        for (Body& b : s->bodies) {
            std::vector<float> cx(b.t);
            std::vector<uint32_t> first(b.t);
            for (uint32_t t = 0; t < b.t; t++) {
                const uint32_t* is = &b.tets[4 * (size_t)t];
                cx[t] = b.xyz[3 * is[0]] + b.xyz[3 * is[1]] + b.xyz[3 * is[2]] + b.xyz[3 * is[3]];
                first[t] = std::min(std::min(is[0], is[1]), std::min(is[2], is[3]));
            }
            SlabOrder(&b.order, b.colour, cx, first);
        }
    // Schedule: colour-major, body-major within a colour, each body's tets
    // of that colour in its sweep order.
    s->colour_begin.assign(s->ncol + 1, 0);
    std::vector<uint4> h_tidx(s->T);
    std::vector<float4> h_tg0(s->T), h_tg1(s->T), h_tg2(s->T);
    s->h_tg3.assign(s->T, make_float4(0, 0, 0, 0));
    s->h_taux.assign(s->T, make_float4(0, 0, 0, 0));
    std::vector<float2> h_tdmp(s->T);
    std::vector<uint32_t> cursor(s->bodies.size(), 0);
    uint32_t pos = 0;
    for (uint32_t c = 0; c < s->ncol; c++) {
        s->colour_begin[c] = pos;
        for (size_t bi = 0; bi < s->bodies.size(); bi++) {
            Body& b = s->bodies[bi];
            if (b.sched_pos.empty()) b.sched_pos.assign(b.t, 0);
            uint32_t& k = cursor[bi];
            while (k < b.t && b.colour[b.order[k]] == c) {
                uint32_t t = b.order[k++];
                const uint32_t* is = &b.tets[4 * (size_t)t];
                const float* q = &b.ec.qi[9 * (size_t)t];
                const float* qq = &b.ec.qq[3 * (size_t)t];
                const float* qr = &b.ec.qr[3 * (size_t)t];
                h_tidx[pos] = make_uint4(b.node_base + is[0], b.node_base + is[1], b.node_base + is[2],
                                         b.node_base + is[3]);
                h_tg0[pos] = make_float4(q[0], q[1], q[2], q[3]);
                h_tg1[pos] = make_float4(q[4], q[5], q[6], q[7]);
                h_tg2[pos] = make_float4(q[8], qq[0], qq[1], qq[2]);
                s->h_tg3[pos] = make_float4(qr[0], qr[1], qr[2], 0.0f);
                uint32_t bid = (uint32_t)bi;
                float bidf;
                memcpy(&bidf, &bid, 4);
                s->h_taux[pos] = make_float4(0.0f, 1.0f, bidf, 0.0f);
                uint32_t local = k - 1;  // position in the body's sweep order
                float localf;
                memcpy(&localf, &local, 4);
                h_tdmp[pos] = make_float2(b.ec.area[t], localf);
                b.sched_pos[t] = pos;
                ++pos;
            }
        }
    }
    s->colour_begin[s->ncol] = pos;
    if (pos != s->T) return Fail(ctx, XF_E_INVALID, "internal: schedule covers %u of %u tets", pos, s->T);

    BlockTables bt;
    {
        int lds = 65536;
        XF_HIP(ctx, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        s->lds_limit = (size_t)lds;
    }
    if (s->opts.sweep_blocks) {
        int prc = PackBlocks(s, s->opts.block_lanes, &bt);
        if (prc) return prc;
    }

    // Nodes.
    std::vector<double4> hX(s->N), hO(s->N), hV(s->N);
    s->h_w.assign(s->N, 0.0);
    s->h_oflags.assign(s->N, 0.0);
    for (size_t bi = 0; bi < s->bodies.size(); bi++) {
        const Body& b = s->bodies[bi];
        for (uint32_t i = 0; i < b.n; i++) {
            uint32_t g = b.node_base + i;
            double flags = (double)(((uint32_t)bi << 1) | (b.pinned[i] ? 1u : 0u));
            s->h_w[g] = (double)b.ec.w[i];
            s->h_oflags[g] = flags;
            // Node data enters as float and is held in double (Geo.cpp:713),
            // or as the caller's FP64 positions.
            double px, py, pz;
            if (b.xyz64.empty()) {
                px = b.xyz[3 * i];
                py = b.xyz[3 * i + 1];
                pz = b.xyz[3 * i + 2];
            } else {
                px = b.xyz64[3 * i];
                py = b.xyz64[3 * i + 1];
                pz = b.xyz64[3 * i + 2];
            }
            hX[g] = make_double4(px, py, pz, s->h_w[g]);
            hO[g] = make_double4(px, py, pz, flags);
            hV[g] = make_double4(b.vel[3 * i], b.vel[3 * i + 1], b.vel[3 * i + 2], 0.0);
        }
    }
    int rc;
    if ((rc = DevAlloc(ctx, &s->X, s->N, &s->device_bytes)) || (rc = DevAlloc(ctx, &s->O, s->N, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->V, s->N, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->tidx, s->T, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->bp, s->bodies.size(), &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->clock, kClockWords, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->counter, 1, &s->device_bytes)) || (rc = DevAlloc(ctx, &s->tdmp, s->T, &s->device_bytes)))
        return rc;
    for (int k = 0; k < 4; k++)
        if ((rc = DevAlloc(ctx, &s->tg[k], s->T, &s->device_bytes))) return rc;
    if ((rc = DevAlloc(ctx, &s->colour_begin_d, s->ncol + 1, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->arrive, 1, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->barrier_err, 1, &s->device_bytes)) ||
        (rc = DevAlloc(ctx, &s->tout, s->T, &s->device_bytes)))
        return rc;
    XF_HIP(ctx, hipDeviceGetAttribute(&s->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    hipStream_t st = ctx->stream;
    XF_HIP(ctx, hipMemcpyAsync(s->X, hX.data(), sizeof(double4) * s->N, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->O, hO.data(), sizeof(double4) * s->N, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->V, hV.data(), sizeof(double4) * s->N, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->tidx, h_tidx.data(), sizeof(uint4) * s->T, hipMemcpyHostToDevice, st));
    s->h_bp.resize(s->bodies.size());
    std::vector<uint32_t> h_tout(s->T);
    {
        uint32_t off = 0;
        for (const Body& b : s->bodies) {
            for (uint32_t t = 0; t < b.t; t++) h_tout[b.sched_pos[t]] = off + t;
            off += b.t;
        }
    }
    XF_HIP(ctx, hipMemcpyAsync(s->tout, h_tout.data(), sizeof(uint32_t) * s->T, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->tg[0], h_tg0.data(), sizeof(float4) * s->T, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->tg[1], h_tg1.data(), sizeof(float4) * s->T, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->tg[2], h_tg2.data(), sizeof(float4) * s->T, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->tdmp, h_tdmp.data(), sizeof(float2) * s->T, hipMemcpyHostToDevice, st));
    if (s->opts.sweep_blocks && (rc = UploadBlocks(s, bt))) return rc;
    uint32_t uzero = 0;
    static const uint32_t kClockZero[kClockWords] = {0u, 0u};  // time 0.0f, tick 0
    XF_HIP(ctx, hipMemcpyAsync(s->clock, kClockZero, sizeof(kClockZero), hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->colour_begin_d, s->colour_begin.data(), sizeof(uint32_t) * (s->ncol + 1),
                               hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->barrier_err, &uzero, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    XF_HIP(ctx, hipStreamSynchronize(st));
    s->time_host = 0.0f;
    s->tick_host = 0;
    s->finalized = true;
    s->material_dirty = true;
    s->params_dirty = true;
    return XF_OK;
}

XF_API int xf_scene_info_get(const xf_scene* s, xf_scene_info* info) {
    if (!s || !info) return XF_E_INVALID;
    info->n_bodies = (uint32_t)s->bodies.size();
    info->n_nodes = s->N;
    info->n_tets = s->T;
    info->n_colours = s->ncol;
    info->device_bytes = s->device_bytes;
    info->n_rings = s->rings_host ? s->R : 0;
    info->n_ring_colours = s->rings_host ? s->rcol : 0;
    info->n_blocks = s->n_blocks;
    info->block_launches = 0;
    for (int q = 0; q < 4; q++) info->block_launches += s->blk[q].nslot ? 1u : 0u;
    info->n_block_nodes = s->blk_nodes;
    info->block_lanes = s->blk_lanes;
    uint32_t slots = 0;
    for (int q = 0; q < 4; q++) slots += s->blk[q].nslot;
    info->n_block_slots = slots;
    bool fused = s->blk_last >= 0 && !s->opts.separate_finish && s->opts.launch_mode == XF_LAUNCH_BLOCKS;
    for (const Body& b : s->bodies)
        if (b.settings.volume_passes != 0 || b.settings.damping > 0.0f || b.settings.pbd_damping > 0.0f) fused = false;
    if (s->nkin || s->manip_on) fused = false;
    info->fused_finish = fused ? 1u : 0u;
    info->tet_lanes = s->opts.tet_lanes;
    info->block_lds_limit = (uint32_t)s->lds_limit;
    info->block_node_stride = s->blk_nstride;
    info->contract = s->opts.contract && s->opts.launch_mode == XF_LAUNCH_BLOCKS ? 1u : 0u;
    return XF_OK;
}

XF_API int xf_scene_set_settings(xf_scene* s, uint32_t body, const xf_settings* st) {
    if (!s || !st) return XF_E_INVALID;
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    int rc = CheckSettings(s->ctx, st);
    if (rc) return rc;
    s->bodies[body].settings = *st;
    s->material_dirty = true;
    s->params_dirty = true;
    return XF_OK;
}

XF_API int xf_scene_set_tet_mu(xf_scene* s, uint32_t body, const float* tet_mu) {
    if (!s || !tet_mu) return XF_E_INVALID;
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    Body& b = s->bodies[body];
    for (uint32_t t = 0; t < b.t; t++)
        if (!(tet_mu[t] > 0.0f) || !std::isfinite(tet_mu[t]))
            return Fail(s->ctx, XF_E_INVALID, "tet_mu must be positive and finite");
    b.tet_mu.assign(tet_mu, tet_mu + b.t);
    s->material_dirty = true;
    return XF_OK;
}

XF_API int xf_scene_step(xf_scene* s, uint32_t substeps, float dt) {
    if (!s) return XF_E_INVALID;
    if (substeps == 0) return XF_OK;
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    int energy;
    bool sim;
    int rc = Prepare(s, dt, &energy, &sim);
    if (rc) return rc;
    if (!s->opts.use_graph || s->opts.launch_mode == XF_LAUNCH_PERSISTENT) {
        rc = Enqueue(s, substeps, dt, energy, sim, nullptr);
    } else {
        if (!s->graph_exec || s->graph_substeps != substeps || s->graph_dt != dt || s->graph_energy != energy ||
            s->graph_sim != sim || s->graph_aux != s->aux || s->graph_vpasses != s->vpasses ||
            s->graph_dproj != s->dproj || s->graph_post != s->post || s->graph_nkin != s->nkin ||
            s->graph_manip != s->manip_on) {
            DropGraph(s);
            XF_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            rc = Enqueue(s, substeps, dt, energy, sim, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return Fail(ctx, XF_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            s->graph = g;
            XF_HIP(ctx, hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
            s->graph_substeps = substeps;
            s->graph_dt = dt;
            s->graph_energy = energy;
            s->graph_sim = sim;
            s->graph_aux = s->aux;
            s->graph_vpasses = s->vpasses;
            s->graph_dproj = s->dproj;
            s->graph_post = s->post;
            s->graph_nkin = s->nkin;
            s->graph_manip = s->manip_on;
        }
        XF_HIP(ctx, hipGraphLaunch(s->graph_exec, ctx->stream));
    }
    if (rc) return rc;
    for (uint32_t k = 0; k < substeps; k++) s->time_host = s->time_host + dt;
    s->tick_host += substeps;
    return XF_OK;
}

namespace {
// Reports a grid-barrier timeout of an earlier persistent launch (the flag
// stays set, so the scene keeps failing loudly).
int CheckBarrier(xf_scene* s) {
    if (!s->barrier_err) return XF_OK;
    uint32_t e = 0;
    XF_HIP(s->ctx, hipMemcpyAsync(&e, s->barrier_err, sizeof(uint32_t), hipMemcpyDeviceToHost, s->ctx->stream));
    XF_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    if (e) return Fail(s->ctx, XF_E_HIP, "persistent launch: grid barrier timed out (grid not co-resident?)");
    return XF_OK;
}
}  // namespace

XF_API int xf_scene_sync(xf_scene* s) {
    if (!s) return XF_E_INVALID;
    XF_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return s->finalized ? CheckBarrier(s) : XF_OK;
}

XF_API int xf_scene_get_state(xf_scene* s, uint32_t body, double* x, double* v) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    const Body& b = s->bodies[body];
    std::vector<double4> tmp(b.n);
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    int brc = CheckBarrier(s);
    if (brc) return brc;
    if (x) {
        XF_HIP(ctx, hipMemcpyAsync(tmp.data(), s->X + b.node_base, sizeof(double4) * b.n, hipMemcpyDeviceToHost,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < b.n; i++) {
            x[3 * i] = tmp[i].x;
            x[3 * i + 1] = tmp[i].y;
            x[3 * i + 2] = tmp[i].z;
        }
    }
    if (v) {
        XF_HIP(ctx, hipMemcpyAsync(tmp.data(), s->V + b.node_base, sizeof(double4) * b.n, hipMemcpyDeviceToHost,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < b.n; i++) {
            v[3 * i] = tmp[i].x;
            v[3 * i + 1] = tmp[i].y;
            v[3 * i + 2] = tmp[i].z;
        }
    }
    return XF_OK;
}

XF_API int xf_scene_set_state(xf_scene* s, uint32_t body, const double* x, const double* v) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    const Body& b = s->bodies[body];
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double4> tmp(b.n);
    if (x) {
        // positions change; inverse masses keep their current values (pins and
        // kinematic nodes may have dropped theirs to 0)
        XF_HIP(ctx, hipMemcpyAsync(tmp.data(), s->X + b.node_base, sizeof(double4) * b.n, hipMemcpyDeviceToHost,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < b.n; i++) tmp[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], tmp[i].w);
        XF_HIP(ctx, hipMemcpyAsync(s->X + b.node_base, tmp.data(), sizeof(double4) * b.n, hipMemcpyHostToDevice,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < b.n; i++) tmp[i].w = s->h_oflags[b.node_base + i];
        XF_HIP(ctx, hipMemcpyAsync(s->O + b.node_base, tmp.data(), sizeof(double4) * b.n, hipMemcpyHostToDevice,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (v) {
        for (uint32_t i = 0; i < b.n; i++) tmp[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0);
        XF_HIP(ctx, hipMemcpyAsync(s->V + b.node_base, tmp.data(), sizeof(double4) * b.n, hipMemcpyHostToDevice,
                                   ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XF_OK;
}

XF_API int xf_scene_get_time(const xf_scene* s, float* t) {
    if (!s || !t) return XF_E_INVALID;
    *t = s->time_host;
    return XF_OK;
}

XF_API int xf_scene_set_time(xf_scene* s, float t) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    XF_HIP(ctx, hipMemcpyAsync(s->clock, &t, sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->time_host = t;
    return XF_OK;
}

XF_API int xf_scene_get_tick(const xf_scene* s, uint32_t* tick) {
    if (!s || !tick) return XF_E_INVALID;
    *tick = s->tick_host;
    return XF_OK;
}

XF_API int xf_scene_set_tick(xf_scene* s, uint32_t tick) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    XF_HIP(ctx, hipMemcpyAsync(s->clock + 1, &tick, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->tick_host = tick;
    return XF_OK;
}

XF_API int xf_scene_get_order(const xf_scene* s, uint32_t body, uint32_t* order, uint32_t* colour_of_tet) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    const Body& b = s->bodies[body];
    if (order) memcpy(order, b.order.data(), sizeof(uint32_t) * b.t);
    if (colour_of_tet) memcpy(colour_of_tet, b.colour.data(), sizeof(uint32_t) * b.t);
    return XF_OK;
}

XF_API int xf_scene_get_ring_order(xf_scene* s, uint32_t body, uint32_t* n_rings, uint32_t* order,
                                   uint32_t* colour_of_ring) {
    if (!s || !n_rings) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    int rc = BuildRingsHost(s);
    if (rc) return rc;
    const Body& b = s->bodies[body];
    *n_rings = b.rings.n;
    if (order) memcpy(order, b.ring_order.data(), sizeof(uint32_t) * b.rings.n);
    if (colour_of_ring) memcpy(colour_of_ring, b.ring_colour.data(), sizeof(uint32_t) * b.rings.n);
    return XF_OK;
}

XF_API int xf_scene_get_elements(const xf_scene* s, uint32_t body, float* qi9, float* qq3, float* qr3, float* vol,
                                 float* w) {
    if (!s) return XF_E_INVALID;
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    const Body& b = s->bodies[body];
    if (qi9) memcpy(qi9, b.ec.qi.data(), sizeof(float) * b.ec.qi.size());
    if (qq3) memcpy(qq3, b.ec.qq.data(), sizeof(float) * b.ec.qq.size());
    if (qr3) memcpy(qr3, b.ec.qr.data(), sizeof(float) * b.ec.qr.size());
    if (vol) memcpy(vol, b.ec.vol.data(), sizeof(float) * b.ec.vol.size());
    if (w) memcpy(w, b.ec.w.data(), sizeof(float) * b.ec.w.size());
    return XF_OK;
}

XF_API int xf_scene_profile(xf_scene* s, uint32_t substeps, float dt, float* ms_project, float* ms_node,
                            uint32_t* launches_project) {
    if (!s) return XF_E_INVALID;
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    int energy;
    bool sim;
    int rc = Prepare(s, dt, &energy, &sim);
    if (rc) return rc;
    LaunchTimer timer;
    rc = Enqueue(s, substeps, dt, energy, sim, &timer);
    hipError_t se = hipStreamSynchronize(ctx->stream);
    float tp = 0.0f, tn = 0.0f;
    auto sum = [&](std::vector<hipEvent_t>& ev, float* acc) {
        for (size_t k = 0; k + 1 < ev.size(); k += 2) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) *acc += ms;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    };
    uint32_t nproj = (uint32_t)(timer.ev_proj.size() / 2);
    sum(timer.ev_proj, &tp);
    sum(timer.ev_node, &tn);
    if (rc) return rc;
    if (se != hipSuccess) return Fail(ctx, XF_E_HIP, "profile sync: %s", hipGetErrorString(se));
    for (uint32_t k = 0; k < substeps; k++) s->time_host = s->time_host + dt;
    s->tick_host += substeps;
    if (ms_project) *ms_project = tp;
    if (ms_node) *ms_node = tn;
    if (launches_project) *launches_project = nproj;
    return XF_OK;
}

XF_API int xf_scene_time_steps(xf_scene* s, uint32_t calls, uint32_t substeps, float dt, float* ms) {
    if (!s || !ms) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XF_HIP(ctx, hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    int rc = XF_OK;
    if (he == hipSuccess) he = hipEventRecord(e0, ctx->stream);
    for (uint32_t k = 0; k < calls && he == hipSuccess && !rc; k++) rc = xf_scene_step(s, substeps, dt);
    if (he == hipSuccess && !rc) he = hipEventRecord(e1, ctx->stream);
    if (he == hipSuccess && !rc) he = hipEventSynchronize(e1);
    float t = 0.0f;
    if (he == hipSuccess && !rc) he = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (he != hipSuccess) return Fail(ctx, XF_E_HIP, "time_steps: %s", hipGetErrorString(he));
    *ms = t;
    return CheckBarrier(s);
}

XF_API int xf_scene_autotune(xf_scene* s, uint32_t substeps, float dt, uint32_t reps, xf_scene_opts* chosen,
                             float* best_ms) {
    if (!s || substeps == 0 || reps == 0) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    int energy;
    bool sim;
    int rc = Prepare(s, dt, &energy, &sim);
    if (rc) return rc;
    // Snapshot of the node state and clock, restored before every trial and
    // at the end: tuning does not advance the simulation.
    double4 *sx = nullptr, *so = nullptr, *sv = nullptr;
    float* sc = nullptr;
    uint64_t dummy = 0;
    if ((rc = DevAlloc(ctx, &sx, s->N, &dummy)) || (rc = DevAlloc(ctx, &so, s->N, &dummy)) ||
        (rc = DevAlloc(ctx, &sv, s->N, &dummy)) || (rc = DevAlloc(ctx, &sc, kClockWords, &dummy))) {
        for (void* p : {(void*)sx, (void*)so, (void*)sv, (void*)sc})
            if (p) (void)hipFree(p);
        return rc;
    }
    hipStream_t st = ctx->stream;
    auto copy_all = [&](bool save) -> hipError_t {
        hipError_t e;
        if ((e = hipMemcpyAsync(save ? sx : s->X, save ? s->X : sx, sizeof(double4) * s->N, hipMemcpyDeviceToDevice,
                                st)))
            return e;
        if ((e = hipMemcpyAsync(save ? so : s->O, save ? s->O : so, sizeof(double4) * s->N, hipMemcpyDeviceToDevice,
                                st)))
            return e;
        if ((e = hipMemcpyAsync(save ? sv : s->V, save ? s->V : sv, sizeof(double4) * s->N, hipMemcpyDeviceToDevice,
                                st)))
            return e;
        if ((e = hipMemcpyAsync(save ? sc : s->clock, save ? s->clock : sc, sizeof(float) * kClockWords,
                                hipMemcpyDeviceToDevice, st)))
            return e;
        return hipStreamSynchronize(st);
    };
    // Candidates: launch mode and workgroup size; for sweep blocks also
    // lanes per tet and the slot cap (block_lanes: the slots are repacked
    // for each cap, caps outermost).
    struct Cand {
        uint32_t mode, size, lanes, cap;
    };
    std::vector<Cand> candidates;
    for (uint32_t size : {64u, 128u, 256u}) candidates.push_back({XF_LAUNCH_PER_COLOUR, size, 1, 0});
    for (uint32_t size : {64u, 128u, 256u, 512u}) candidates.push_back({XF_LAUNCH_PERSISTENT, size, 1, 0});
    if (s->opts.sweep_blocks && energy != kRingMode)
        for (uint32_t cap : {32u, 64u, 128u, 256u})
            for (uint32_t lanes : {1u, 4u})
                for (uint32_t size : {64u, 128u, 256u}) candidates.push_back({XF_LAUNCH_BLOCKS, size, lanes, cap});
    xf_scene_opts keep = s->opts;
    const uint32_t cap_keep = s->blk_lanes;
    Cand best_c = {keep.launch_mode, keep.launch_mode == XF_LAUNCH_PERSISTENT ? keep.persistent_blocks : keep.block_size,
                   keep.tet_lanes, cap_keep};
    const bool quad_ok = (energy == 0 || energy == 4 || energy == 7) && sim && !s->dproj;
    int maxb = 0;  // ring, Pixar and Cube energies run per colour only
    if (energy != kRingMode && energy != 20 && energy != 22 && energy != 30)
        XF_HIP(ctx, xf_frame_max_blocks(energy, sim, s->aux, s->dproj, &maxb));
    float best = -1.0f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = copy_all(true);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    float time_keep = s->time_host;
    uint32_t tick_keep = s->tick_host;
    uint32_t packed = cap_keep;  // the slot cap the uploaded tables hold
    auto repack = [&](uint32_t cap) -> int {
        if (cap == packed) return XF_OK;
        DropGraph(s);
        BlockTables bt;
        int r = PackBlocks(s, cap, &bt);
        if (!r) r = UploadBlocks(s, bt);
        if (!r) packed = cap;
        return r;
    };
    for (const Cand& cand : candidates) {
        if (he != hipSuccess || rc) break;
        if (cand.mode == XF_LAUNCH_PERSISTENT && (int)cand.size > maxb) continue;
        if (cand.lanes == 4 && !quad_ok) continue;
        if (cand.mode == XF_LAUNCH_BLOCKS && (rc = repack(cand.cap))) break;
        s->opts.launch_mode = cand.mode;
        s->opts.tet_lanes = cand.lanes;
        if (cand.mode == XF_LAUNCH_PERSISTENT) s->opts.persistent_blocks = cand.size;
        else s->opts.block_size = cand.size;
        DropGraph(s);
        if ((he = copy_all(false)) != hipSuccess) break;
        if ((rc = xf_scene_step(s, substeps, dt))) break;  // warm-up and graph capture
        if ((he = hipEventRecord(e0, st)) != hipSuccess) break;
        for (uint32_t r = 0; r < reps && !rc; r++) rc = xf_scene_step(s, substeps, dt);
        if (rc) break;
        if ((he = hipEventRecord(e1, st)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess) break;
        float ms = 0.0f;
        if ((he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
        ms /= (float)reps;
        if (best < 0.0f || ms < best) {
            best = ms;
            best_c = cand;
            if (cand.mode != XF_LAUNCH_BLOCKS) best_c.cap = cap_keep;
        }
        if ((rc = CheckBarrier(s))) break;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    hipError_t he2 = copy_all(false);
    for (void* p : {(void*)sx, (void*)so, (void*)sv, (void*)sc}) (void)hipFree(p);
    s->time_host = time_keep;
    s->tick_host = tick_keep;
    s->opts = keep;
    s->opts.launch_mode = best_c.mode;
    s->opts.tet_lanes = best_c.lanes;
    if (best_c.mode == XF_LAUNCH_PERSISTENT) s->opts.persistent_blocks = best_c.size;
    else s->opts.block_size = best_c.size;
    DropGraph(s);
    if (!rc && s->opts.sweep_blocks) {
        rc = repack(best_c.cap);
        if (!rc && best_c.cap != cap_keep) s->opts.block_lanes = best_c.cap;
    }
    if (rc) return rc;
    if (he != hipSuccess) return Fail(ctx, XF_E_HIP, "autotune: %s", hipGetErrorString(he));
    if (he2 != hipSuccess) return Fail(ctx, XF_E_HIP, "autotune restore: %s", hipGetErrorString(he2));
    if (chosen) *chosen = s->opts;
    if (best_ms) *best_ms = best;
    return XF_OK;
}


namespace {
int Reserve(xf_scene* s, xf_scene::AdjBuf* b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b->bytes >= bytes) return XF_OK;
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
    hipError_t e = hipMalloc(&b->p, bytes);
    if (e != hipSuccess) return Fail(s->ctx, XF_E_NOMEM, "adjoint buffer of %zu bytes: %s", bytes, hipGetErrorString(e));
    b->bytes = bytes;
    return XF_OK;
}
}  // namespace

namespace {

// Device buffers of one adjoint call.
struct AdjBufs {
    double4 *ckx = nullptr, *ckv = nullptr;  // window-start checkpoints of X and V
    double* tape = nullptr;                  // 12 doubles per tet and substep of a window
    double4 *xs = nullptr, *oo = nullptr;    // X after the sweep and O, per substep of a window
    double4 *Xb = nullptr, *Vb = nullptr, *Ob = nullptr;
    void* gacc = nullptr;                    // per-tet gradient within a window (float; double when fp64)
    double* glog = nullptr;                  // per-tet gradient over windows
    double4 *xloss = nullptr, *tgt = nullptr;
    double* lw = nullptr;                    // per-node loss-gradient weight
    double4 *save_x = nullptr, *save_v = nullptr;
};

int AdjointCheck(xf_scene* s, uint32_t substeps, uint32_t n_loss, const uint32_t* loss_at, const double* targets,
                 uint32_t window) {
    xf_ctx* ctx = s->ctx;
    if (!s->finalized) return Fail(ctx, XF_E_STATE, "scene is not finalized");
    if (substeps == 0 || window == 0) return Fail(ctx, XF_E_INVALID, "substeps and window must be positive");
    if (n_loss == 0 || !loss_at || !targets) return Fail(ctx, XF_E_INVALID, "need at least one loss substep");
    for (uint32_t j = 0; j < n_loss; j++)
        if (loss_at[j] < 1 || loss_at[j] > substeps || (j && loss_at[j] <= loss_at[j - 1]))
            return Fail(ctx, XF_E_INVALID, "loss_at must ascend within [1, substeps]");
    if (s->opts.contract && s->opts.launch_mode == XF_LAUNCH_BLOCKS)
        return Fail(ctx, XF_E_UNSUPPORTED,
                    "contract rounds the forward sweep differently from the adjoint's exact replay; "
                    "turn contract off for the adjoint");
    for (const Body& b : s->bodies) {
        const xf_settings& st = b.settings;
        if (!st.simultaneous || (st.energy != 4 && st.energy != 7) || (st.collide & XF_COLLIDE_SPHERE) ||
            st.volume_passes != 0 || st.damping > 0.0f || st.pbd_damping > 0.0f)
            return Fail(ctx, XF_E_UNSUPPORTED,
                        "adjoint covers simultaneous MixedSel/YeohSkinFast with pins and ground contact "
                        "(no volume passes, no damping)");
        if (st.unlock_pins || b.pin_first_substep)
            return Fail(ctx, XF_E_UNSUPPORTED, "adjoint takes pins locked from the first substep");
    }
    if (s->nkin || s->manip_on)
        return Fail(ctx, XF_E_UNSUPPORTED, "adjoint takes no kinematic nodes or manipulator");
    return XF_OK;
}

// `substeps` substeps of the FP64 element model from (X, V): prediction,
// the colours in sweep order, pins / ground contact / velocity. Advances
// the clock as xf_scene_step does.
int Forward64(xf_scene* s, uint32_t substeps, float dt) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t N = s->N;
    for (uint32_t k = 0; k < substeps; k++) {
        XF_HIP(ctx, xf_launch_predict64(st, N, s->X, s->O, s->V, s->bp, dt, s->clock));
        for (uint32_t c = 0; c < s->ncol; c++) {
            uint32_t begin = s->colour_begin[c], count = s->colour_begin[c + 1] - begin;
            XF_HIP(ctx, xf_launch_project64(st, s->aux, false, begin, count, s->tidx, s->tg[0], s->tg[1], s->tg[2],
                                            s->tg[3], s->taux, s->X, nullptr, s->bp, s->comp1_u, s->a_u, dt));
        }
        XF_HIP(ctx, xf_launch_finish64(st, N, s->X, s->O, s->V, s->bp, nullptr, nullptr));
        s->time_host = s->time_host + dt;
        s->tick_host += 1;
    }
    return XF_OK;
}

// Forward rollout with checkpoints at window starts and X at the loss
// substeps; graph: through xf_scene_step (its hipGraph cache), else eager
// launches (no graph instantiation); fp64: the FP64 element model
// (Forward64). Leaves the end state saved in save_x / save_v and the end
// clock in *clock_end.
int AdjointForward(xf_scene* s, uint32_t substeps, float dt, int energy, bool sim, uint32_t n_loss,
                   const uint32_t* loss_at, uint32_t window, const AdjBufs& b, bool graph, bool fp64,
                   float* clock_end) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t N = s->N;
    const size_t n4 = sizeof(double4) * N;
    int rc;
    uint32_t k = 0, j = 0;
    while (k < substeps) {
        if (k % window == 0) {
            XF_HIP(ctx, hipMemcpyAsync(b.ckx + (size_t)N * (k / window), s->X, n4, hipMemcpyDeviceToDevice, st));
            XF_HIP(ctx, hipMemcpyAsync(b.ckv + (size_t)N * (k / window), s->V, n4, hipMemcpyDeviceToDevice, st));
        }
        uint32_t next = std::min(substeps, (k / window + 1) * window);
        if (j < n_loss) next = std::min(next, loss_at[j]);
        if (fp64) {
            if ((rc = Forward64(s, next - k, dt))) return rc;
        } else if (graph) {
            if ((rc = xf_scene_step(s, next - k, dt))) return rc;
        } else {
            if ((rc = Enqueue(s, next - k, dt, energy, sim, nullptr))) return rc;
            for (uint32_t q = k; q < next; q++) s->time_host = s->time_host + dt;
            s->tick_host += next - k;
        }
        k = next;
        if (j < n_loss && loss_at[j] == k) {
            XF_HIP(ctx, hipMemcpyAsync(b.xloss + (size_t)N * j, s->X, n4, hipMemcpyDeviceToDevice, st));
            j++;
        }
    }
    XF_HIP(ctx, hipMemcpyAsync(b.save_x, s->X, n4, hipMemcpyDeviceToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(b.save_v, s->V, n4, hipMemcpyDeviceToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(clock_end, s->clock, sizeof(float), hipMemcpyDeviceToHost, st));
    XF_HIP(ctx, hipStreamSynchronize(st));
    return CheckBarrier(s);
}

// Launches of the reverse pass, window by window from the end: each window
// is replayed from its checkpoint (recording the sweep tape and the node
// states before the contact pass), then swept back. pipelined: the replay
// of window w - 1 runs interleaved with the reverse of window w (their
// colour launches paired into one launch each, xf_launch_adj_pair), with
// the tape and node records double-buffered by window parity; else one
// window after the other in single buffers. glog receives -dL/dlog mu per
// schedule position.
int ReverseLaunches(xf_scene* s, uint32_t substeps, float dt, int energy, uint32_t n_loss, const uint32_t* loss_at,
                    uint32_t window, const AdjBufs& b, bool fp64, bool pipelined) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const uint32_t N = s->N, T = s->T;
    const uint32_t nwin = (substeps + window - 1) / window;
    const uint32_t wlen = std::min(window, substeps);
    auto k0_of = [&](uint32_t w) { return w * window; };
    auto len_of = [&](uint32_t w) { return std::min(substeps, (w + 1) * window) - w * window; };
    // per-window records: tape (12 doubles a tet), X after the sweep and O (a node each), per substep
    auto slot = [&](uint32_t w) { return pipelined ? (w & 1u) : 0u; };
    auto tape_of = [&](uint32_t w, uint32_t i) { return b.tape + 12 * (size_t)T * (slot(w) * (size_t)wlen + i); };
    auto xs_of = [&](uint32_t w, uint32_t i) { return b.xs + (size_t)N * (slot(w) * (size_t)wlen + i); };
    auto oo_of = [&](uint32_t w, uint32_t i) { return b.oo + (size_t)N * (slot(w) * (size_t)wlen + i); };
    const double4* W = b.ckx;  // inverse masses (.w) for the reverse sweeps, constant over the rollout
    XF_HIP(ctx, xf_launch_zero_d4(st, N, b.Xb));
    XF_HIP(ctx, xf_launch_zero_d4(st, N, b.Vb));
    XF_HIP(ctx, xf_launch_zero_u32(st, fp64 ? 2 * T : T, reinterpret_cast<uint32_t*>(b.gacc)));
    XF_HIP(ctx, xf_launch_zero_u32(st, 2 * T, reinterpret_cast<uint32_t*>(b.glog)));
    const bool single = s->bodies.size() == 1;
    // replay of window w: start (state from the checkpoint, first prediction)
    auto replay_start = [&](uint32_t w) -> int {
        XF_HIP(ctx, xf_launch_copy2_d4(st, N, s->X, b.ckx + (size_t)N * w, s->V, b.ckv + (size_t)N * w));
        if (fp64) XF_HIP(ctx, xf_launch_predict64(st, N, s->X, s->O, s->V, s->bp, dt, nullptr));
        else XF_HIP(ctx, xf_launch_predict(st, N, s->X, s->O, s->V, s->bp));
        return XF_OK;
    };
    // end of replayed substep i of window w (records X and O, then the next prediction)
    auto replay_end = [&](uint32_t w, uint32_t i) -> int {
        const bool next = i + 1 < len_of(w);
        if (fp64) {
            XF_HIP(ctx, xf_launch_finish64(st, N, s->X, s->O, s->V, s->bp, xs_of(w, i), oo_of(w, i)));
            if (next) XF_HIP(ctx, xf_launch_predict64(st, N, s->X, s->O, s->V, s->bp, dt, nullptr));
        } else {
            XF_HIP(ctx, xf_launch_finish_snap(st, next, single, N, s->X, s->O, s->V, s->bp, s->clock, xs_of(w, i),
                                              oo_of(w, i)));
        }
        return XF_OK;
    };
    // colour c of a replay substep (record) and colour v of a reverse substep, either absent (nullptr tape)
    auto colours = [&](double* rec_tape, uint32_t c, const double* rev_tape, uint32_t v) -> int {
        const uint32_t rb = s->colour_begin[c], rc = rec_tape ? s->colour_begin[c + 1] - rb : 0;
        const uint32_t vb = s->colour_begin[v], vc = rev_tape ? s->colour_begin[v + 1] - vb : 0;
        if (rc && vc) {
            XF_HIP(ctx, xf_launch_adj_pair(st, s->aux, fp64, rb, rc, s->X, rec_tape, vb, vc, rev_tape, W, b.Xb,
                                           b.gacc, s->tidx, s->tg[0], s->tg[1], s->tg[2], s->tg[3], s->taux, s->bp,
                                           s->comp1_u, s->a_u, dt));
        } else if (rc) {
            if (fp64)
                XF_HIP(ctx, xf_launch_project64(st, s->aux, true, rb, rc, s->tidx, s->tg[0], s->tg[1], s->tg[2],
                                                s->tg[3], s->taux, s->X, rec_tape, s->bp, s->comp1_u, s->a_u, dt));
            else
                XF_HIP(ctx, xf_launch_project_record(st, energy, s->aux, rb, rc, s->tidx, s->tg[0], s->tg[1],
                                                     s->tg[2], s->tg[3], s->taux, s->X, rec_tape, s->bp, s->comp1_u,
                                                     s->a_u, dt, nullptr));
        } else if (vc) {
            XF_HIP(ctx, xf_launch_project_reverse(st, s->aux, fp64, vb, vc, s->tidx, s->tg[0], s->tg[1], s->tg[2],
                                                  s->tg[3], s->taux, rev_tape, W, b.Xb, b.gacc, s->bp, s->comp1_u,
                                                  s->a_u, dt));
        }
        return XF_OK;
    };
    uint32_t jl = n_loss;
    // the loss injected at the end of substep kk, if kk ends a loss frame
    auto loss_of = [&](uint32_t kk, const double4** xl, const double4** tg) {
        *xl = nullptr;
        *tg = nullptr;
        if (jl > 0 && loss_at[jl - 1] == kk + 1) {
            jl--;
            *xl = b.xloss + (size_t)N * jl;
            *tg = b.tgt + (size_t)N * jl;
        }
    };
    int rc;
    if (pipelined) {
        // the last window's replay runs alone
        const uint32_t wl = nwin - 1;
        if ((rc = replay_start(wl))) return rc;
        for (uint32_t i = 0; i < len_of(wl); i++) {
            for (uint32_t c = 0; c < s->ncol; c++)
                if ((rc = colours(tape_of(wl, i), c, nullptr, 0))) return rc;
            if ((rc = replay_end(wl, i))) return rc;
        }
    }
    for (uint32_t w = nwin; w-- > 0;) {
        // the window to replay alongside this reverse (pipelined: the previous one), or this one first
        const bool rep = pipelined ? w > 0 : true;
        const uint32_t wr = pipelined ? w - 1 : w;
        if (!pipelined) {
            if ((rc = replay_start(w))) return rc;
            for (uint32_t i = 0; i < len_of(w); i++) {
                for (uint32_t c = 0; c < s->ncol; c++)
                    if ((rc = colours(tape_of(w, i), c, nullptr, 0))) return rc;
                if ((rc = replay_end(w, i))) return rc;
            }
        } else if (rep && (rc = replay_start(wr))) {
            return rc;
        }
        const uint32_t lv = len_of(w), lr = pipelined && rep ? len_of(wr) : 0, k0 = k0_of(w);
        const double4 *xl, *tg;
        loss_of(k0 + lv - 1, &xl, &tg);
        XF_HIP(ctx, xf_launch_adj_nodes(st, fp64, false, true, N, s->O, xs_of(w, lv - 1), oo_of(w, lv - 1), b.Xb,
                                        b.Vb, b.Ob, s->bp, xl, tg, b.lw));
        for (uint32_t i = 0; i < std::max(lv, lr); i++) {
            const bool von = i < lv, ron = i < lr;
            const uint32_t kk = k0 + lv - 1 - i;  // reverse substep (when von)
            for (uint32_t c = 0; c < s->ncol; c++)
                if ((rc = colours(ron ? tape_of(wr, i) : nullptr, c, von ? tape_of(w, lv - 1 - i) : nullptr,
                                  s->ncol - 1 - c)))
                    return rc;
            if (ron && (rc = replay_end(wr, i))) return rc;
            if (von) {
                const bool more = kk > k0;
                xl = tg = nullptr;
                if (more) loss_of(kk - 1, &xl, &tg);
                const uint32_t ip = more ? lv - 2 - i : 0;
                XF_HIP(ctx, xf_launch_adj_nodes(st, fp64, true, more, N, s->O, xs_of(w, ip), oo_of(w, ip), b.Xb,
                                                b.Vb, b.Ob, s->bp, xl, tg, b.lw));
            }
        }
        XF_HIP(ctx, xf_launch_adj_flush(st, fp64, T, b.gacc, b.glog));
    }
    return XF_OK;
}

// Reverse pass: the launches of ReverseLaunches, eagerly or (graph) as one
// hipGraph kept for the call shape and replayed by later calls of that
// shape. Leaves the scene at the end of the forward rollout and glog
// holding -dL/dlog mu per schedule position.
int AdjointReverse(xf_scene* s, uint32_t substeps, float dt, int energy, uint32_t n_loss, const uint32_t* loss_at,
                   uint32_t window, const AdjBufs& b, bool fp64, bool graph, bool pipelined,
                   const float* clock_end) {
    xf_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const size_t n4 = sizeof(double4) * s->N;
    if (!graph) {
        int rc = ReverseLaunches(s, substeps, dt, energy, n_loss, loss_at, window, b, fp64, pipelined);
        if (rc) return rc;
    } else {
        auto bits = [](float f) {
            uint32_t u;
            memcpy(&u, &f, 4);
            return (uint64_t)u;
        };
        std::vector<uint64_t> key = {substeps, window, n_loss, fp64, pipelined, (uint64_t)energy, s->aux, bits(dt),
                                     bits(s->comp1_u), bits(s->a_u), s->ncol, s->N, s->T};
        key.insert(key.end(), loss_at, loss_at + n_loss);
        for (const void* q : {(const void*)b.ckx, (const void*)b.ckv, (const void*)b.tape, (const void*)b.xs,
                              (const void*)b.oo, (const void*)b.Xb, (const void*)b.Vb, (const void*)b.Ob,
                              (const void*)b.gacc,
                              (const void*)b.glog, (const void*)b.xloss, (const void*)b.tgt, (const void*)b.lw,
                              (const void*)s->X, (const void*)s->O, (const void*)s->V, (const void*)s->bp,
                              (const void*)s->clock})
            key.push_back((uint64_t)(uintptr_t)q);
        if (!s->adj_exec || s->adj_key != key) {
            DropAdjGraph(s);
            XF_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = ReverseLaunches(s, substeps, dt, energy, n_loss, loss_at, window, b, fp64, pipelined);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return Fail(ctx, XF_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&s->adj_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->adj_exec = nullptr;
                return Fail(ctx, XF_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
            }
            s->adj_key = key;
        }
        XF_HIP(ctx, hipGraphLaunch(s->adj_exec, st));
    }
    // Leave the scene at the end of the forward rollout.
    XF_HIP(ctx, hipMemcpyAsync(s->X, b.save_x, n4, hipMemcpyDeviceToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->V, b.save_v, n4, hipMemcpyDeviceToDevice, st));
    XF_HIP(ctx, hipMemcpyAsync(s->clock, clock_end, sizeof(float), hipMemcpyHostToDevice, st));
    return XF_OK;
}

}  // namespace

namespace {
// xf_scene_adjoint (fp64 false) and xf_scene_adjoint_fp64 (fp64 true).
int AdjointRun(xf_scene* s, uint32_t substeps, float dt, uint32_t n_loss, const uint32_t* loss_at,
               const double* targets, uint32_t window, double* loss, double* grad_log_mu, double* body_loss,
               bool fp64) {
    if (!s) return XF_E_INVALID;
    xf_ctx* ctx = s->ctx;
    int rc = AdjointCheck(s, substeps, n_loss, loss_at, targets, window);
    if (rc) return rc;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    int energy;
    bool sim;
    rc = Prepare(s, dt, &energy, &sim);
    if (rc) return rc;
    const uint32_t N = s->N, T = s->T;
    const uint32_t nwin = (substeps + window - 1) / window;
    const uint32_t wlen = std::min(window, substeps);
    const size_t n4 = sizeof(double4) * N;
    // more than one window: the replay of each window runs beside the reverse of the next
    const bool pipelined = nwin > 1;
    const size_t nbuf = pipelined ? 2 : 1;
    if ((rc = Reserve(s, &s->adj_ck_x, n4 * nwin)) || (rc = Reserve(s, &s->adj_ck_v, n4 * nwin)) ||
        (rc = Reserve(s, &s->adj_tape, sizeof(double) * 12 * (size_t)T * wlen * nbuf)) ||
        (rc = Reserve(s, &s->adj_xs, n4 * wlen * nbuf)) || (rc = Reserve(s, &s->adj_o, n4 * wlen * nbuf)) ||
        (rc = Reserve(s, &s->adj_xb, n4)) || (rc = Reserve(s, &s->adj_vb, n4)) || (rc = Reserve(s, &s->adj_ob, n4)) ||
        (rc = Reserve(s, &s->adj_gacc, (fp64 ? sizeof(double) : sizeof(float)) * T)) ||
        (rc = Reserve(s, &s->adj_glog, sizeof(double) * T)) ||
        (rc = Reserve(s, &s->adj_xloss, n4 * n_loss)) || (rc = Reserve(s, &s->adj_tgt, n4 * n_loss)) ||
        (rc = Reserve(s, &s->adj_lw, sizeof(double) * N)) || (rc = Reserve(s, &s->adj_save_x, n4)) ||
        (rc = Reserve(s, &s->adj_save_v, n4)) ||
        (rc = Reserve(s, &s->adj_traw, sizeof(double) * 3 * (size_t)N * n_loss)) ||
        (rc = Reserve(s, &s->adj_eterm, sizeof(double) * s->bodies.size() * n_loss)) ||
        (rc = Reserve(s, &s->adj_lout, sizeof(double) * (1 + s->bodies.size()))) ||
        (rc = Reserve(s, &s->adj_gout, sizeof(double) * T)))
        return rc;
    AdjBufs b;
    b.ckx = (double4*)s->adj_ck_x.p;
    b.ckv = (double4*)s->adj_ck_v.p;
    b.tape = (double*)s->adj_tape.p;
    b.xs = (double4*)s->adj_xs.p;
    b.oo = (double4*)s->adj_o.p;
    b.Xb = (double4*)s->adj_xb.p;
    b.Vb = (double4*)s->adj_vb.p;
    b.Ob = (double4*)s->adj_ob.p;
    b.gacc = s->adj_gacc.p;
    b.glog = (double*)s->adj_glog.p;
    b.xloss = (double4*)s->adj_xloss.p;
    b.tgt = (double4*)s->adj_tgt.p;
    b.lw = (double*)s->adj_lw.p;
    b.save_x = (double4*)s->adj_save_x.p;
    b.save_v = (double4*)s->adj_save_v.p;
    hipStream_t st = ctx->stream;
    const uint32_t B = (uint32_t)s->bodies.size();
    double* traw = (double*)s->adj_traw.p;
    double* eterm = (double*)s->adj_eterm.p;
    double* lout = (double*)s->adj_lout.p;
    double* gout = (double*)s->adj_gout.p;

    // Targets (body-concatenated xyz) to the node layout and per-node
    // loss-gradient weights 2/n_b, on the device.
    XF_HIP(ctx, hipMemcpyAsync(traw, targets, sizeof(double) * 3 * (size_t)N * n_loss, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, xf_launch_expand_targets(st, N * n_loss, traw, b.tgt));
    XF_HIP(ctx, xf_launch_loss_weights(st, N, s->O, s->bp, b.lw));

    float clock_end = 0.0f;
    if ((rc = AdjointForward(s, substeps, dt, energy, sim, n_loss, loss_at, window, b, true, fp64, &clock_end)))
        return rc;
    // Loss: sum over loss substeps, then bodies, of (1/n_b) sum |x - target|^2.
    XF_HIP(ctx, xf_launch_loss(st, n_loss, B, N, b.xloss, b.tgt, s->bp, eterm, lout));
    if ((rc = AdjointReverse(s, substeps, dt, energy, n_loss, loss_at, window, b, fp64, s->opts.use_graph != 0,
                             pipelined, &clock_end)))
        return rc;
    XF_HIP(ctx, xf_launch_grad_out(st, T, b.glog, s->tout, gout));
    if (loss) XF_HIP(ctx, hipMemcpyAsync(loss, lout, sizeof(double), hipMemcpyDeviceToHost, st));
    if (body_loss) XF_HIP(ctx, hipMemcpyAsync(body_loss, lout + 1, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    if (grad_log_mu) XF_HIP(ctx, hipMemcpyAsync(grad_log_mu, gout, sizeof(double) * T, hipMemcpyDeviceToHost, st));
    XF_HIP(ctx, hipStreamSynchronize(st));
    return XF_OK;
}
}  // namespace

XF_API int xf_scene_adjoint(xf_scene* s, uint32_t substeps, float dt, uint32_t n_loss, const uint32_t* loss_at,
                            const double* targets, uint32_t window, double* loss, double* grad_log_mu,
                            double* body_loss) {
    return AdjointRun(s, substeps, dt, n_loss, loss_at, targets, window, loss, grad_log_mu, body_loss, false);
}

XF_API int xf_scene_adjoint_fp64(xf_scene* s, uint32_t substeps, float dt, uint32_t n_loss, const uint32_t* loss_at,
                                 const double* targets, uint32_t window, double* loss, double* grad_log_mu,
                                 double* body_loss) {
    return AdjointRun(s, substeps, dt, n_loss, loss_at, targets, window, loss, grad_log_mu, body_loss, true);
}

namespace {
// Scratch bytes xf_scene_adjoint_noalloc carves for one call (each region
// rounded up to 256 bytes).
size_t AdjointScratchBytes(const xf_scene* s, uint32_t substeps, uint32_t n_loss, uint32_t window) {
    const size_t N = s->N, T = s->T, B = s->bodies.size();
    const size_t nwin = (substeps + window - 1) / window, wlen = std::min(window, substeps);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t n4 = al(sizeof(double4) * N);
    return 2 * al(n4 * nwin) + al(sizeof(double) * 12 * T * wlen) + 2 * al(n4 * wlen) + 5 * n4 +
           al(sizeof(float) * T) + 2 * al(sizeof(double) * T) + 2 * al(n4 * n_loss) +
           al(sizeof(double) * 3 * N * n_loss) + al(sizeof(double) * N) + al(sizeof(double) * B * n_loss) +
           al(sizeof(double) * (1 + B));
}
}  // namespace

XF_API int xf_scene_adjoint_noalloc(xf_scene* s, uint32_t substeps, float dt, uint32_t n_loss,
                                    const uint32_t* loss_at, const double* targets, uint32_t window, double* loss,
                                    double* grad_log_mu, double* body_loss) {
    if (!s) return XF_E_INVALID;
    xf_ctx* ctx = s->ctx;
    int rc = AdjointCheck(s, substeps, n_loss, loss_at, targets, window);
    if (rc) return rc;
    const size_t need = AdjointScratchBytes(s, substeps, n_loss, window);
    if (need > kAdjointScratchBytes)
        return Fail(ctx, XF_E_INVALID,
                    "adjoint needs %zu bytes of scratch, more than the %zu of the allocation-free entry point "
                    "(use xf_scene_adjoint)",
                    need, kAdjointScratchBytes);
    XF_HIP(ctx, hipSetDevice(ctx->device));
    // Per-tet material storage is allocated on first use; this entry point
    // does not allocate it.
    if (s->material_dirty && !s->taux) {
        bool had = s->aux;
        ComputeMaterial(s);
        bool need_taux = s->aux || s->vpasses > 0;
        s->aux = had;
        if (need_taux)
            return Fail(ctx, XF_E_STATE,
                        "the scene's per-tet material storage is not allocated yet: run xf_scene_step or "
                        "xf_scene_adjoint once after changing settings");
    }
    int energy;
    bool sim;
    rc = Prepare(s, dt, &energy, &sim);
    if (rc) return rc;
    const uint32_t N = s->N, T = s->T, B = (uint32_t)s->bodies.size();
    const uint32_t nwin = (substeps + window - 1) / window;
    const uint32_t wlen = std::min(window, substeps);
    const size_t n4 = sizeof(double4) * N;
    // carve the scratch (each region 256-byte aligned)
    unsigned char* p = ctx->scratch;
    auto take = [&p](size_t bytes) {
        unsigned char* r = p;
        p += (bytes + 255) & ~size_t(255);
        return r;
    };
    AdjBufs b;
    b.ckx = (double4*)take(n4 * nwin);
    b.ckv = (double4*)take(n4 * nwin);
    b.tape = (double*)take(sizeof(double) * 12 * (size_t)T * wlen);
    b.xs = (double4*)take(n4 * wlen);
    b.oo = (double4*)take(n4 * wlen);
    b.Xb = (double4*)take(n4);
    b.Vb = (double4*)take(n4);
    b.Ob = (double4*)take(n4);
    b.gacc = take(sizeof(float) * T);
    b.glog = (double*)take(sizeof(double) * T);
    b.xloss = (double4*)take(n4 * n_loss);
    b.tgt = (double4*)take(n4 * n_loss);
    b.lw = (double*)take(sizeof(double) * N);
    b.save_x = (double4*)take(n4);
    b.save_v = (double4*)take(n4);
    double* traw = (double*)take(sizeof(double) * 3 * (size_t)N * n_loss);
    double* eterm = (double*)take(sizeof(double) * (size_t)B * n_loss);
    double* lout = (double*)take(sizeof(double) * (1 + (size_t)B));
    double* gout = (double*)take(sizeof(double) * T);
    hipStream_t st = ctx->stream;

    // Targets and per-node loss-gradient weights 2/n_b, built on the device.
    XF_HIP(ctx, hipMemcpyAsync(traw, targets, sizeof(double) * 3 * (size_t)N * n_loss, hipMemcpyHostToDevice, st));
    XF_HIP(ctx, xf_launch_expand_targets(st, N * n_loss, traw, b.tgt));
    XF_HIP(ctx, xf_launch_loss_weights(st, N, s->O, s->bp, b.lw));

    float clock_end = 0.0f;
    if ((rc = AdjointForward(s, substeps, dt, energy, sim, n_loss, loss_at, window, b, false, false, &clock_end)))
        return rc;
    XF_HIP(ctx, xf_launch_loss(st, n_loss, B, N, b.xloss, b.tgt, s->bp, eterm, lout));
    if ((rc = AdjointReverse(s, substeps, dt, energy, n_loss, loss_at, window, b, false, false, false, &clock_end)))
        return rc;
    XF_HIP(ctx, xf_launch_grad_out(st, T, b.glog, s->tout, gout));
    if (loss) XF_HIP(ctx, hipMemcpyAsync(loss, lout, sizeof(double), hipMemcpyDeviceToHost, st));
    if (body_loss) XF_HIP(ctx, hipMemcpyAsync(body_loss, lout + 1, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    if (grad_log_mu) XF_HIP(ctx, hipMemcpyAsync(grad_log_mu, gout, sizeof(double) * T, hipMemcpyDeviceToHost, st));
    XF_HIP(ctx, hipStreamSynchronize(st));
    return XF_OK;
}

namespace {
// Rebuilds the scene's kinematic list from the bodies' lists (global ids).
int UploadKinematic(xf_scene* s) {
    xf_ctx* ctx = s->ctx;
    std::vector<uint32_t> idx;
    std::vector<double4> pos;
    for (const Body& b : s->bodies)
        for (size_t k = 0; k < b.kin_nodes.size(); k++) {
            idx.push_back(b.node_base + b.kin_nodes[k]);
            pos.push_back(make_double4(b.kin_xyz[3 * k], b.kin_xyz[3 * k + 1], b.kin_xyz[3 * k + 2], 0.0));
        }
    const uint32_t n = (uint32_t)idx.size();
    XF_HIP(ctx, hipSetDevice(ctx->device));
    if (n > s->kin_cap) {
        // the stream may still read the old lists
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (s->kin_idx) (void)hipFree(s->kin_idx);
        if (s->kin_pos) (void)hipFree(s->kin_pos);
        s->kin_idx = nullptr;
        s->kin_pos = nullptr;
        s->kin_cap = 0;
        int rc;
        if ((rc = DevAlloc(ctx, &s->kin_idx, n, &s->device_bytes)) || (rc = DevAlloc(ctx, &s->kin_pos, n, &s->device_bytes)))
            return rc;
        s->kin_cap = n;
    }
    if (n) {
        XF_HIP(ctx, hipMemcpyAsync(s->kin_idx, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, ctx->stream));
        XF_HIP(ctx, hipMemcpyAsync(s->kin_pos, pos.data(), sizeof(double4) * n, hipMemcpyHostToDevice, ctx->stream));
        XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    s->nkin = n;
    return XF_OK;
}
}  // namespace

XF_API int xf_scene_set_kinematic(xf_scene* s, uint32_t body, uint32_t n, const uint32_t* nodes, const double* xyz) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    if (body >= s->bodies.size()) return Fail(s->ctx, XF_E_INVALID, "body %u out of range", body);
    if (n && (!nodes || !xyz)) return Fail(s->ctx, XF_E_INVALID, "nodes and xyz are needed");
    Body& b = s->bodies[body];
    std::vector<uint8_t> seen(b.n, 0);
    for (uint32_t k = 0; k < n; k++) {
        if (nodes[k] >= b.n) return Fail(s->ctx, XF_E_INVALID, "kinematic node %u out of range", nodes[k]);
        if (seen[nodes[k]]++) return Fail(s->ctx, XF_E_INVALID, "kinematic node %u given twice", nodes[k]);
        for (int r = 0; r < 3; r++)
            if (!std::isfinite(xyz[3 * (size_t)k + r])) return Fail(s->ctx, XF_E_INVALID, "non-finite target");
    }
    b.kin_nodes.assign(nodes, nodes + n);
    b.kin_xyz.assign(xyz, xyz + 3 * (size_t)n);
    return UploadKinematic(s);
}

XF_API int xf_scene_set_manip(xf_scene* s, uint32_t body, uint32_t node, const float target[3]) {
    if (!s) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    uint32_t g = xf::kNoNode;
    float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (node != xf::kNoNode) {
        if (body >= s->bodies.size()) return Fail(ctx, XF_E_INVALID, "body %u out of range", body);
        if (node >= s->bodies[body].n) return Fail(ctx, XF_E_INVALID, "node %u out of range", node);
        if (!target || !std::isfinite(target[0]) || !std::isfinite(target[1]) || !std::isfinite(target[2]))
            return Fail(ctx, XF_E_INVALID, "the manipulator needs a finite target");
        g = s->bodies[body].node_base + node;
        m = make_float4(target[0], target[1], target[2], 0.0f);
    }
    memcpy(&m.w, &g, sizeof(float));
    XF_HIP(ctx, hipSetDevice(ctx->device));
    if (!s->manip) {
        if (g == xf::kNoNode) return XF_OK;  // never enabled
        int rc = DevAlloc(ctx, &s->manip, 1, &s->device_bytes);
        if (rc) return rc;
    }
    XF_HIP(ctx, hipMemcpyAsync(s->manip, &m, sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->manip_on = g != xf::kNoNode;
    return XF_OK;
}

XF_API int xf_scene_count_nonfinite(xf_scene* s, uint32_t* count) {
    if (!s || !count) return XF_E_INVALID;
    if (!s->finalized) return Fail(s->ctx, XF_E_STATE, "scene is not finalized");
    xf_ctx* ctx = s->ctx;
    XF_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t zero = 0;
    XF_HIP(ctx, hipMemcpyAsync(s->counter, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    XF_HIP(ctx, xf_launch_count_nonfinite(ctx->stream, s->N, s->X, s->counter));
    XF_HIP(ctx, hipMemcpyAsync(count, s->counter, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    XF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XF_OK;
}

XF_API int xf_mesh_tet_block(uint32_t W, uint32_t H, uint32_t D, const float scale[3], const float origin[3],
                             uint32_t mirrored, float* xyz, uint32_t* tets) {
    if (!W || !H || !D || !scale || !origin || !xyz || !tets) return XF_E_INVALID;
    if ((uint64_t)W * H * D * 6 >= (1ull << 31)) return XF_E_INVALID;
    xf::TetBlock(W, H, D, scale, origin, mirrored != 0, xyz, tets);
    return XF_OK;
}

XF_API int xf_mesh_tet_block_colours(uint32_t W, uint32_t H, uint32_t D, uint32_t* colour) {
    if (!W || !H || !D || !colour) return XF_E_INVALID;
    if ((uint64_t)W * H * D * 6 >= (1ull << 31)) return XF_E_INVALID;
    size_t t = 0;
    for (uint32_t z = 0; z < D; z++)
        for (uint32_t y = 0; y < H; y++)
            for (uint32_t x = 0; x < W; x++)
                for (uint32_t k = 0; k < 6; k++) colour[t++] = 6 * ((x + y + z) % 4) + k;
    return XF_OK;
}

XF_API int xf_minstd_order(uint32_t n, uint32_t* order) {
    if (!order) return XF_E_INVALID;
    xf::MinstdOrder(n, order);
    return XF_OK;
}

XF_API int xf_normal_fill(uint64_t seed, float sigma, uint32_t n, float* out) {
    if (!out) return XF_E_INVALID;
    xf::NormalFill(seed, sigma, n, out);
    return XF_OK;
}

XF_API int xf_mesh_colour(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, uint32_t colour_order,
                          uint32_t* order, uint32_t* colour, uint32_t* n_colours) {
    if (!tets || !order || !n_colours || colour_order > XF_COLOUR_MINSTD) return XF_E_INVALID;
    for (uint64_t k = 0; k < 4ull * n_tets; k++)
        if (tets[k] >= n_nodes) return XF_E_INVALID;
    std::vector<uint32_t> col, ord;
    uint32_t nc = xf::ColourSweepOrder(n_nodes, tets, n_tets, colour_order == XF_COLOUR_MINSTD, &col, &ord);
    if (nc == 0) return XF_E_UNSUPPORTED;
    memcpy(order, ord.data(), sizeof(uint32_t) * n_tets);
    if (colour) memcpy(colour, col.data(), sizeof(uint32_t) * n_tets);
    *n_colours = nc;
    return XF_OK;
}

namespace {
thread_local std::string g_mesh_error;
}

XF_API int xf_mesh_sweep_blocks(uint32_t n_nodes, const float* xyz, uint32_t n_tets, const uint32_t* tets,
                                uint32_t colour_order, uint32_t* n_blocks, uint32_t* block_of_tet,
                                uint32_t* phase_of_tet, uint32_t* launch_of_block, uint32_t* max_phases) {
    if (!xyz || !tets || !n_blocks || !block_of_tet || !phase_of_tet || !launch_of_block ||
        colour_order > XF_COLOUR_MINSTD)
        return XF_E_INVALID;
    std::vector<uint32_t> visit(n_tets);
    if (colour_order == XF_COLOUR_MINSTD) xf::MinstdOrder(n_tets, visit.data());
    else
        for (uint32_t t = 0; t < n_tets; t++) visit[t] = t;
    xf::SweepBlocks sb;
    std::string err;
    if (!xf::BuildSweepBlocks(xyz, n_nodes, tets, n_tets, visit.data(), &sb, &err)) {
        g_mesh_error = err;
        return err == "tet index out of range" ? XF_E_INVALID : XF_E_UNSUPPORTED;
    }
    *n_blocks = sb.n_blocks;
    memcpy(block_of_tet, sb.block.data(), sizeof(uint32_t) * n_tets);
    memcpy(phase_of_tet, sb.phase.data(), sizeof(uint32_t) * n_tets);
    uint32_t mp = 0;
    for (uint32_t k = 0; k < sb.n_blocks; k++) {
        launch_of_block[k] = sb.launch[k];
        mp = std::max(mp, sb.phases[k]);
    }
    if (max_phases) *max_phases = mp;
    return XF_OK;
}

XF_API int xf_mesh_rings(uint32_t n_tets, const uint32_t* tets, uint32_t* n_rings, uint32_t* node_total,
                         uint32_t* node_off, uint32_t* nodes, uint32_t* tet_off, uint32_t* ring_tets, uint8_t* idx,
                         uint32_t* largest) {
    if (!tets || !n_rings || !node_total) return XF_E_INVALID;
    xf::Rings r;
    if (!xf::BuildRings(tets, n_tets, &r)) return XF_E_UNSUPPORTED;
    *n_rings = r.n;
    *node_total = (uint32_t)r.nodes.size();
    if (largest) *largest = r.largest;
    if (node_off) memcpy(node_off, r.node_off.data(), sizeof(uint32_t) * r.node_off.size());
    if (nodes) memcpy(nodes, r.nodes.data(), sizeof(uint32_t) * r.nodes.size());
    if (tet_off) memcpy(tet_off, r.tet_off.data(), sizeof(uint32_t) * r.tet_off.size());
    if (ring_tets) memcpy(ring_tets, r.tets.data(), sizeof(uint32_t) * r.tets.size());
    if (idx) memcpy(idx, r.idx.data(), r.idx.size());
    return XF_OK;
}

XF_API const char* xf_mesh_error(void) { return g_mesh_error.c_str(); }

XF_API int xf_mesh_read_tetgen(const char* node_path, const char* ele_path, uint32_t orient, uint32_t* n_nodes,
                               uint32_t* n_tets, float* xyz, uint32_t* tets) {
    if (!node_path || !ele_path || !n_nodes || !n_tets) {
        g_mesh_error = "null argument";
        return XF_E_INVALID;
    }
    std::vector<float> x;
    std::vector<uint32_t> t;
    if (!xf::ReadTetgen(node_path, ele_path, orient != 0, &x, &t, &g_mesh_error)) return XF_E_INVALID;
    if (x.size() / 3 > 0xFFFFFFFFull || t.size() / 4 > 0xFFFFFFFFull) {
        g_mesh_error = "mesh too large";
        return XF_E_INVALID;
    }
    if (xyz || tets) {
        if (!xyz || !tets || *n_nodes != x.size() / 3 || *n_tets != t.size() / 4) {
            g_mesh_error = "output arrays must match the counts of the first call";
            return XF_E_INVALID;
        }
        memcpy(xyz, x.data(), sizeof(float) * x.size());
        memcpy(tets, t.data(), sizeof(uint32_t) * t.size());
    }
    *n_nodes = (uint32_t)(x.size() / 3);
    *n_tets = (uint32_t)(t.size() / 4);
    return XF_OK;
}

XF_API int xf_mesh_tets_from_stream(const uint32_t* idx, uint32_t count, uint32_t* n_tets, uint32_t* tets) {
    if ((!idx && count) || !n_tets) {
        g_mesh_error = "null argument";
        return XF_E_INVALID;
    }
    std::vector<uint32_t> t;
    if (!xf::TetsFromStream(idx, count, &t)) {
        g_mesh_error = "index stream holds an element other than a 4-node tet";
        return XF_E_UNSUPPORTED;
    }
    if (tets) {
        if (*n_tets != t.size() / 4) {
            g_mesh_error = "output array must match the count of the first call";
            return XF_E_INVALID;
        }
        memcpy(tets, t.data(), sizeof(uint32_t) * t.size());
    }
    *n_tets = (uint32_t)(t.size() / 4);
    return XF_OK;
}

XF_API int xf_ldlt_solve(uint32_t n, const float* A, const float* b, float* x) {
    if (!A || !b || !x) return XF_E_INVALID;
    return xf::LdltSolve(n, A, b, x) == 0 ? XF_OK : XF_E_INVALID;
}

}  // extern "C"
