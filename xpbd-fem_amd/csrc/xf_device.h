// xf_device.h -- layouts shared by the host code and the gfx950 kernels.
//
// Device memory of a scene (all bodies concatenated):
//   nodes  X[N]  double4 (x, y, z, w)        w = inverse mass, 0 when pinned
//          O[N]  double4 (x, y, z, flags)    flags bits: 0 pinned, 1.. body id
//          V[N]  double4 (vx, vy, vz, 0)     written at the end of each step
//                                            call (every substep with damping)
//   tets in sweep order (colour-major, body-major inside a colour):
//          tidx[T] uint4 (global node ids; node 3 is the root node)
//          tg0..tg3[T] float4: Qi (column-major) | QQ | QR | comp0
//            tg0 = Qi00 Qi01 Qi02 Qi10   tg1 = Qi11 Qi12 Qi20 Qi21
//            tg2 = Qi22 QQ0  QQ1  QQ2    tg3 = QR0  QR1  QR2  comp0
//          taux[T] float4 (comp1, a, body id bits, compliance * volume) --
//            read by the projection only when the scene is compressible or
//            holds more than one body; always by the volume and damping passes
//          tdmp[T] float2 (surface area, position in the body's sweep order
//            as uint bits) -- damping passes only
//   vertex rings in sweep order (colour-major, body-major inside a colour),
//   built when a body first uses a vertex-ring energy:
//          rmeta[R] uint4 (node offset, node count, ring-tet offset, tet count)
//          rinfo[R] uint4 (body id, position in the body's ring sweep order,
//            ring volume as float bits, 0)
//          rnodes[]  uint32 global node ids, the ring's own vertex first
//          rincb[] uint16 per ring node, rinc[4 * ring tets] uint16: each ring
//            node's (ring tet << 2 | corner) slots in ring-tet order
//          rtet[2 * ring tets] float4 pairs (QQ0 QQ1 QQ2 QR0 | QR1 QR2 V idx)
//            idx = four ring-local node indices, one byte each, as float bits
//   body table  XfBodyParams[B]
#pragma once
#include <stdint.h>

struct XfBodyParams {
    float gdx, gdy;      // gravity * dt (Geo.cpp:308)
    float dragf;         // 1 - timeCorrectedDrag (Geo.cpp:309)
    float dt;            // substep length
    float invdt;         // 1.0f / dt (Geo.cpp:343)
    uint32_t collide;    // XF_COLLIDE_* bits
    float ground_y;
    float friction;
    float sc0[3];
    float sv[3];
    float sr;
    uint32_t energy;     // 4 or 7 (per-body dispatch in multi-body scenes)
    float yeoh[3];
    uint32_t volume_passes;  // Settings::volumePasses (Geo.cpp:779-782)
    // Damping (GeoLinear3d::Damp, Geo.cpp:790-811; Xpbd.h:87-348)
    float damping;           // Settings::damping (Rayleigh gamma)
    uint32_t rtype;          // Rayleigh_Paper 0, Limit 1, Post 2, PostAmortized 3
    float pbd;               // volumeAndTimeCorrectedPbdDamping (pass on when > 0)
    float pbd_amortized;     // amortizedVolumeAndTimeCorrectedPbdDamping
    uint32_t n_tets;         // tets of the body (amortized range)
    uint32_t simultaneous;   // Settings_XpbdSolveBit (RayleighDamp form)
    // Vertex-ring energies (SolveSimplex, Geo.cpp:13-64)
    uint32_t n_rings;        // rings of the body (amortized range of the ring damping pass)
    float compliance;        // Settings::compliance
    float poisson;           // Settings::poissonsRatio
    // Node range of the body in the scene (adjoint loss terms)
    uint32_t node_base;
    uint32_t n_nodes;
    uint32_t unlock;         // 1: pins are not locked this step (Settings_LockLeft clear, Geo.cpp:318)
    // Settings as given, for the FP64 element model of the adjoint (which
    // forms gravity * dt and 1 - drag in double)
    float gravity[2];
    float drag;
};

// The scene clock buffer: [0] simulation time (float), [1] tick (uint32,
// Settings::tickId of the next substep). The first colour launch of a
// substep advances both; kernels later in the same substep use tick - 1.
constexpr int kClockWords = 2;

// Kernel selector for the projection launch.
struct XfProjectArgs {
    uint32_t begin;      // first tet of the colour in sweep order
    uint32_t count;      // tets in the colour
};
