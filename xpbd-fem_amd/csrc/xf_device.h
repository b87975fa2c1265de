// xf_device.h -- layouts shared by the host code and the gfx950 kernels.
//
// Device memory of a scene (all bodies concatenated):
//   nodes  X[N]  float4 (x, y, z, w)        w = inverse mass, 0 when pinned
//          O[N]  float4 (x, y, z, flags)    flags bits: 0 pinned, 1.. body id
//          V[N]  float4 (vx, vy, vz, 0)     written at the end of each step call
//   tets in sweep order (colour-major, body-major inside a colour):
//          tidx[T] uint4 (global node ids; node 3 is the root node)
//          tg0..tg3[T] float4: Qi (column-major) | QQ | QR | comp0
//            tg0 = Qi00 Qi01 Qi02 Qi10   tg1 = Qi11 Qi12 Qi20 Qi21
//            tg2 = Qi22 QQ0  QQ1  QQ2    tg3 = QR0  QR1  QR2  comp0
//          taux[T] float4 (comp1, a, body id bits, 0) -- read only when the
//            scene is compressible or holds more than one body
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
    float pad;
};

// Kernel selector for the projection launch.
struct XfProjectArgs {
    uint32_t begin;      // first tet of the colour in sweep order
    uint32_t count;      // tets in the colour
};
