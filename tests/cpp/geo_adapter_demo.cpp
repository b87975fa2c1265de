// Drives GeoGpuLinear3d the way the reference's Sim::Update drives its Geo
// objects (Demo.cpp:67-90): a Settings-shaped struct, the [4, i0..i3]* index
// stream of GenerateTetBlock, and one Substep call per substep. Writes the
// final FP64 positions and velocities to the file given as argv[1].
// Usage: geo_adapter_demo OUT FRAMES SUBSTEPS ENERGY [RAYLEIGH_TYPE DAMPING PBD
// PBD_AMORTIZED [AUTORESIZE [SWEEP_BLOCKS]]]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../xpbd-fem_amd/csrc/xf_geo_gpu.h"

namespace {
struct Vec2 {
    float x, y;
};
// Field names of the reference's Settings (Settings.h:79-102).
struct DemoSettings {
    float substepsPerSecond = 1200.0f;
    uint32_t volumePasses = 0;
    Vec2 gravity = {0.0f, -9.81f};
    float compliance = 1e-5f;
    float damping = 0.0f;
    float pbdDamping = 0.0f;
    float poissonsRatio = 0.5f;
    uint32_t flags = 0;
    float volumeAndTimeCorrectedPbdDamping = 0.0f;
    float amortizedVolumeAndTimeCorrectedPbdDamping = 0.0f;
    float timeCorrectedDrag = 0.0f;
    uint32_t tickId = 0;
};
struct DemoManipulator {
    void* pickedGeo = nullptr;
};
}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr,
                "usage: %s OUT FRAMES SUBSTEPS ENERGY [RAYLEIGH_TYPE DAMPING PBD PBD_AMORTIZED [AUTORESIZE "
                "[SWEEP_BLOCKS]]]\n",
                argv[0]);
        return 2;
    }
    const int frames = atoi(argv[2]), substeps = atoi(argv[3]);
    const uint32_t energy = (uint32_t)atoi(argv[4]);
    const uint32_t W = 12, H = 4, D = 4;
    std::vector<float> xyz(3 * (W + 1) * (H + 1) * (D + 1));
    std::vector<uint32_t> tets(24 * W * H * D);
    const float scale[3] = {0.1f, 0.1f, 0.1f}, origin[3] = {0.0f, -0.2f, -0.2f};
    if (xf_mesh_tet_block(W, H, D, scale, origin, 0, xyz.data(), tets.data())) return 1;
    std::vector<uint32_t> stream;  // the reference's index stream layout
    for (size_t t = 0; t < tets.size() / 4; t++) {
        stream.push_back(4);
        for (int c = 0; c < 4; c++) stream.push_back(tets[4 * t + c]);
    }
    xfgpu::GeoGpuLinear3d geo;
    const bool auto_resize = argc >= 10 && atoi(argv[9]) != 0;  // GeoLinear3d::Init's autoResize
    int rc = geo.Init(1000.0f, xyz.data(), (uint32_t)xyz.size(), stream.data(), (uint32_t)stream.size(), auto_resize);
    if (rc) {
        fprintf(stderr, "Init failed (%d): %s\n", rc, geo.LastError());
        return 1;
    }
    geo.SetGround(-0.5f, 0.5f);
    if (argc >= 11 && atoi(argv[10]) != 0 && (rc = geo.SetSweepBlocks(true))) return 1;
    DemoSettings settings;
    settings.flags = (energy << xfgpu::kEnergyBit) | (1u << xfgpu::kXpbdSolveBit) | xfgpu::kLockLeft;
    if (argc >= 9) {  // damping: Settings flag bits 20-21 and the derived PBD factors
        settings.flags |= ((uint32_t)atoi(argv[5]) & xfgpu::kRayleighMask) << xfgpu::kRayleighBit;
        settings.damping = (float)atof(argv[6]);
        settings.volumeAndTimeCorrectedPbdDamping = (float)atof(argv[7]);
        settings.amortizedVolumeAndTimeCorrectedPbdDamping = (float)atof(argv[8]);
        settings.pbdDamping = settings.volumeAndTimeCorrectedPbdDamping > 0.0f ? 1.0f : 0.0f;
    }
    DemoManipulator manip;
    const float dt = 1.0f / settings.substepsPerSecond;
    float v0 = geo.CalculateVolume();
    for (int f = 0; f < frames; f++)
        for (int s = 0; s < substeps; s++)
            if ((rc = geo.Substep(settings, manip, dt))) {
                fprintf(stderr, "Substep failed (%d): %s\n", rc, geo.LastError());
                return 1;
            } else {
                ++settings.tickId;  // Sim::Update advances tickId once per substep (Demo.cpp:81-89)
            }
    std::vector<double> X(3 * geo.VertCount()), V(3 * geo.VertCount());
    if ((rc = geo.GetState(X.data(), V.data()))) return 1;
    FILE* f = fopen(argv[1], "wb");
    if (!f) return 1;
    fwrite(X.data(), sizeof(double), X.size(), f);
    fwrite(V.data(), sizeof(double), V.size(), f);
    fclose(f);
    printf("nodes %u elements %u volume %.6f -> %.6f\n", geo.VertCount(), geo.ElementCount(), v0,
           geo.CalculateVolume());
    return 0;
}
