// xf_geo_gpu.h -- header-only C++ adapter with the step interface of the
// reference's GeoLinear3d (Geo.h:146-164; Geo3d::Substep Geo.cpp:305-356),
// backed by the gfx950 library through the C ABI of include/xpbdfem.h.
//
// The reference calls geos[i]->Substep(settings, manip, dt) once per substep
// (Demo.cpp:86-88). GeoGpuLinear3d::Substep has the same signature and
// meaning; Frame() runs several substeps as one captured hipGraph.
//
// The settings and manipulator types are template parameters: any struct
// with the reference's Settings field names (flags, compliance,
// poissonsRatio, gravity.x/.y, damping, pbdDamping,
// volumeAndTimeCorrectedPbdDamping,
// amortizedVolumeAndTimeCorrectedPbdDamping, timeCorrectedDrag,
// volumePasses, tickId; Settings.h:79-102) works, so the adapter needs none
// of the reference's headers. What needs the reference's own vector math
// (Transform, the Settings_LockRight targets, the manipulator's target
// point) is done by the Geo subclass in GeoGpu.h, which hands the results
// over through SetPositions, SetKinematic and SetManip; a step whose
// settings ask for Settings_LockRight, or whose manipulator picked this
// body, without them fails with XF_E_UNSUPPORTED.
#pragma once

#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xpbdfem.h"

namespace xfgpu {

// Flag layout of Settings::flags (Settings.h:9-29).
constexpr uint32_t kEnergyBit = 6, kEnergyMask = 31, kXpbdSolveBit = 11, kRayleighBit = 20, kRayleighMask = 3,
                   kLockLeft = 1u << 26, kLockRight = 1u << 27;
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

class GeoGpuLinear3d {
   public:
    GeoGpuLinear3d() = default;
    GeoGpuLinear3d(const GeoGpuLinear3d&) = delete;
    GeoGpuLinear3d& operator=(const GeoGpuLinear3d&) = delete;
    ~GeoGpuLinear3d() {
        if (scene_) xf_scene_destroy(scene_);
        if (ctx_) xf_ctx_destroy(ctx_);
    }

    // GeoLinear3d::Init (Geo.cpp:697): node data xyz, index stream
    // [4, i0, i1, i2, i3]* (only tets). Nodes whose normalised x is below 2%
    // get the Left flag (Geo.cpp:724); they are pinned when the first
    // Substep sees Settings_LockLeft. Returns an xf_status.
    // autoResize: positions become 0.085 (x - centre) / largest extent in FP64,
    // as GeoLinear3d::Init computes them (Geo.cpp:712-729), after the 2 % pin
    // rule has been applied to the original positions.
    int Init(float density, const float* nodeData, uint32_t nodeDataCount, const uint32_t* idxData,
             uint32_t idxDataCount, bool autoResize = false, int device = 0) {
        if (nodeDataCount % 3) return XF_E_INVALID;
        nodes_.assign(nodeData, nodeData + nodeDataCount);
        uint32_t nt = 0;
        int rc = xf_mesh_tets_from_stream(idxData, idxDataCount, &nt, nullptr);  // tets only
        if (rc) return rc;
        tets_.assign(4 * (size_t)nt, 0u);
        if ((rc = xf_mesh_tets_from_stream(idxData, idxDataCount, &nt, tets_.data()))) return rc;
        density_ = density;
        n_ = nodeDataCount / 3;
        float mn = 1e24f, mx = -1e24f;
        for (uint32_t i = 0; i < n_; i++) {
            mn = std::fmin(mn, nodes_[3 * i]);
            mx = std::fmax(mx, nodes_[3 * i]);
        }
        left_.assign(n_, 0);
        right_.assign(n_, 0);
        for (uint32_t i = 0; i < n_; i++) {
            if ((nodes_[3 * i] - mn) / (mx - mn) < 0.02f) left_[i] = 1;
            if ((nodes_[3 * i] - mn) / (mx - mn) > 0.98f) right_[i] = 1;  // Geo.cpp:725
        }
        nodes64_.clear();
        if (autoResize) {
            float lo[3] = {1e24f, 1e24f, 1e24f}, hi[3] = {-1e24f, -1e24f, -1e24f};
            for (uint32_t i = 0; i < n_; i++)
                for (int r = 0; r < 3; r++) {
                    lo[r] = std::fmin(lo[r], nodes_[3 * i + r]);
                    hi[r] = std::fmax(hi[r], nodes_[3 * i + r]);
                }
            float dims[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            double max_dim = std::fmax(dims[0], std::fmax(dims[1], dims[2]));
            nodes64_.resize(3 * (size_t)n_);
            for (uint32_t i = 0; i < n_; i++)
                for (int r = 0; r < 3; r++) {
                    double centre = 0.5 * (double)(lo[r] + hi[r]);
                    nodes64_[3 * i + r] = 0.085 * (((double)nodes_[3 * i + r] - centre) / max_dim);
                }
        }
        return xf_ctx_create(device, &ctx_);
    }

    // Geo3d::Substep: one substep of length dt. owner: the object the
    // caller's manipulator points at when it picked this body (the Geo that
    // holds the adapter), or null.
    template <class SettingsT, class ManipT>
    int Substep(const SettingsT& settings, const ManipT& manip, float dt, const void* owner = nullptr) {
        if (owner && static_cast<const void*>(manip.pickedGeo) == owner && !manip_set_) return XF_E_UNSUPPORTED;
        int rc = Frame(settings, 1, dt);
        manip_set_ = kinematic_set_ = false;
        return rc;
    }

    // `substeps` consecutive Substep calls, enqueued as one hipGraph.
    template <class SettingsT>
    int Frame(const SettingsT& settings, uint32_t substeps, float dt) {
        if ((settings.flags & kLockRight) && !kinematic_set_) return XF_E_UNSUPPORTED;
        int rc = EnsureScene(settings.flags);
        if (rc) return rc;
        xf_settings s;
        xf_settings_default(&s);
        s.energy = (settings.flags >> kEnergyBit) & kEnergyMask;
        s.simultaneous = (settings.flags >> kXpbdSolveBit) & 1u;
        s.compliance = settings.compliance;
        s.poissons_ratio = settings.poissonsRatio;
        s.gravity[0] = settings.gravity.x;
        s.gravity[1] = settings.gravity.y;
        s.drag = settings.timeCorrectedDrag;
        s.damping = settings.damping;
        s.rayleigh_type = (settings.flags >> kRayleighBit) & kRayleighMask;
        // the PBD pass runs when pbdDamping > 0 (Geo.cpp:806) with the
        // volume- and time-corrected factors derived from it (Demo.cpp:53-61)
        s.pbd_damping = settings.pbdDamping > 0.0f ? settings.volumeAndTimeCorrectedPbdDamping : 0.0f;
        s.pbd_damping_amortized = settings.amortizedVolumeAndTimeCorrectedPbdDamping;
        s.volume_passes = settings.volumePasses;
        s.collide = collide_;
        s.ground_y = ground_y_;
        s.friction = friction_;
        // with the reference's pins the Left nodes are pinned whenever the
        // step's settings lock them (Geo.cpp:318)
        s.unlock_pins = reference_pins_ && !(settings.flags & kLockLeft) ? 1u : 0u;
        if ((rc = xf_scene_set_settings(scene_, 0, &s))) return rc;
        uint32_t tick = 0;
        if ((rc = xf_scene_get_tick(scene_, &tick))) return rc;
        if (tick != settings.tickId && (rc = xf_scene_set_tick(scene_, settings.tickId))) return rc;
        return xf_scene_step(scene_, substeps, dt);
    }

    // Ground-plane contact (no reference counterpart).
    void SetGround(float y, float friction) {
        collide_ |= XF_COLLIDE_GROUND;
        ground_y_ = y;
        friction_ = friction;
    }

    // GeoLinear3d::CalculateVolume (Geo.cpp:827-832): sum of signed tet volumes.
    float CalculateVolume() {
        std::vector<double> X(3 * (size_t)n_);
        if (GetState(X.data(), nullptr)) return NAN;
        float volume = 0.0f;
        for (size_t t = 0; t < tets_.size() / 4; t++) {
            const uint32_t* is = &tets_[4 * t];
            float m[3][3];
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) m[c][r] = (float)(X[3 * is[c] + r] - X[3 * is[3] + r]);
            float det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                        m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                        m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
            volume += (1.0f / 6.0f) * det;
        }
        return volume;
    }

    uint32_t VertCount() const { return n_; }
    uint32_t ElementCount() const { return (uint32_t)(tets_.size() / 4); }

    // FP64 positions / velocities (the reference's dvec3 X, V).
    int GetState(double* X, double* V) {
        if (!scene_) {
            if (X)
                for (uint32_t i = 0; i < 3 * n_; i++) X[i] = nodes64_.empty() ? (double)nodes_[i] : nodes64_[i];
            if (V) memset(V, 0, sizeof(double) * 3 * n_);
            return XF_OK;
        }
        int rc = xf_scene_sync(scene_);
        return rc ? rc : xf_scene_get_state(scene_, 0, X, V);
    }

    const char* LastError() const { return ctx_ ? xf_last_error(ctx_) : "no context"; }

    // The reference's pin behaviour (before the scene exists): the Left nodes
    // are pinned in the substeps whose settings have Settings_LockLeft, and
    // keep inverse mass 1/m until the end of their first locked substep
    // (GeoLinear3d::Init sets w = 1/m for every node, Geo.cpp:744; the lock
    // sets it to 0, Geo.cpp:318-322). Off: the Left nodes are pinned from the
    // first substep with inverse mass 0 when its settings lock them.
    int SetReferencePins(bool on) {
        if (scene_) return XF_E_STATE;
        reference_pins_ = on;
        return XF_OK;
    }

    // Creates the scene now (it is otherwise created by the first Substep,
    // whose settings decide the pins unless SetReferencePins is on).
    int Prepare(uint32_t settings_flags) { return EnsureScene(settings_flags); }

    // FP64 positions of every node (Geo3d::Transform's result, Geo.cpp:358-364).
    int SetPositions(const double* X) {
        int rc = EnsureScene(kLockLeft);
        return rc ? rc : xf_scene_set_state(scene_, 0, X, nullptr);
    }

    // Nodes held at FP64 targets after the sweep of the next Substep
    // (Settings_LockRight, Geo.cpp:324-331); n = 0 clears.
    int SetKinematic(uint32_t n, const uint32_t* nodes, const double* xyz) {
        int rc = EnsureScene(kLockLeft);
        if (rc) return rc;
        if (n || kinematic_n_) rc = xf_scene_set_kinematic(scene_, 0, n, nodes, xyz);
        kinematic_n_ = n;
        kinematic_set_ = rc == XF_OK;
        return rc;
    }

    // The manipulation constraint on `node` toward `target` in the next
    // Substep (Geo.cpp:334-339); node = kNoNode: none.
    int SetManip(uint32_t node, const float target[3]) {
        int rc = EnsureScene(kLockLeft);
        if (rc) return rc;
        if (node != kNoNode || manip_on_) rc = xf_scene_set_manip(scene_, 0, node, target);
        manip_on_ = node != kNoNode;
        manip_set_ = rc == XF_OK;
        return rc;
    }

    // The scene's sweep order (the reference's tOrder that reproduces it).
    int Order(uint32_t* order) {
        int rc = EnsureScene(kLockLeft);
        return rc ? rc : xf_scene_get_order(scene_, 0, order, nullptr);
    }

    // Rest positions (X0 of GeoLinear3d::Init, after autoResize), node flags.
    double RestPosition(uint32_t i, int r) const {
        return nodes64_.empty() ? (double)nodes_[3 * i + r] : nodes64_[3 * i + r];
    }
    bool IsLeft(uint32_t i) const { return left_[i] != 0; }
    bool IsRight(uint32_t i) const { return right_[i] != 0; }
    const std::vector<uint32_t>& Tets() const { return tets_; }

    // The sweep-block schedule (xf_scene_opts.sweep_blocks with
    // XF_LAUNCH_BLOCKS: 4 projection launches per substep, each workgroup
    // sweeping one block through its phases in LDS) for the scene created at
    // the first Substep. Vertex-ring energies (Settings energy 8..10) then
    // fail: they run per colour only. Returns XF_E_STATE after the first
    // Substep.
    int SetSweepBlocks(bool on) {
        if (scene_) return XF_E_STATE;
        sweep_blocks_ = on;
        return XF_OK;
    }

   private:
    int EnsureScene(uint32_t flags) {
        if (scene_) return XF_OK;
        if (!ctx_) return XF_E_STATE;
        int rc = xf_scene_create(ctx_, &scene_);
        if (rc) return rc;
        xf_body_desc d;
        memset(&d, 0, sizeof(d));
        if (nodes64_.empty()) d.xyz = nodes_.data();
        else d.xyz64 = nodes64_.data();
        d.n_nodes = n_;
        d.tets = tets_.data();
        d.n_tets = (uint32_t)(tets_.size() / 4);
        d.density = density_;
        d.pinned = (reference_pins_ || (flags & kLockLeft)) ? left_.data() : nullptr;
        d.pin_first_substep = reference_pins_ ? 1u : 0u;
        if ((rc = xf_scene_add_body(scene_, &d, nullptr))) return rc;
        xf_scene_opts o;
        memset(&o, 0, sizeof(o));
        o.colour_order = XF_COLOUR_NATURAL;
        o.block_size = 64;
        o.use_graph = 1;
        o.launch_mode = sweep_blocks_ ? XF_LAUNCH_BLOCKS : XF_LAUNCH_PER_COLOUR;
        o.sweep_blocks = sweep_blocks_ ? 1u : 0u;
        return xf_scene_finalize(scene_, &o);
    }

    xf_ctx* ctx_ = nullptr;
    xf_scene* scene_ = nullptr;
    std::vector<float> nodes_;
    std::vector<double> nodes64_;  // autoresized FP64 positions (empty without autoResize)
    std::vector<uint32_t> tets_;
    std::vector<uint8_t> left_, right_;
    float density_ = 0.0f;
    uint32_t n_ = 0;
    uint32_t collide_ = 0;
    float ground_y_ = 0.0f, friction_ = 0.0f;
    bool sweep_blocks_ = false;
    bool reference_pins_ = false;
    uint32_t kinematic_n_ = 0;
    bool kinematic_set_ = false, manip_on_ = false, manip_set_ = false;
};

}  // namespace xfgpu
