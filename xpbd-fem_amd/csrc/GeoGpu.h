// GeoGpu.h -- a Geo (Geo.h:15-35) whose substeps run on an MI355X.
//
// Drop-in for the reference's GeoLinear3d on linear-tet meshes: it sits
// beside Geo.h in the reference's source tree (it includes the reference's
// headers: Geo.h, DebugGeo.h), Sim::AddBlock (Demo.cpp:127-150) creates it
// for Element_T4 with GeoLinear3d::Init's arguments, and every other caller
// -- FinishAddingBlocks' Transform (Demo.cpp:153-163), Sim::Update's
// geos[i]->Substep loop (Demo.cpp:67-90), picking, Render, CalculateVolume --
// stays as it is. The substep itself runs through xfgpu::GeoGpuLinear3d
// (xf_geo_gpu.h) on the C ABI of include/xpbdfem.h; what needs the
// reference's vector math runs here, with the reference's own types, so the
// results are the reference's bits:
//   * Transform: positions mapped as Geo3d::Transform maps them (Geo.cpp:
//     358-364), origin with them;
//   * Settings_LockRight: the Right nodes' targets origin +
//     lockedRightTransform3d X0 (Geo.cpp:324-331), handed over per substep;
//   * the manipulator's target point (Geo.cpp:334-339) when it picked this
//     body;
//   * Pick, Render and CalculateVolume on a host copy of the positions.
// Pins follow the reference exactly: Left nodes are locked in the substeps
// whose settings have Settings_LockLeft and keep inverse mass 1/m until
// their first lock. Element arithmetic, sweep order semantics and
// precision are those of the reference: with the sweep order the GPU
// reports (Order) as the GeoLinear3d's tOrder, both give the same bits.
// Errors of the GPU calls (which Geo's void methods cannot return) are kept
// in Status() / LastError().
#pragma once

#include <cfloat>
#include <cstdint>
#include <vector>

#include "DebugGeo.h"
#include "Geo.h"
#include "xf_geo_gpu.h"

struct GeoGpu : public Geo {
    // GeoLinear3d::Init's arguments (Geo.h:147). The connectivity (edges
    // for Render) is built with the reference's own Connectivity3d::Init in
    // `alloc`; the node and element data live on the GPU.
    int Init(Allocator* alloc, float density, const float* nodeData, uint32_t nodeDataCount, const uint32_t* idxData,
             uint32_t idxDataCount, bool autoResize, int device = 0) {
        con.Init(alloc, idxData, idxDataCount);
        if ((status = impl.Init(density, nodeData, nodeDataCount, idxData, idxDataCount, autoResize, device)))
            return status;
        if ((status = impl.SetReferencePins(true)) || (status = impl.Prepare(0u))) return status;
        const uint32_t n = impl.VertCount();
        X0.resize(n);
        for (uint32_t i = 0; i < n; i++)
            X0[i] = dvec3(impl.RestPosition(i, 0), impl.RestPosition(i, 1), impl.RestPosition(i, 2));
        for (uint32_t i = 0; i < n; i++)
            if (impl.IsRight(i)) right.push_back(i);
        everLocked.assign(n, 0);
        return status = FetchPositions();
    }

    void Substep(const Settings& settings, const Manipulator& manip, float dt) override {
        if (status) return;
        // Settings_LockRight: the Right nodes follow the animated transform
        if (settings.flags & Settings_LockRight) {
            std::vector<double> target(3 * right.size());
            for (size_t k = 0; k < right.size(); k++) {
                dvec3 p = dvec3(origin + (settings.lockedRightTransform3d * vec3(X0[right[k]])));
                target[3 * k] = p.x;
                target[3 * k + 1] = p.y;
                target[3 * k + 2] = p.z;
                everLocked[right[k]] = 1;
            }
            if ((status = impl.SetKinematic((uint32_t)right.size(), right.data(), target.data()))) return;
        } else if ((status = impl.SetKinematic(0, nullptr, nullptr))) {
            return;
        }
        // the manipulator's target point on the plane of the original pick
        if (manip.pickedGeo == this) {
            float t = dot(manip.manipPlaneNormal, manip.pick0 - manip.pos) /
                      dot(manip.manipPlaneNormal, manip.pickDirTarget);
            vec3 target = manip.pos + t * manip.pickDirTarget;
            const float tp[3] = {target.x, target.y, target.z};
            if ((status = impl.SetManip(manip.pickedPointIdx, tp))) return;
        } else if ((status = impl.SetManip(xfgpu::kNoNode, nullptr))) {
            return;
        }
        if (settings.flags & Settings_LockLeft)
            for (uint32_t i = 0; i < impl.VertCount(); i++)
                if (impl.IsLeft(i)) everLocked[i] = 1;
        if ((status = impl.Substep(settings, manip, dt, this))) return;
        fresh = false;
    }

    // Geo3d::Transform: every node moves as the reference maps it; O is set
    // by the next prediction, the rest positions X0 stay.
    void Transform(mat3 t3) override {
        if (status || (status = FetchPositions())) return;
        mat4 t4 = mat4(vec4(t3[0], 0.0f), vec4(t3[1], 0.0f), vec4(0.0f, 0.0f, 1.0f, 0.0f), vec4(t3[2].xy, 0.0f, 1.0f));
        for (size_t i = 0; i < X.size(); i++) X[i] = dvec3((t4 * vec4(vec3(X[i]), 1.0f)).xyz);
        origin = (t4 * vec4(origin, 1.0f)).xyz;
        std::vector<double> flat(3 * X.size());
        for (size_t i = 0; i < X.size(); i++) {
            flat[3 * i] = X[i].x;
            flat[3 * i + 1] = X[i].y;
            flat[3 * i + 2] = X[i].z;
        }
        status = impl.SetPositions(flat.data());
    }

    // Geo3d::Pick over the current positions: nodes whose inverse mass is 0
    // (locked Left / Right nodes) are not pickable.
    void Pick(vec3 rayOrigin, vec3 rayDir, vec3* outNearestPoint, uint32_t* outNearestPointIdx,
              float* outDistance) override {
        if (status || (status = FetchPositions())) return;
        float weightedDist = FLT_MAX;
        for (uint32_t i = 0; i < (uint32_t)X.size(); i++) {
            if (everLocked[i]) continue;
            vec3 p = vec3(X[i]);
            // nodes behind the ray origin count as infinitely far, as in
            // Geo3d::Pick (so one of them is taken when nothing is in front)
            float iDist = FLT_MAX, iWeighted = FLT_MAX;
            float t = dot(rayDir, p - rayOrigin);
            if (t >= 0.0f) {
                iDist = distance(p, rayOrigin + t * rayDir);
                iWeighted = t * iDist;
            }
            if (weightedDist >= iWeighted) {
                weightedDist = iWeighted;
                *outDistance = iDist;
                *outNearestPoint = p;
                *outNearestPointIdx = i;
            }
        }
    }

    void Constrain(const Settings&, float) override {}  // inside Substep on the GPU
    void Damp(const Settings&, float) override {}       // inside Substep on the GPU

    // GeoLinear3d::Render's edges (Geo.cpp:813-825) from the current positions.
    void Render(const Settings& settings) override {
        if (status || (status = FetchPositions())) return;
        bool tooBig = con.edgeCount > 10000;
        for (uint32_t i = 0; i < con.edgeCount; i++) {
            const ConEdge3d& e = con.edges[i];
            bool isSurface = con.surfaceEdges[i];
            if (!isSurface && tooBig) continue;
            vec3 c = isSurface ? color : mix(color, vec3(1.0f), 0.7f);
            bool left = impl.IsLeft(e.v[0]) && impl.IsLeft(e.v[1]);
            bool rgt = impl.IsRight(e.v[0]) && impl.IsRight(e.v[1]);
            if (((settings.flags & Settings_LockLeft) && left) || ((settings.flags & Settings_LockRight) && rgt))
                c = vec3(0.7f, 0.0f, 0.0f);
            DLINE(vec3(X[e.v[0]]), vec3(X[e.v[1]]), c);
        }
    }

    // GeoLinear3d::CalculateVolume: the reference's linear-tet element volume
    // summed over the tets (Geo.cpp:827-832, Fem.cpp:1067).
    float CalculateVolume() const override {
        GeoGpu* self = const_cast<GeoGpu*>(this);
        if (self->status || (self->status = self->FetchPositions())) return NAN;
        const std::vector<uint32_t>& tets = impl.Tets();
        float volume = 0.0f;
        for (size_t k = 0; k < tets.size() / 4; k++) {
            const uint32_t* is = &tets[4 * k];
            volume += (1.0f / 6.0f) * determinant(mat3(vec3(X[is[0]] - X[is[3]]), vec3(X[is[1]] - X[is[3]]),
                                                       vec3(X[is[2]] - X[is[3]])));
        }
        return volume;
    }

    uint32_t VertCount() const override { return impl.VertCount(); }
    uint32_t ElementCount() const override { return impl.ElementCount(); }

    // The sweep order of the GPU scene: set as a GeoLinear3d's tOrder, the
    // reference reproduces this body bit for bit.
    int Order(uint32_t* tOrder) { return impl.Order(tOrder); }
    // FP64 positions and velocities (the reference's X, V).
    int GetState(double* x, double* v) { return impl.GetState(x, v); }
    int Status() const { return status; }
    const char* LastError() const { return impl.LastError(); }

    xfgpu::GeoGpuLinear3d impl;
    Connectivity3d con;
    std::vector<dvec3> X0;           // rest positions (GeoLinear3d::X0)
    std::vector<dvec3> X;            // host copy of the positions
    std::vector<uint32_t> right;     // the Right nodes (Geo.cpp:725)
    std::vector<uint8_t> everLocked;  // nodes whose inverse mass the locks have set to 0
    vec3 origin = vec3(0.0f);        // Geo3d::origin
    int status = 0;
    bool fresh = false;              // X holds the current positions

   private:
    int FetchPositions() {
        if (fresh) return 0;
        const uint32_t n = impl.VertCount();
        std::vector<double> flat(3 * (size_t)n);
        int rc = impl.GetState(flat.data(), nullptr);
        if (rc) return rc;
        X.resize(n);
        for (uint32_t i = 0; i < n; i++) X[i] = dvec3(flat[3 * i], flat[3 * i + 1], flat[3 * i + 2]);
        fresh = true;
        return 0;
    }
};
