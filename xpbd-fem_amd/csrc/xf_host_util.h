// xf_host_util.h -- host-side mesh, element-constant and colouring code
// (no HIP dependency).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace xf {

constexpr uint32_t kMaxColours = 256;

// Linear-tet block (GenerateHexBlock + GenerateTetBlock, MeshGen.cpp:156-244).
void TetBlock(uint32_t W, uint32_t H, uint32_t D, const float scale[3], const float origin[3], bool mirrored,
              float* xyz, uint32_t* tets);

// The reference's MINSTD sweep order (Geo.cpp:759-771).
void MinstdOrder(uint32_t n, uint32_t* order);

// SplitMix64 + Box-Muller normal deviates.
void NormalFill(uint64_t seed, float sigma, uint32_t n, float* out);

struct ElementConstants {
    std::vector<float> qi;   // 9 per tet, column-major inverse rest map
    std::vector<float> qq;   // 3 per tet
    std::vector<float> qr;   // 3 per tet
    std::vector<float> vol;   // rest volume per tet
    std::vector<float> area;  // rest surface area per tet (PBD damping)
    std::vector<float> w;     // inverse lumped mass per node
};

// InitFiniteElement for T4 over a whole mesh (Fem.cpp:197-224, 269;
// Geo.cpp:732-744). Returns false if a tet index is out of range.
bool InitElements(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                  ElementConstants* out);
// The same from FP64 positions (GeoLinear3d::Init after autoResize holds
// positions that float cannot represent; Geo.cpp:712-729).
bool InitElements(const double* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, float density,
                  ElementConstants* out);

// Greedy node-disjoint colouring in visit order. Returns the colour count
// (0 if more than kMaxColours would be needed).
uint32_t GreedyColour(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, const uint32_t* visit,
                      std::vector<uint32_t>* colour);

// Colour-major sweep order of a greedy colouring: tets of colour 0 in visit
// order, then colour 1, ... Returns the colour count (0 on failure).
uint32_t ColourSweepOrder(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, bool minstd_visit,
                          std::vector<uint32_t>* colour, std::vector<uint32_t>* order);

// Sweep blocks: a partition of the tets into blocks whose launch colours
// (0..3) make blocks of one launch colour node-disjoint, and a node-disjoint
// colouring (phases) inside each block. Node levels along a direction d are
// breadth-first distances over tet edges from the nodes of least d . p
// (within 1e-4 of its range); a component the search did not reach restarts
// from its node of least d . p. The nodes of one tet differ by at most one
// level, so a tet with minimum levels (k1, k2) in two directions touches
// levels k1..k1+1 and k2..k2+1; its block is (k1, k2) and its launch colour
// (k1 & 1) | (k2 & 1) << 1, and two blocks of one launch colour differ by at
// least 2 in k1 or in k2 and share no node. The direction pair (x with y,
// y - z, y + z, z; y with z, x - z, x + z; z with x - y, x + y) is the first
// with the fewest tets of one block on one node (the least phases a block
// can have); a pair reaching a quarter of the largest node valence ends the
// search. Blocks are numbered by (launch colour, k1, k2). Phases: DSatur
// inside each block (ties: more neighbours, then earlier in visit order),
// then tets of phases above ceil(tets / phases) move, in visit order, to the
// first smaller phase none of their nodes touches. Returns false (with a
// message) if a block needs more than 64 phases.
constexpr uint32_t kMaxBlockPhases = 64;
struct SweepBlocks {
    uint32_t n_blocks = 0;
    int axes[2] = {0, 1};          // direction pair (indices into x, y, z, y-z, y+z, x-z, x+z, x-y, x+y)
    std::vector<uint32_t> block;   // per tet
    std::vector<uint32_t> phase;   // per tet: colour inside its block
    std::vector<uint8_t> launch;   // per block: launch colour 0..3
    std::vector<uint32_t> phases;  // per block: phase count
};
bool BuildSweepBlocks(const float* xyz, uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets,
                      const uint32_t* visit, SweepBlocks* out, std::string* err);

// Vertex rings of a linear-tet mesh as GeoLinear3d::Init builds them
// (OneRing, Connectivity.h:106-155; Geo.cpp:745-757), in CSR form: ring v
// has nodes nodes[node_off[v] .. node_off[v+1]) (v first, then every node of
// its tets in first appearance; the tets of v in increasing index) and tets
// tets[tet_off[v] .. tet_off[v+1]); idx holds four ring-local node indices
// per ring tet. Rings exist for vertices 0 .. max index in the tets.
constexpr uint32_t kMaxRingNodes = 256;  // Geo.cpp:10
constexpr uint32_t kMaxRingTets = 128;   // OneRingConstructor::polys (Connectivity.h:151)
struct Rings {
    uint32_t n = 0, largest = 0;
    std::vector<uint32_t> node_off, nodes, tet_off, tets;
    std::vector<uint8_t> idx;
};
// Returns false if a ring would exceed kMaxRingNodes nodes or kMaxRingTets tets.
bool BuildRings(const uint32_t* tets, uint32_t n_tets, Rings* out);

// The reference's vOrder (Geo.cpp:759-771): the MINSTD state continues
// from the n_tets draws of the tet shuffle, then shuffles n_rings vertices.
void MinstdRingOrder(uint32_t n_tets, uint32_t n_rings, uint32_t* order);

// DSatur colouring of rings (rings sharing a node get different colours;
// ties: more conflicting rings, then earlier in visit order) and the
// colour-major ring order, stable in visit order. Returns the colour count
// (0 if more than kMaxColours would be needed).
uint32_t RingColourSweepOrder(const Rings& r, uint32_t n_nodes, const uint32_t* visit,
                              std::vector<uint32_t>* colour, std::vector<uint32_t>* order);

// LDLTDecomposeAndSolve (Ldlt.h:88-95). Returns 0 on success.
int LdltSolve(uint32_t n, const float* A, const float* b, float* x);

// Tet mesh from a tetgen .node / .ele pair: xyz (3 per node) and tets (4 per
// tet, 0-based; tetgen numbering may start at 0 or 1). '#' starts a comment.
// Only 4-node tets are accepted. If orient, tets of negative volume get their
// first two nodes swapped (positive rest volume, as generated blocks have).
// Returns false with a message in *err.
bool ReadTetgen(const char* node_path, const char* ele_path, bool orient, std::vector<float>* xyz,
                std::vector<uint32_t>* tets, std::string* err);

// The element index stream of GeoLinear3d::Init (Geo.cpp:697) -- [4, i0, i1,
// i2, i3]* for tets -- to 4 indices per tet. Returns false on any element
// that is not a tet.
bool TetsFromStream(const uint32_t* idx, uint32_t count, std::vector<uint32_t>* tets);

}  // namespace xf
