/*
 * xpbdfem.h -- C ABI of the MI355X (gfx950) small-step XPBD substep for
 * linear tetrahedra (Neo-Hookean / Yeoh, incompressible mixed form).
 *
 * The reference (jak-xyz/xpbd-fem) has no C ABI for this path: its step
 * API is the C++ virtual interface Geo::Substep / Constrain / Damp
 * (Geo.h:15-35) over GeoLinear3d::Init (Geo.h:147, Geo.cpp:697-772), with
 * the per-element entry SolveElement (Fem.h:124, Fem.cpp:892). The entry
 * points below replace those; the C++ adapter in
 * xpbd-fem_amd/csrc/xf_geo_gpu.h keeps the Geo3d-shaped interface on top.
 *
 * Conventions: plain pointers and sizes, host buffers are caller-owned and
 * copied, device memory is library-owned. Every call returns an xf_status;
 * no call aborts or throws. xf_last_error() gives the message of the last
 * failing call on a context. A context owns one HIP stream and is not
 * thread-safe; separate contexts may be used from separate threads.
 */
#ifndef XPBDFEM_H
#define XPBDFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XF_API __attribute__((visibility("default")))

typedef enum xf_status {
    XF_OK = 0,
    XF_E_INVALID = -1,     /* bad argument or mesh */
    XF_E_HIP = -2,         /* HIP runtime error */
    XF_E_NOMEM = -3,       /* device or host allocation failed */
    XF_E_STATE = -4,       /* call not valid in the object's current state */
    XF_E_UNSUPPORTED = -5, /* option outside the implemented scope */
    XF_E_NONFINITE = -6    /* NaN/Inf detected in the node state */
} xf_status;

/* Energy ids of the 5-bit energy field (Settings.h:41-53). */
enum {
    /* Every 5-bit value is accepted and dispatched as
     * TSolveElement does for linear tets (Fem.cpp:870-889): Pixar and
     * PixarReduced coincide on a linear tet (one quadrature point), the Cube
     * energies constrain quads and hexes only (tets stay free), values above
     * 12 take the default, MixedSelective. Pixar and Cube run with per-colour
     * launches; Pixar takes settings.compliance (no per-tet mu). */
    XF_ENERGY_PIXAR = 0,
    XF_ENERGY_PIXAR_REDUCED = 1,
    XF_ENERGY_PIXAR_SEL = 2,
    XF_ENERGY_MIXED = 3,
    XF_ENERGY_MIXED_SEL = 4,
    XF_ENERGY_YEOH_SKIN = 5,
    XF_ENERGY_YEOH_SKIN_SEL = 6,
    XF_ENERGY_YEOH_SKIN_FAST = 7,
    /* Vertex-ring ("continuous") energies: the sweep runs over vertex rings
     * instead of tets (GeoLinear3d::Constrain, Geo.cpp:783-786; SolveRing and
     * SolveSimplex, Geo.cpp:13-64, 148-226). They take settings.compliance
     * and poissons_ratio (no per-tet mu), ignore volume_passes and
     * simultaneous as the reference does, and run with per-colour launches.
     * All bodies of a scene use ring energies or none does. */
    XF_ENERGY_CONTINUOUS_PIXAR = 8,
    XF_ENERGY_CONTINUOUS_MIXED = 9,
    XF_ENERGY_CONTINUOUS_SKIN = 10,
    XF_ENERGY_CUBE_NEO = 11,
    XF_ENERGY_CUBE_SKIN = 12
};

/* Collision flags (no reference counterpart: the reference has no
 * collision code). */
enum { XF_COLLIDE_GROUND = 1, XF_COLLIDE_SPHERE = 2 };
/* Rayleigh damping types (Settings.h:69-72). */
enum { XF_RAYLEIGH_PAPER = 0, XF_RAYLEIGH_LIMIT = 1, XF_RAYLEIGH_POST = 2, XF_RAYLEIGH_POST_AMORTIZED = 3 };

/* Colouring visit orders for the greedy node-disjoint tet colouring. */
enum {
    XF_COLOUR_NATURAL = 0, /* tets in input order */
    XF_COLOUR_MINSTD = 1   /* the reference's MINSTD sweep order (Geo.cpp:759-771) */
};

typedef struct xf_ctx xf_ctx;
typedef struct xf_scene xf_scene;

/* Per-body runtime settings: the hot-path fields of struct Settings
 * (Settings.h:79-102) plus Yeoh coefficients and collisions.
 *   energy, simultaneous  <- flags bits 6-10 and bit 11
 *   compliance            <- Settings::compliance (mu = 1/compliance,
 *                            Fem.cpp:445); ignored for tets with a per-tet mu
 *   poissons_ratio        <- Settings::poissonsRatio (0.5 = incompressible)
 *   gravity[2]            <- Settings::gravity (z is always 0)
 *   drag                  <- Settings::timeCorrectedDrag (velocity scale
 *                            1 - drag per substep, Geo.cpp:309)
 *   yeoh[3]               Yeoh C used by energies 5/6/7; the reference
 *                         hard-codes {0.1095, 14.95, 4.595} (Fem.cpp:471,526).
 *                         A Yeoh material C10, C20, C30 maps to
 *                         compliance = k/C10, yeoh = k*{1, C20/C10, C30/C10}
 *                         for any k > 0; k = 0.1095 (the reference's skin
 *                         scale) keeps the simultaneous solve stable where
 *                         k = 1 does not (same energy, different XPBD
 *                         conditioning).
 *   volume_passes         <- Settings::volumePasses (at most 64): extra
 *                            SolveElementVolume sweeps per substep with
 *                            compliance * tet volume (Fem.cpp:840-867),
 *                            using this compliance even for per-tet mu
 *   damping               <- Settings::damping (Rayleigh gamma)
 *   rayleigh_type         <- flag bits 20-21: XF_RAYLEIGH_PAPER and _LIMIT
 *                            damp inside the projection (Xpbd.h:87-214);
 *                            _POST and _POST_AMORTIZED run RayleighDamp on the
 *                            velocities after the velocity update
 *                            (GeoLinear3d::Damp, Geo.cpp:790-811), the
 *                            amortized form on 1/8 of the sweep order per
 *                            substep (tick % 8) with damping x 8
 *   pbd_damping           <- Settings::volumeAndTimeCorrectedPbdDamping, the
 *                            per-volume factor of the PBD velocity damping
 *                            (Xpbd.h:306-348, min(1, pbd_damping / area));
 *                            the pass runs when it is > 0
 *   pbd_damping_amortized <- amortizedVolumeAndTimeCorrectedPbdDamping, used
 *                            instead under XF_RAYLEIGH_POST_AMORTIZED
 *   unlock_pins           <- Settings_LockLeft clear (Geo.cpp:318): 1 leaves
 *                            the body's pinned nodes free in these substeps
 *                            (a node that was locked before keeps inverse
 *                            mass 0, as in the reference) */
typedef struct xf_settings {
    uint32_t energy;
    uint32_t simultaneous;
    float compliance;
    float poissons_ratio;
    float gravity[2];
    float drag;
    float damping;
    float pbd_damping;
    uint32_t volume_passes;
    float yeoh[3];
    uint32_t collide;      /* XF_COLLIDE_* bits */
    float ground_y;        /* plane y = ground_y, normal +y */
    float friction;        /* Coulomb coefficient (static = kinetic) */
    float sphere_c0[3];    /* sphere centre at simulation time 0 */
    float sphere_v[3];     /* sphere velocity (kinematic) */
    float sphere_r;
    uint32_t rayleigh_type;       /* XF_RAYLEIGH_* (Settings flag bits 20-21) */
    float pbd_damping_amortized;  /* amortizedVolumeAndTimeCorrectedPbdDamping */
    uint32_t unlock_pins;         /* 1: pins not locked (Settings_LockLeft clear) */
} xf_settings;

/* Fills defaults: MixedSel, simultaneous, compliance 0.01, nu 0.45,
 * gravity (0, -9.81), skin Yeoh constants, no collisions (the
 * Settings.h:79-92 defaults with the solve bit set). */
XF_API void xf_settings_default(xf_settings* s);

/* One body: a linear-tet mesh. Mirrors GeoLinear3d::Init (Geo.cpp:697):
 * xyz[3*n_nodes], tets[4*n_tets] (the index stream [4,i0..i3]* of the
 * reference without the leading 4s), density -> lumped masses rho*V/4
 * (Fem.cpp:217-223), w = 1/m (Geo.cpp:744).
 * pinned[n_nodes] (nullable): explicit pins with the Settings_LockLeft
 * semantics (X reset to O after the sweep, w = 0 from the first substep).
 * velocity[3*n_nodes] (nullable): initial velocity, else 0.
 * tet_mu[n_tets] (nullable): per-tet shear modulus mu (compliance 1/mu);
 * overrides settings.compliance for this body. */
typedef struct xf_body_desc {
    const float* xyz;
    uint32_t n_nodes;
    const uint32_t* tets;
    uint32_t n_tets;
    float density;
    const uint8_t* pinned;
    const float* velocity;
    const float* tet_mu;
    const uint32_t* tet_colour; /* nullable: caller's colouring (< 256 colours, node-disjoint
                                   classes, validated); else the scene's greedy colouring */
    const double* xyz64;        /* nullable: FP64 positions (3 per node). When set they replace
                                   xyz (which may then be NULL) for the initial node state and
                                   the element constants -- the positions GeoLinear3d::Init
                                   holds after autoResize (Geo.cpp:712-729), which float
                                   cannot represent. */
    uint32_t pin_first_substep; /* 1: pinned nodes keep inverse mass 1/m until the end of their
                                   first locked substep, as the reference's Left nodes do
                                   (w = 1/m from GeoLinear3d::Init, Geo.cpp:744, set to 0 by the
                                   lock, Geo.cpp:318-322), so the first sweep moves them before
                                   they are reset to O; 0: inverse mass 0 from the start */
} xf_body_desc;

/* Launch modes of xf_scene_step. */
enum {
    XF_LAUNCH_PER_COLOUR = 0, /* one launch per colour phase and node phase (hipGraph-captured if use_graph) */
    XF_LAUNCH_PERSISTENT = 1, /* one cooperative launch per step call, grid barriers between phases */
    XF_LAUNCH_BLOCKS = 2      /* needs sweep_blocks: one projection launch per block launch colour
                                 (at most 4 per substep), each workgroup of block_size threads
                                 sweeping one block through its phases with the block's nodes in
                                 LDS; other passes run per colour */
};

typedef struct xf_scene_opts {
    uint32_t colour_order;      /* XF_COLOUR_* */
    uint32_t block_size;        /* workgroup size of per-colour and sweep-block projection
                                   launches (64, 128, 192, 256), 0 = default (64) */
    uint32_t use_graph;         /* capture each per-colour xf_scene_step call shape in a hipGraph */
    uint32_t launch_mode;       /* XF_LAUNCH_* */
    uint32_t persistent_blocks; /* grid size of the persistent launch, 0 = one per compute unit */
    uint32_t xcd_slabs;         /* 1: inside each colour, cut each body's tets (rings) into eighths
                                   by centroid (vertex) x, node-index ordered inside an eighth, and
                                   let workgroup b work on eighth (b mod 8) of the colour, so each
                                   of the 8 XCDs (workgroups are dispatched round-robin over them)
                                   sweeps one spatial slab and node lines are fetched by one L2.
                                   Changes the sweep order inside colours (a valid serial order
                                   still; xf_scene_get_order / xf_scene_get_ring_order report it). */
    uint32_t ring_lanes;        /* lanes per vertex ring in ring-energy launches: 16, 32, or 0 = 32
                                   when some ring has more than 16 tets, else 16 */
    uint32_t sweep_blocks;      /* 1: colour the tets by sweep blocks. Node levels are breadth-first
                                   distances over tet edges from the nodes of least d . p, for a
                                   pair of directions d (x with y, y - z, y + z, z; y with z,
                                   x -+ z; z with x -+ y: the pair whose blocks put the fewest tets
                                   on one node); a tet's block is its pair of minimum levels and its
                                   launch colour the pair's parities, so blocks of one launch
                                   colour share no node. Inside a block, tets take DSatur phases
                                   (ties in visit order). Uniform blocks of Freudenthal cells get
                                   6 phases per block, 24 per substep.
                                   Colour (q, j) holds phase j of every block of launch colour q;
                                   the order is colour-major, then body, block, visit order (a
                                   valid serial order; xf_scene_get_order reports it). Needed by
                                   XF_LAUNCH_BLOCKS; excludes xcd_slabs and tet_colour. The nodes
                                   of a block must fit the LDS of one workgroup (at most 2048). */
    uint32_t block_lanes;       /* sweep blocks: consecutive blocks of a launch colour share one
                                   workgroup while each phase holds at most this many tets (1 to
                                   256; a block above it gets a workgroup of its own), or 0: the
                                   largest of 64, 128, 256 that leaves at least 4 workgroups per
                                   compute unit in every launch colour (else 64) */
    uint32_t separate_finish;   /* sweep blocks: 1 keeps the end of each substep (pinning, contact,
                                   velocity, next prediction) as a node launch of its own; 0: the
                                   last block launch of a substep does it for every node when at
                                   least two launch colours hold blocks, the last has at most two
                                   workgroups per compute unit (latency-bound) whose nodes fit LDS
                                   with their O (68 B a node), no body has volume passes or
                                   damping and the energy constrains tets (not in the last
                                   substep of a step call, which stores the velocities) */
    uint32_t tet_lanes;         /* sweep-block launches: lanes per tet, 1 (0 = 1) or 4 (four lanes of a
                                   quad share one tet, each owning one node: the per-node work is
                                   split four ways, for launches whose phases are latency-bound; the
                                   simultaneous MixedSel / YeohSkinFast projection without damping,
                                   else step calls fail with XF_E_UNSUPPORTED). Same results. */
    uint32_t contract;          /* sweep-block launches: 1 runs kernels compiled with multiply-add
                                   contraction (fused multiply-adds in the element chain and the
                                   node phase: about a quarter fewer instructions per tet). Results
                                   are then no longer bit-identical to the reference's serial FP32
                                   sweep (each product-sum rounds once); positions stay within a
                                   relative 1e-4 of it after a frame. 0 (default): the reference's
                                   rounding. Other launch modes are unaffected; the adjoint entry
                                   points reject a scene with it set in XF_LAUNCH_BLOCKS. */
} xf_scene_opts;

typedef struct xf_scene_info {
    uint32_t n_bodies;
    uint32_t n_nodes;
    uint32_t n_tets;
    uint32_t n_colours;      /* colour phases per substep (max over bodies) */
    uint64_t device_bytes;   /* device memory held by the scene */
    uint32_t n_rings;        /* vertex rings (0 until a ring energy or xf_scene_get_ring_order builds them) */
    uint32_t n_ring_colours; /* ring colour phases per substep (max over bodies) */
    uint32_t n_blocks;       /* sweep blocks over all bodies (0 without sweep_blocks) */
    uint32_t block_launches; /* block launch colours that hold blocks (0 without sweep_blocks) */
    uint32_t n_block_nodes;  /* nodes of all sweep-block workgroups (slots), each counting its own (a
                                node shared by slots of different launch colours counts once per slot) */
    uint32_t block_lanes;    /* the slot lane cap in use (xf_scene_opts.block_lanes or its automatic
                                choice; 0 without sweep_blocks) */
    uint32_t n_block_slots;  /* workgroups of the block launches of one substep */
    uint32_t fused_finish;   /* 1: under the current settings, step calls in XF_LAUNCH_BLOCKS end
                                their substeps inside the last block launch (see separate_finish;
                                the energy is not checked here) */
    uint32_t tet_lanes;      /* lanes per tet of sweep-block launches in use (xf_scene_opts.tet_lanes,
                                or the autotuner's choice) */
    uint32_t block_lds_limit; /* LDS bytes one sweep-block workgroup may use: the device's
                                 per-workgroup limit (160 KiB on gfx950); it caps a slot at
                                 block_lds_limit / 32 nodes and decides whether the node phase
                                 fits the last launch */
    uint32_t contract;        /* 1: step calls run the contracted sweep-block kernels
                                 (xf_scene_opts.contract in XF_LAUNCH_BLOCKS) */
    uint32_t block_node_stride; /* LDS bytes per slot node in the last block launch run: 48 where
                                   every block launch has at most two workgroups per compute unit,
                                   the projection is simultaneous MixedSel / YeohSkinFast without
                                   damping, one lane per tet, and the nodes fit (fewer LDS bank
                                   conflicts); else 32 (also before the first step) */
} xf_scene_info;

/* ------------------------------------------------------------ context */
XF_API int xf_ctx_create(int device, xf_ctx** out);
/* Destroys the context and every scene still open on it; handles of those
 * scenes become invalid and must not be passed to xf_scene_destroy again. */
XF_API void xf_ctx_destroy(xf_ctx* ctx);
XF_API const char* xf_last_error(const xf_ctx* ctx);
XF_API const char* xf_version(void);

/* -------------------------------------------------------------- scene */
XF_API int xf_scene_create(xf_ctx* ctx, xf_scene** out);
XF_API void xf_scene_destroy(xf_scene* scene);
/* Adds a body (before xf_scene_finalize). *body_id receives its index. */
XF_API int xf_scene_add_body(xf_scene* scene, const xf_body_desc* desc, uint32_t* body_id);
/* Element constants, colouring, device layout and upload. opts nullable. */
XF_API int xf_scene_finalize(xf_scene* scene, const xf_scene_opts* opts);
XF_API int xf_scene_info_get(const xf_scene* scene, xf_scene_info* info);
/* Runtime settings of one body (any time after add_body). */
XF_API int xf_scene_set_settings(xf_scene* scene, uint32_t body, const xf_settings* s);
/* Replaces the per-tet mu of a body (n_tets values). */
XF_API int xf_scene_set_tet_mu(xf_scene* scene, uint32_t body, const float* tet_mu);

/* Advances every body by `substeps` calls of the substep of Geo3d::Substep
 * (Geo.cpp:305-356): predict, tet sweep in colour order, pin, collisions,
 * velocity. dt is the substep length (Demo.cpp:39: 1/substepsPerSecond).
 * Returns when the work is enqueued; xf_scene_sync waits for it. */
XF_API int xf_scene_step(xf_scene* scene, uint32_t substeps, float dt);
XF_API int xf_scene_sync(xf_scene* scene);

/* Node state of one body: x[3*n], v[3*n] (either nullable). The state is
 * FP64 like the reference's dvec3 X, V (Geo.h:61-64). Synchronous. */
XF_API int xf_scene_get_state(xf_scene* scene, uint32_t body, double* x, double* v);
XF_API int xf_scene_set_state(xf_scene* scene, uint32_t body, const double* x, const double* v);
/* Simulation clock (seconds) that drives kinematic colliders. */
XF_API int xf_scene_get_time(const xf_scene* scene, float* t);
XF_API int xf_scene_set_time(xf_scene* scene, float t);
/* Settings::tickId of the next substep (the amortized damping picks its
 * 1/8 of the sweep order by tick % 8). Advances by one per substep; starts
 * at 0. Sim::Update keeps one counter for all bodies (Demo.cpp:81-89). */
XF_API int xf_scene_get_tick(const xf_scene* scene, uint32_t* tick);
XF_API int xf_scene_set_tick(xf_scene* scene, uint32_t tick);

/* The sweep order of one body: its tets (body-local ids) in colour order.
 * Sweeping this order serially reproduces the GPU result. */
XF_API int xf_scene_get_order(const xf_scene* scene, uint32_t body, uint32_t* order, uint32_t* colour_of_tet);

/* The vertex-ring sweep order of one body (used by the ring energies 8..10):
 * *n_rings = rings of the body (vertices 0 .. largest tet index); order
 * (nullable) gets them in sweep order, colour_of_ring (nullable) each
 * ring's colour. Sweeping this order serially reproduces the GPU result.
 * Replaces GeoLinear3d::vOrder (Geo.h:162, Geo.cpp:759-771), which the
 * reference shuffles with MINSTD. Rings are coloured by DSatur (rings that
 * share a node get different colours; ties: more conflicting rings, then
 * earlier in the visit order: vertex order, or with XF_COLOUR_MINSTD the
 * reference's vOrder); inside a colour they follow the visit order. Builds
 * the rings on first use. */
XF_API int xf_scene_get_ring_order(xf_scene* scene, uint32_t body, uint32_t* n_rings, uint32_t* order,
                                   uint32_t* colour_of_ring);
/* Element constants of one body as computed by InitFiniteElement
 * (Fem.cpp:269): qi9 column-major inverse rest map, qq3, qr3, vol per tet,
 * w per node (0 where pinned). Any pointer may be null. */
XF_API int xf_scene_get_elements(const xf_scene* scene, uint32_t body, float* qi9, float* qq3, float* qr3,
                                 float* vol, float* w);

/* Runs `substeps` substeps eagerly with a HIP event pair around every
 * launch and reports device time: ms_project = sum over projection
 * kernels, ms_node = sum over node kernels, launches_project = number of
 * projection launches. */
XF_API int xf_scene_profile(xf_scene* scene, uint32_t substeps, float dt, float* ms_project, float* ms_node,
                            uint32_t* launches_project);
/* Runs `calls` calls of xf_scene_step(substeps, dt) between two HIP events
 * recorded on the context's stream and returns the device time between them
 * in *ms (synchronous). The benchmark's timed region; no reference
 * counterpart (the reference prints wall-clock sim time, Demo.cpp:92-117). */
XF_API int xf_scene_time_steps(xf_scene* scene, uint32_t calls, uint32_t substeps, float dt, float* ms);
/* Autotuner: times `reps` calls of xf_scene_step(substeps, dt) for each
 * launch shape -- per-colour launches with workgroup sizes 64, 128, 256 and
 * the persistent launch with 64, 128, 256 and 512 blocks (where co-resident),
 * and, when the scene has sweep_blocks, sweep-block launches with
 * workgroup sizes 64, 128, 256, 1 or 4 lanes per tet (4 where the
 * projection allows it) and slot caps (block_lanes) 32, 64, 128 and 256,
 * the slots repacked for each cap -- and keeps the fastest in the scene
 * options (block_lanes is set when the chosen cap differs from the one in
 * use). The node state and clock are restored afterwards, so tuning does
 * not advance the simulation. chosen / best_ms (ms per call) nullable.
 * No reference counterpart (the reference is single-threaded). */
XF_API int xf_scene_autotune(xf_scene* scene, uint32_t substeps, float dt, uint32_t reps, xf_scene_opts* chosen,
                             float* best_ms);
/* Adjoint (no reference counterpart). Rolls every body forward `substeps`
 * substeps of length dt from the current state (as xf_scene_step would) and
 * returns the trajectory loss
 *   L = sum_j sum_b (1/n_b) sum_{i in b} |x_i(after substep loss_at[j]) - target_j,i|^2
 * and grad_log_mu[t] = dL/d(log mu_t) for every tet of every body (bodies
 * concatenated, each in its own tet order). loss_at: n_loss ascending
 * substep counts in [1, substeps]; targets: n_loss x 3 x (total nodes), body
 * order. The forward states are checkpointed every `window` substeps and each
 * window is replayed once (recording) before its reverse sweep; per-tet
 * gradients are summed in FP32 within a window and in FP64 across windows.
 * Scope: simultaneous MixedSel / YeohSkinFast, pins, ground contact (sphere
 * contact and serial mode return XF_E_UNSUPPORTED). The scene is left at the
 * end of the forward rollout. body_loss (nullable) receives each body's
 * term of L (one double per body), so a caller can tell which bodies of a
 * batch went non-finite. With the scene option use_graph, the forward runs
 * through xf_scene_step's graphs and the reverse pass (window replays and
 * reverse sweeps) is captured as one hipGraph, kept for the call shape
 * (substeps, window, loss substeps, buffers) and replayed by later calls of
 * that shape; without it every launch is eager. */
XF_API int xf_scene_adjoint(xf_scene* scene, uint32_t substeps, float dt, uint32_t n_loss, const uint32_t* loss_at,
                            const double* targets, uint32_t window, double* loss, double* grad_log_mu,
                            double* body_loss);
/* xf_scene_adjoint in the FP64 element model (no reference counterpart):
 * the rollout and its adjoint with every element, contact and settings
 * product in double -- element constants and settings widened from their
 * FP32 values, gravity * dt and 1 - drag and 1 / dt formed in double,
 * inverse masses in FP32 as stored -- and per-tet window sums in double.
 * The loss and gradient are those of this FP64 model (the gradient of the
 * FP32 element arithmetic carries its rounding noise; this one matches
 * finite differences of the FP64 model). Same arguments, scope and end
 * state (the FP64 rollout's) as xf_scene_adjoint; one colour launch per
 * colour and substep (no sweep blocks, no hipGraph). */
XF_API int xf_scene_adjoint_fp64(xf_scene* scene, uint32_t substeps, float dt, uint32_t n_loss,
                                 const uint32_t* loss_at, const double* targets, uint32_t window, double* loss,
                                 double* grad_log_mu, double* body_loss);
/* xf_scene_adjoint without allocating: no hipMalloc, hipFree or other
 * allocating call. No counterpart in the reference; for callers that cannot
 * allocate (tight fitting loops, harnesses that forbid device allocation
 * during a run). Same arguments, results and scope as xf_scene_adjoint
 * (loss, gradients and body_loss equal it bit for bit), with two
 * differences: the forward rollout runs as eager launches (no hipGraph is
 * instantiated), and the device buffers of the call -- checkpoints, the
 * window's sweep tape, adjoint state, loss staging -- come from a fixed
 * 32 MiB scratch array that each device holds for the life of the process:
 *   2 N 32 ceil(substeps / window) + 96 T min(window, substeps)
 *   + 64 N min(window, substeps) + 160 N + 20 T + 88 N n_loss + 8 N
 *   + 8 B (n_loss + 1) + 8 bytes (N nodes, T tets, B bodies; each of the 19
 *   regions rounded up to 256 bytes)
 * must fit, else XF_E_INVALID (the message gives the bytes needed); the
 * 16x4x4 beam (425 nodes, 1,536 tets) over 600 substeps with window 20 needs
 * about 4.3 MiB. XF_E_STATE when the scene's per-tet material storage would
 * have to be allocated first (a multi-body or compressible scene whose
 * settings changed since its last step): run xf_scene_step or
 * xf_scene_adjoint once. Every call on a device shares that device's
 * scratch and leaves its contents behind (what must start at zero is
 * cleared by each call); calls that overlap on one device, from different
 * contexts or threads, race. */
XF_API int xf_scene_adjoint_noalloc(xf_scene* scene, uint32_t substeps, float dt, uint32_t n_loss,
                                    const uint32_t* loss_at, const double* targets, uint32_t window, double* loss,
                                    double* grad_log_mu, double* body_loss);
/* Nodes held at kinematic targets: the n body-local nodes get X = O =
 * xyz[3 k ..] (FP64) and inverse mass 0 after the sweep of every substep of
 * later step calls, before the velocity update -- what Settings_LockRight
 * does to the reference's Right nodes with their animated target
 * origin + lockedRightTransform3d X0 (Geo.cpp:324-331). n = 0 clears. A
 * caller that animates the targets steps one substep at a time and sets them
 * before each. These nodes skip contact; the adjoint and the persistent
 * launch reject scenes with kinematic nodes. */
XF_API int xf_scene_set_kinematic(xf_scene* scene, uint32_t body, uint32_t n, const uint32_t* nodes,
                                  const double* xyz);
/* The manipulation constraint of Geo3d::Substep (Geo.cpp:334-339) on one
 * body-local node: after the sweep of every substep of later step calls,
 * X += (target - X) w / (max(1e-6, w) + 1.8 / dt^2) in FP32 as there.
 * node = 0xFFFFFFFF switches it off. The reference recomputes the target per
 * substep from the manipulator (Demo.cpp:84-85); a caller following it steps
 * one substep at a time. */
XF_API int xf_scene_set_manip(xf_scene* scene, uint32_t body, uint32_t node, const float target[3]);
/* Counts non-finite node positions (0 = all finite). */
XF_API int xf_scene_count_nonfinite(xf_scene* scene, uint32_t* count);

/* ---------------------------------------------------- host utilities */
/* Linear-tet block generator (GenerateHexBlock + GenerateTetBlock,
 * MeshGen.cpp:156-244, wonkiness 0): node (x,y,z) at origin + scale*(x,y,z),
 * index x + y(W+1) + z(W+1)(H+1); 6 tets per cell. xyz[3*(W+1)(H+1)(D+1)],
 * tets[24*W*H*D]. mirrored selects Pattern_Mirrored. */
XF_API int xf_mesh_tet_block(uint32_t W, uint32_t H, uint32_t D, const float scale[3], const float origin[3],
                             uint32_t mirrored, float* xyz, uint32_t* tets);
/* The 24-colouring of a uniform-pattern xf_mesh_tet_block mesh (the node
 * valence lower bound): colour = 6 * ((x + y + z) mod 4) + split index, for
 * the tet of split index s in cell (x, y, z). Every split tet spans its cell
 * diagonal in unit steps, so two tets of one split index share a node only
 * in cells whose offset has coordinate sum 1..3, never 0 mod 4. colour[6WHD].
 * No reference counterpart. */
XF_API int xf_mesh_tet_block_colours(uint32_t W, uint32_t H, uint32_t D, uint32_t* colour);
/* The reference's MINSTD sweep order for n tets (Geo.cpp:759-771). */
XF_API int xf_minstd_order(uint32_t n, uint32_t* order);
/* n normal deviates N(0, sigma^2) from SplitMix64(seed) + Box-Muller. */
XF_API int xf_normal_fill(uint64_t seed, float sigma, uint32_t n, float* out);
/* The greedy node-disjoint colouring xf_scene_finalize applies, on the host:
 * order[n_tets] = tets in sweep order (colour-major, visit order inside a
 * colour), colour[n_tets] (nullable), *n_colours. No reference counterpart:
 * the reference sweeps its MINSTD order serially (Geo.cpp:759-777). */
XF_API int xf_mesh_colour(uint32_t n_nodes, const uint32_t* tets, uint32_t n_tets, uint32_t colour_order,
                          uint32_t* order, uint32_t* colour, uint32_t* n_colours);
/* The sweep blocks xf_scene_finalize builds with sweep_blocks, on the host
 * (no reference counterpart): per tet its block (block_of_tet[n_tets]) and its
 * phase inside the block (phase_of_tet[n_tets]); per block its launch colour
 * (launch_of_block, capacity n_tets); *n_blocks; *max_phases (nullable) the
 * most phases of any block. Visit order per colour_order (XF_COLOUR_*). */
XF_API int xf_mesh_sweep_blocks(uint32_t n_nodes, const float* xyz, uint32_t n_tets, const uint32_t* tets,
                                uint32_t colour_order, uint32_t* n_blocks, uint32_t* block_of_tet,
                                uint32_t* phase_of_tet, uint32_t* launch_of_block, uint32_t* max_phases);
/* The vertex rings the ring energies sweep, on the host: the reference's
 * OneRing of every vertex 0 .. largest index (Connectivity.h:106-155, built by
 * GeoLinear3d::Init, Geo.cpp:745-757) in CSR form. *n_rings, *node_total and
 * *largest (nullable; nodes of the largest ring) always; where set,
 * node_off[n_rings + 1], nodes[node_total] (each ring's own vertex first),
 * tet_off[n_rings + 1], ring_tets[4 n_tets] and idx[16 n_tets] (four
 * ring-local node indices per ring tet). XF_E_UNSUPPORTED if a ring would
 * exceed 256 nodes (kMaxRingNodes, Geo.cpp:10) or 128 tets (the capacity of
 * OneRingConstructor, Connectivity.h:151). */
XF_API int xf_mesh_rings(uint32_t n_tets, const uint32_t* tets, uint32_t* n_rings, uint32_t* node_total,
                         uint32_t* node_off, uint32_t* nodes, uint32_t* tet_off, uint32_t* ring_tets, uint8_t* idx,
                         uint32_t* largest);
/* Tet mesh ingestion (no reference counterpart: the reference compiles its
 * tetgen-made Armadillo volume mesh in, MeshGen.cpp:23-80). Reads a tetgen
 * .node / .ele pair: call with xyz = tets = NULL for *n_nodes / *n_tets,
 * then again with xyz[3 * n_nodes] and tets[4 * n_tets]. Numbering may
 * start at 0 or 1; only 4-node tets; '#' comments. orient != 0 swaps the
 * first two nodes of negatively oriented tets so every rest volume is
 * positive. Errors: XF_E_INVALID with the text in xf_mesh_error(). */
XF_API int xf_mesh_read_tetgen(const char* node_path, const char* ele_path, uint32_t orient, uint32_t* n_nodes,
                               uint32_t* n_tets, float* xyz, uint32_t* tets);
/* The element index stream GeoLinear3d::Init takes (Geo.cpp:697),
 * [4, i0, i1, i2, i3]*, to 4 indices per tet: tets NULL for *n_tets only.
 * XF_E_UNSUPPORTED for streams holding other elements. */
XF_API int xf_mesh_tets_from_stream(const uint32_t* idx, uint32_t count, uint32_t* n_tets, uint32_t* tets);
/* Message of the last failed xf_mesh_* call on this thread. */
XF_API const char* xf_mesh_error(void);
/* LDLTDecomposeAndSolve (Ldlt.h:88-95): packed lower-triangular A, n <= 25. */
XF_API int xf_ldlt_solve(uint32_t n, const float* A, const float* b, float* x);

#ifdef __cplusplus
}
#endif

#endif /* XPBDFEM_H */
