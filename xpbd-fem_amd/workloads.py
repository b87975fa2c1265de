"""Workload definitions of the substep benchmarks and parity cases.

Each builder returns a dict with the mesh (xyz float32[3n], tets uint32[4t]),
density, pins, initial velocity, per-tet mu (or None), and the settings as
keyword arguments of xpbdfem.default_settings(), plus dt, substeps per frame
and frame count. Meshes come from xpbdfem.tet_block (the GenerateTetBlock
layout of MeshGen.cpp:156-244) placed with x starting at 0.
"""
import numpy as np

import xpbdfem as xf

FRAME_DT = 1.0 / 60.0


def _pins_x0(xyz):
    return (xyz[0::3] == 0.0).astype(np.uint8)


def beam_neo_12x4x4(energy=xf.ENERGY_MIXED_SEL, simultaneous=True, collide=True):
    """12x4x4 cells of 0.1 m (1,152 tets, 325 nodes), Neo-Hookean mu = 1e5 Pa
    (compliance 1e-5), incompressible, density 1000, x = 0 face pinned,
    gravity -9.81 y, ground y = -0.5 with friction 0.5; 1/60 s frames of 20
    substeps, 120 frames, from rest."""
    xyz, tets = xf.tet_block(12, 4, 4, (0.1, 0.1, 0.1), (0.0, -0.2, -0.2))
    substeps = 20
    return dict(
        name="beam12x4x4_neo", xyz=xyz, tets=tets, density=1000.0, pinned=_pins_x0(xyz), velocity=None, tet_mu=None,
        settings=dict(energy=energy, simultaneous=1 if simultaneous else 0, compliance=1e-5, poissons_ratio=0.5,
                      gravity=(0.0, -9.81), collide=xf.COLLIDE_GROUND if collide else 0, ground_y=-0.5,
                      friction=0.5),
        dt=np.float32(1.0) / np.float32(60 * substeps), substeps=substeps, frames=120)


def beam_yeoh_96x24x24(seed=1, h=0.1, structured=True):
    return beam_yeoh(96, 24, 24, seed=seed, h=h, structured=structured)


def beam_yeoh(W, H, D, seed=1, h=0.1, structured=True):
    """96x24x24 cells of h = 0.1 m (331,776 tets, 60,625 nodes; the beam spans
    x in [0, 9.6], y and z in [-1.2, 1.2]), Yeoh energy
    C10 (IM + 0.1 IM^2 + 0.01 IM^3) with C10 = 5e4, C20 = 5e3, C30 = 5e2 Pa,
    incompressible, density 1000, initial velocity N(0, 0.05^2) per component
    from SplitMix64(seed) + Box-Muller, x = 0 face pinned, ground y = -2.2
    (1 m under the beam, friction 0.5) and a kinematic sphere r = 0.3 m
    starting at (0, -1.5, 0), touching the underside at the pinned end and
    moving +x at 1 m/s; 1/60 s frames of 30 substeps, 60 frames.

    The energy goes in the YeohSkinFast form with the reference's skin
    normalisation (Fem.cpp:526, linear coefficient 0.1095): C = 0.1095 (1,
    0.1, 0.01) and compliance = 0.1095 / C10. Passing C = (1, 0.1, 0.01) with
    compliance = 1 / C10 describes the same polynomial, but the reference's
    simultaneous solve then turns chaotic within a few frames (node speeds
    above 100 m/s; also from rest, also without contacts), and it does so
    for the skin-normalised form too at cell sizes of 0.02 m and below.

    Other W x H x D cell counts give the same material, pins and contacts
    (the ground 1 m under the beam, the sphere at its underside)."""
    xyz, tets = xf.tet_block(W, H, D, (h, h, h), (0.0, -0.5 * H * h, -0.5 * D * h))
    n = xyz.size // 3
    vel = xf.normal_fill(seed, 0.05, 3 * n)
    substeps = 30
    c10, c20, c30 = 5e4, 5e3, 5e2
    k = xf.SKIN_YEOH[0]
    y_low = -0.5 * H * h
    return dict(
        name=f"beam{W}x{H}x{D}_yeoh", xyz=xyz, tets=tets, density=1000.0, pinned=_pins_x0(xyz), velocity=vel,
        tet_mu=None, tet_colour=xf.tet_block_colours(W, H, D) if structured else None,
        settings=dict(energy=xf.ENERGY_YEOH_SKIN_FAST, simultaneous=1, compliance=k / c10, poissons_ratio=0.5,
                      gravity=(0.0, -9.81), yeoh=(k, k * c20 / c10, k * c30 / c10),
                      collide=xf.COLLIDE_GROUND | xf.COLLIDE_SPHERE, ground_y=y_low - 1.0, friction=0.5,
                      sphere_c0=(0.0, y_low - 0.3, 0.0), sphere_v=(1.0, 0.0, 0.0), sphere_r=0.3),
        dt=np.float32(1.0) / np.float32(60 * substeps), substeps=substeps, frames=60)


def beam_fit_16x4x4(mu_split=(2e5, 2e4), x_split=0.8):
    """16x4x4 cells of 0.1 m (1,536 tets, 425 nodes), Neo-Hookean with per-tet
    mu: mu_split[0] where the tet centroid has x < x_split, else mu_split[1].
    Pins, gravity and ground as the 12x4x4 beam; 30 frames of 20 substeps."""
    xyz, tets = xf.tet_block(16, 4, 4, (0.1, 0.1, 0.1), (0.0, -0.2, -0.2))
    cx = xyz[0::3][tets.reshape(-1, 4)].mean(axis=1)
    mu = np.where(cx < x_split, mu_split[0], mu_split[1]).astype(np.float32)
    substeps = 20
    return dict(
        name="beam16x4x4_fit", xyz=xyz, tets=tets, density=1000.0, pinned=_pins_x0(xyz), velocity=None, tet_mu=mu,
        settings=dict(energy=xf.ENERGY_MIXED_SEL, simultaneous=1, compliance=1e-5, poissons_ratio=0.5,
                      gravity=(0.0, -9.81), collide=xf.COLLIDE_GROUND, ground_y=-0.5, friction=0.5),
        dt=np.float32(1.0) / np.float32(60 * substeps), substeps=substeps, frames=30)


class SplitMix64:
    """Deterministic uniform stream for workload draws (portable across
    numpy versions)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) / 9007199254740992.0)

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) % (hi - lo + 1)

    def normal(self):
        u1 = ((self.next() >> 11) + 1.0) / 9007199254740992.0
        u2 = (self.next() >> 11) / 9007199254740992.0
        return float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def random_rotation(rng):
    q = np.array([rng.normal() for _ in range(4)])
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def boxes_batch(n_bodies=256, seed=2, cells=(3, 24), h=0.1, frames=60):
    """n_bodies independent boxes: per-axis cell counts uniform in
    [cells[0], cells[1]] (cells of h m, 6 tets each), MixedSel or
    YeohSkinFast 50/50, mu (or C10) log-uniform over 1e4-1e6 Pa,
    incompressible, density uniform over 500-2000, a uniform random rotation,
    dropped from 0.2-1.0 m onto a ground plane y = 0 (friction 0.5) of its
    own; gravity -9.81 y; 1/60 s frames of 20 substeps. Yeoh bodies use the
    reference's skin shape constants (Fem.cpp:526) scaled so that the linear
    coefficient equals C10 (compliance = 0.1095 / C10). Bodies are placed 5 m
    apart along z (they never interact)."""
    rng = SplitMix64(seed)
    bodies = []
    for b in range(n_bodies):
        W, H, D = (rng.integer(*cells) for _ in range(3))
        yeoh = rng.uniform() < 0.5
        mod = float(10.0 ** rng.uniform(4.0, 6.0))
        rho = rng.uniform(500.0, 2000.0)
        drop = rng.uniform(0.2, 1.0)
        R = random_rotation(rng)
        xyz, tets = xf.tet_block(W, H, D, (h, h, h), (-0.5 * W * h, -0.5 * H * h, -0.5 * D * h))
        p = xyz.reshape(-1, 3).astype(np.float64) @ R.T
        p[:, 1] += drop - p[:, 1].min()
        p[:, 2] += 5.0 * b
        settings = dict(energy=xf.ENERGY_YEOH_SKIN_FAST if yeoh else xf.ENERGY_MIXED_SEL, simultaneous=1,
                        compliance=(0.1095 / mod) if yeoh else 1.0 / mod, poissons_ratio=0.5, gravity=(0.0, -9.81),
                        yeoh=xf.SKIN_YEOH, collide=xf.COLLIDE_GROUND, ground_y=0.0, friction=0.5)
        bodies.append(dict(name=f"box{b}_{W}x{H}x{D}", xyz=p.astype(np.float32).reshape(-1), tets=tets, density=rho,
                           pinned=None, velocity=None, tet_mu=None, tet_colour=xf.tet_block_colours(W, H, D),
                           settings=settings))
    substeps = 20
    return dict(name=f"boxes{n_bodies}", bodies=bodies, dt=np.float32(1.0) / np.float32(60 * substeps),
                substeps=substeps, frames=frames)


def fit_batch(n_bodies=64, seed=4, cells=(4, 16), h=0.08, frames=30, substeps=40):
    """Batched fitting problem: n_bodies cantilever beams, per-axis cell
    counts uniform in [cells[0], cells[1]] (cells of h m, 6 tets each),
    density 1000, Yeoh energy C10 (IM + r IM^2 + (r^2 / 10) IM^3) with the
    C20/C10 ratio r uniform over 0.01-0.2 (C30/C10 = r^2 / 10 is this
    workload's choice), and a per-tet C10 from a hidden log field: the sum of
    4 Gaussian bumps (centres uniform in the beam's box, widths 0.15-0.5 of
    its diagonal, weights 0.5-1.5) evaluated at the tet centroids and mapped
    affinely onto [log 1e4, log 1e6] Pa.

    The energy is passed in the YeohSkinFast form with the reference's skin
    normalisation (Fem.cpp:526, linear coefficient 0.1095): C = 0.1095 (1, r,
    r^2 / 10) and per-tet mu = 1 / compliance = C10 / 0.1095. Passing
    C = (1, r, r^2 / 10) with mu = C10 instead describes the same polynomial
    but drives the reference's simultaneous solve unstable from rest (node
    speeds of ~100 m/s after 40-100 substeps).

    The cell size h = 0.08 m keeps every body's rollout smooth for both the
    hidden field and the initial guess (node speeds below 4 m/s). At 0.05 m,
    5 of the 64 hidden fields drive the solve unstable in their stiff
    regions, in FP64 as well: the clamped deviatoric term's compliance,
    which scales as 1 / (mu h^3 dt^2), then no longer dominates the mass-
    weighted gradient term, which scales as 1 / (rho h^5).

    Incompressible, x = 0 face pinned, gravity -9.81 y, ground y = -1.0 with
    friction 0.5, from rest; 1/60 s frames of `substeps` substeps. Bodies sit
    2 m apart along z. Each body's tet_mu holds the hidden field and
    mu_guess the initial guess C10 = 1e5 Pa, both as the library's mu;
    mu_range is the hidden fields' range in the same units."""
    rng = SplitMix64(seed)
    bodies = []
    for b in range(n_bodies):
        W, H, D = (rng.integer(*cells) for _ in range(3))
        r = rng.uniform(0.01, 0.2)
        xyz, tets = xf.tet_block(W, H, D, (h, h, h), (0.0, -0.5 * H * h, -0.5 * D * h))
        p = xyz.reshape(-1, 3).astype(np.float64)
        lo, hi = p.min(axis=0), p.max(axis=0)
        diag = float(np.linalg.norm(hi - lo))
        cen = p[tets.reshape(-1, 4)].mean(axis=1)
        field = np.zeros(cen.shape[0])
        for _ in range(4):
            c = np.array([rng.uniform(lo[a], hi[a]) for a in range(3)])
            sig = rng.uniform(0.15, 0.5) * diag
            wgt = rng.uniform(0.5, 1.5)
            field += wgt * np.exp(-np.sum((cen - c) ** 2, axis=1) / (2.0 * sig * sig))
        f = (field - field.min()) / max(field.max() - field.min(), 1e-30)
        c10 = np.exp(np.log(1e4) + f * (np.log(1e6) - np.log(1e4)))
        p[:, 2] += 2.0 * b
        k = xf.SKIN_YEOH[0]
        settings = dict(energy=xf.ENERGY_YEOH_SKIN_FAST, simultaneous=1, compliance=k / 1e5, poissons_ratio=0.5,
                        gravity=(0.0, -9.81), yeoh=(k, k * r, k * r * r / 10.0), collide=xf.COLLIDE_GROUND,
                        ground_y=-1.0, friction=0.5)
        bodies.append(dict(name=f"beam{b}_{W}x{H}x{D}", xyz=p.astype(np.float32).reshape(-1), tets=tets,
                           density=1000.0, pinned=_pins_x0(xyz), velocity=None,
                           tet_mu=(c10 / k).astype(np.float32), mu_guess=np.full(c10.size, 1e5 / k, np.float32),
                           tet_colour=xf.tet_block_colours(W, H, D), settings=settings, yeoh_ratio=r))
    k = xf.SKIN_YEOH[0]
    return dict(name=f"fit{n_bodies}", bodies=bodies, dt=np.float32(1.0) / np.float32(60 * substeps),
                substeps=substeps, frames=frames, mu_range=(1e4 / k, 1e6 / k))


def make_batch_scene(ctx, batch, bodies=None, colour_order=xf.COLOUR_NATURAL, use_graph=True, block_size=64,
                     launch_mode=xf.LAUNCH_PER_COLOUR, persistent_blocks=0, xcd_slabs=False, ring_lanes=0,
                     sweep_blocks=False, block_lanes=0, separate_finish=False, tet_lanes=0, contract=False):
    """One scene holding the given bodies of a batch (all by default). With
    sweep_blocks the scene colours the tets itself (tet_colour is dropped)."""
    sc = xf.Scene(ctx)
    chosen = batch["bodies"] if bodies is None else [batch["bodies"][i] for i in bodies]
    for wl in chosen:
        sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], velocity=wl["velocity"],
                    tet_mu=wl["tet_mu"], tet_colour=None if sweep_blocks else wl.get("tet_colour"))
    sc.finalize(colour_order=colour_order, use_graph=use_graph, block_size=block_size, launch_mode=launch_mode,
                persistent_blocks=persistent_blocks, xcd_slabs=xcd_slabs, ring_lanes=ring_lanes,
                sweep_blocks=sweep_blocks, block_lanes=block_lanes, separate_finish=separate_finish,
                tet_lanes=tet_lanes, contract=contract)
    for i, wl in enumerate(chosen):
        sc.set_settings(i, **wl["settings"])
    return sc


def make_scene(ctx, wl, colour_order=xf.COLOUR_NATURAL, use_graph=True, block_size=64,
               launch_mode=xf.LAUNCH_PER_COLOUR, persistent_blocks=0, xcd_slabs=False, ring_lanes=0,
               sweep_blocks=False, block_lanes=0, separate_finish=False, tet_lanes=0, contract=False):
    """Builds and finalizes a one-body scene of a workload. With sweep_blocks
    the scene colours the tets itself (the workload's tet_colour is dropped)."""
    sc = xf.Scene(ctx)
    sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], velocity=wl["velocity"],
                tet_mu=wl["tet_mu"], tet_colour=None if sweep_blocks else wl.get("tet_colour"))
    sc.finalize(colour_order=colour_order, use_graph=use_graph, block_size=block_size, launch_mode=launch_mode,
                persistent_blocks=persistent_blocks, xcd_slabs=xcd_slabs, ring_lanes=ring_lanes,
                sweep_blocks=sweep_blocks, block_lanes=block_lanes, separate_finish=separate_finish,
                tet_lanes=tet_lanes, contract=contract)
    sc.set_settings(0, **wl["settings"])
    return sc
